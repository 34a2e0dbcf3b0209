"""uc_transport_props_noalloc, the allocation-free variant of uc_transport_props,
through the C ABI: its outputs equal the context entry point's bit for bit on two calls
in a row (the second with a smaller field), and again after a call with other scales
replaced the tables in its fixed device array."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from tests.golden_io import GOLDEN
from unicorn_amd import chem
from unicorn_amd.chem import ChemStep, DeviceField
from unicorn_amd.fields import ChemField

pytestmark = pytest.mark.gpu
KEYS = ("vmu", "vtc", "vdf", "stdf")


def field(y, tk, li):
    m = tk.size
    lj = m // li
    z = np.zeros(m)
    return ChemField(li, lj, np.ones(m), np.ascontiguousarray(y), np.ascontiguousarray(tk), z.copy(), z.copy(),
                     z.copy(), np.zeros(m, dtype=np.int32), 1e-3)


def noalloc(scales, dev):
    n = dev.rho.numel()
    e = lambda *s: torch.full(s, float("nan"), dtype=torch.float64, device="cuda")  # noqa: E731
    out = {"vmu": e(n), "vtc": e(n), "vdf": e(52, n), "stdf": e(n)}
    fc = dev.c()
    rc = chem.library().uc_transport_props_noalloc(C.byref(scales), C.byref(fc), *[out[k].data_ptr() for k in KEYS],
                                                   C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, chem.library().uc_last_error()
    return out


def with_context(scales, dev):
    return ChemStep(dev.rho.numel(), scales=scales).transport_props(dev)


def test_noalloc_transport_equals_the_context_entry_point():
    z = np.load(GOLDEN / "transport.npz", allow_pickle=False)
    y, tk = z["tr_y"], z["tr_tk"]
    big = DeviceField(field(y, tk, 50))                     # 3000 cells
    small = DeviceField(field(y[:, 1000:1960], tk[1000:1960], 30))   # 960 cells
    sc = chem.scales_from_card()
    other = chem.scales_from_card(vref=5.0, tref=300.0)
    for label, scales, dev in (("first", sc, big), ("second, smaller", sc, small),
                               ("other scales", other, small), ("back", sc, big)):
        print(f"no-alloc call: {label}, {dev.rho.numel()} cells", flush=True)
        a, b = noalloc(scales, dev), with_context(scales, dev)
        torch.cuda.synchronize()
        for k in KEYS:
            assert torch.equal(a[k], b[k]), (label, k)
    assert not torch.equal(noalloc(other, small)["vmu"], noalloc(sc, small)["vmu"])


def test_noalloc_transport_refuses_bad_sizes():
    """Refused on the host before any device work: the pointers are never read."""
    sc = chem.scales_from_card()
    fc = chem.UcField(2, 5, 0, 8, 8, 0, 0, 0, 0, 0, 0)
    rc = chem.library().uc_transport_props_noalloc(C.byref(sc), C.byref(fc), 8, 8, 8, 8, None)
    assert rc == -1 and b"at least 3" in chem.library().uc_last_error()
    fc = chem.UcField(4001, 1000, 8, 8, 8, 0, 0, 0, 0, 0, 0)
    rc = chem.library().uc_transport_props_noalloc(C.byref(sc), C.byref(fc), 8, 8, 8, 8, None)
    assert rc == -1 and b"4,000,000" in chem.library().uc_last_error()
