"""SOLVEH's T(h) failure conditions on the GPU (solveh.f:1170-1227): the thfail fixture
(tests/golden/make_golden.py) holds nine cells with enthalpies planted where the Newton
iteration finds no root in 100 steps (six) or the enthalpy lies above the high-range fit
at 5000 K (three), each outcome stable under 1e-13 relative changes of the enthalpy.

* uc_step: uc_stats counts both conditions and names the first cell without a root, as
  the reference's DEBUG(9427) lines and the oracle do, and uc_th_flags gives the oracle's
  per-cell flags; the field to the parity bar.
* uc_step_batch with a second grid: the counts and the first cell stay per grid.
* The Fortran drop-in (oracle/_ref/ref_driver_shim, libunicorn_fshim.so) prints the
  reference's FORMAT 714 lines, character for character, under its IOVER rule.
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import assert_state_close, load
from oracle import pyoracle, refrun

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def expected(z, lj):
    first = min((i - 1) * lj + (j - 1) for i, j in z["th_debug_ij"])
    return len(z["th_debug_ij"]), first


def test_gpu_th_failure_counts(torch_cuda):
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    f, z = load("thfail")
    print("thfail 31x21: one step on the GPU", flush=True)
    dev = DeviceField(f)
    ctx = ChemStep(f.n)
    st = ctx.step(dev, Controls(dt=f.dt))
    got = dev.to_host()
    o = pyoracle.step(f.copy())
    assert np.array_equal(ctx.th_flags(dev).cpu().numpy(), o.thflag)
    nfail, first = expected(z, f.lj)
    assert (st.isor1, st.isor2) == (int(z["s1_isor1"][0]), int(z["s1_isor2"][0]))
    assert (st.t_out_of_limit, st.newton_failed, st.first_failed) == (3, nfail, first)
    # the field to the parity bar; T and rho apart from the cells without a root, where
    # the reference's T is the 100th Newton iterate of an oscillation that any last-bit
    # difference in its inputs sends elsewhere
    e = {k: z[f"s1_{k}"] for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    act = f.active()
    nr = [(i - 1) * f.lj + (j - 1) for i, j in z["th_debug_ij"]]
    assert_state_close(got, e, act[~np.isin(act, nr)], "thfail, cells with a root")
    for k in ("tk", "rho"):
        getattr(got, k)[nr] = e[k][nr]
    assert_state_close(got, e, act, "thfail")


def test_gpu_th_failure_counts_per_grid(torch_cuda):
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import flame_field
    f, z = load("thfail")
    other = flame_field(41, 25, "CH4", 5)
    print("thfail + a 41x25 CH4 grid: one batched step", flush=True)
    devs = [DeviceField(other), DeviceField(f)]
    ctx = ChemStep(other.n + f.n, max_grids=2)
    st = ctx.step_batch(devs, [Controls(dt=other.dt), Controls(dt=f.dt)])
    nfail, first = expected(z, f.lj)
    assert (st[0].t_out_of_limit, st[0].newton_failed, st[0].first_failed) == (0, 0, -1)
    assert (st[1].t_out_of_limit, st[1].newton_failed, st[1].first_failed) == (3, nfail, first)
    assert not ctx.th_flags(devs[0], 0).any()
    assert np.array_equal(ctx.th_flags(devs[1], 1).cpu().numpy(), pyoracle.step(f.copy()).thflag)


@pytest.mark.skipif(not refrun.available("shim"), reason="drop-in driver not built (make -C oracle shim)")
def test_gpu_drop_in_prints_the_reference_messages(torch_cuda):
    f, z = load("thfail")
    print("thfail through the Fortran drop-in", flush=True)
    r = refrun.run(f.copy(), driver="shim")
    lines = [l for l in r.stdout.splitlines() if "UNABLE TO FIND THE ROOT" in l or "OUT-OF-LIMIT" in l]
    assert lines == [str(m) for m in z["th_msg"]]
    assert np.array_equal(r.isor1, z["s1_isor1"]) and np.array_equal(r.isor2, z["s1_isor2"])
