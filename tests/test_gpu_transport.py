"""Mixture-averaged transport properties on the GPU (uc_transport_props, main:829-1005)
against what PROGRAM UNICORN itself computed (tests/golden/transport.npz, dumped from
the reference program's /CB07/) and against the oracle (oracle/chem_oracle.c
uco_transport, pinned bit for bit to the program by tests/test_program.py) on states
the program's flames do not reach: sub-mixtures of up to all 52 species, pure species,
pure N2, and temperatures across both polynomial ranges and every collision-integral
branch.  The bar is bit for bit: the kernel keeps the reference's operation order with
contraction off, and FP64 division and square root are correctly rounded."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import pyoracle
from tests.golden_io import GOLDEN
from unicorn_amd.chem import ChemStep, DeviceField
from unicorn_amd.fields import ChemField

pytestmark = pytest.mark.gpu
KEYS = ("vmu", "vtc", "vdf", "stdf")


def as_field(y, tk, li):
    """m cells laid out as an li x (m/li) grid (every cell is evaluated)."""
    m = tk.size
    lj = m // li
    assert li * lj == m and li >= 3 and lj >= 3
    z = np.zeros(m)
    return ChemField(li, lj, np.ones(m), np.ascontiguousarray(y, dtype=np.float64), np.ascontiguousarray(tk),
                     z.copy(), z.copy(), z.copy(), np.zeros(m, dtype=np.int32), 1e-3)


def gpu(f):
    ctx = ChemStep(f.n)
    out = ctx.transport_props(DeviceField(f))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def oracle(y, tk):
    n = tk.size
    o = {"vmu": np.zeros(n), "vtc": np.zeros(n), "vdf": np.zeros((52, n)), "stdf": np.zeros(n)}
    L = pyoracle.lib()
    L.uco_transport.restype = C.c_int
    r = np.array(pyoracle.default_ref())
    y = np.ascontiguousarray(y, dtype=np.float64)
    tk = np.ascontiguousarray(tk, dtype=np.float64)
    assert L.uco_transport(pyoracle.ptr(r), C.c_int64(n), pyoracle.ptr(y), pyoracle.ptr(tk),
                           *[pyoracle.ptr(o[k]) for k in KEYS]) == 0
    return o


def test_gpu_transport_matches_the_program():
    z = np.load(GOLDEN / "transport.npz", allow_pickle=False)
    f = as_field(z["tr_y"], z["tr_tk"], 50)
    print(f"transport fixture: {f.n} cells of the reference program's /CB07/", flush=True)
    g = gpu(f)
    for k in KEYS:
        bad = g[k] != z[k]
        assert not bad.any(), (k, int(bad.sum()), float(np.max(np.abs(g[k] - z[k]) / np.abs(z[k]))))


def synthetic_states(m=4096, seed=5):
    rng = np.random.default_rng(seed)
    y = np.zeros((52, m))
    tk = np.empty(m)
    for c in range(m):
        kind = c % 8
        if kind == 0:          # many species above the 1e-3 threshold, up to all 51
            nbig = int(rng.integers(8, 52))
            sel = rng.choice(51, size=nbig, replace=False)
            y[sel, c] = rng.uniform(1e-3, 0.9 / nbig + 1e-3, nbig)
        elif kind == 1:        # a pure species, or pure N2
            k = int(rng.integers(0, 52))
            y[k, c] = 1.0
        elif kind == 2:        # all 51 species just above the threshold (LSPT = 52)
            y[:51, c] = rng.uniform(1e-3, 1.5e-2, 51)
        else:                  # a few majors and minors below the threshold
            nbig = int(rng.integers(1, 7))
            sel = rng.choice(51, size=nbig, replace=False)
            y[sel, c] = rng.dirichlet(np.ones(nbig)) * rng.uniform(0.05, 0.6)
            minors = rng.choice(51, size=20, replace=False)
            y[minors, c] += 10.0 ** rng.uniform(-12, -3.2, 20)
        s = y[:51, c].sum()
        if s > 0.999:
            y[:51, c] *= 0.999 / s if kind != 1 else 1.0
        if kind != 1 or y[51, c] == 0.0:
            y[51, c] = 1.0 - y[:51, c].sum()
        tk[c] = rng.uniform(200.0, 3500.0) / 294.0
    tk[:8] = np.array([999.999, 1000.0, 1000.001, 250.0, 300.0, 2500.0, 3499.0, 700.0]) / 294.0
    return y, tk


def test_gpu_transport_matches_oracle_beyond_the_flames():
    y, tk = synthetic_states()
    f = as_field(y, tk, 64)
    g = gpu(f)
    o = oracle(y, tk)
    for k in KEYS:
        bad = g[k] != o[k]
        assert not bad.any(), (k, int(bad.sum()))
    lspt = (y[:51] >= 1e-3).sum(axis=0) + 1
    assert lspt.min() == 1 and lspt.max() == 52
    assert np.isfinite(g["vdf"]).all() and (g["vmu"] > 0).all() and (g["vtc"] > 0).all()
    # the reference's zero read (DIFSP(LSPT, LSPT)) shows in pure N2 cells? No: N2 takes the
    # trace form; a pure species other than N2 reads DIFSP(k, N2) with X(N2) = 0
    pure = [c for c in range(tk.size) if (y[:, c] == 1.0).any() and y[51, c] == 0.0]
    assert pure, "no pure-species cell in the sample"
