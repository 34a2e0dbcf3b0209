"""The species global stop (solvesp.f:2834: stop at the first sweep with RSD/SPNORM <=
TOLRSP) at near-ties, through the C ABI.

The device forms each sweep's RSD from its own per-cell terms |S - FNEW|, summed per tile
and then over tiles in a fixed order; the reference adds the same terms one by one, J
outer, I inner.  Near convergence a term is a small difference of nearly equal iterates,
so the last-bit differences of the device's FNEW (FMA contraction, device libm in the
rates) move single terms by up to about 1e-8 of their size.  Over the field those
differences, not the order of summation (about 1e-14 relative on these fields), set how
far the device's RSD may lie from the reference's.  The tests measure that distance
against the oracle (bit for bit with the reference on the fixtures; measured at most
9.3e-13 relative) and place TOLRSP 1e-9 and 1e-11 (relative) above and below the oracle's
own residual of a sweep: the device must stop where the oracle stops on both sides.
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import load
from oracle import pyoracle

pytestmark = pytest.mark.gpu

NOISE = 1e-11       # largest relative distance allowed between device and oracle residuals


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def gpu_step(f, tol):
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    dev = DeviceField(f)
    ctx = ChemStep(f.n)
    st = ctx.step(dev, Controls(dt=f.dt, tolrsp=tol))
    return st, ctx.sweep_residuals()


@pytest.mark.parametrize("name", ["cfg0", "c7band", "sooty"])
def test_sweep_residuals_match_the_oracle(torch_cuda, name):
    """The residual history the device's stop reads, sweep by sweep, against the oracle's."""
    f0, _ = load(name)
    out = pyoracle.step(f0.copy())
    st, hist = gpu_step(f0.copy(), 1e-8)
    assert st.isor1 == out.isor1 and len(hist) == out.isor1
    ref = out.rsd_hist[:out.isor1]
    rel = np.abs(hist - ref) / ref
    print(f"{name}: {len(hist)} sweeps, max relative distance to the oracle's RSD {rel.max():.2e}", flush=True)
    assert rel.max() <= NOISE, rel


@pytest.mark.parametrize("margin", [1e-9, 1e-11])
@pytest.mark.parametrize("name", ["cfg0", "c7band"])
def test_global_stop_at_near_ties(torch_cuda, name, margin):
    """TOLRSP set to the oracle's residual of sweep k = N*-1, scaled by 1 +- margin: the oracle
    stops at k above the tie and at k+1 below it, and so must the device."""
    f0, _ = load(name)
    base = pyoracle.step(f0.copy())
    k = base.isor1 - 1
    rk = base.rsd_hist[k - 1]
    for sign, want in ((+1, k), (-1, k + 1)):
        tol = rk * (1.0 + sign * margin)
        ref = pyoracle.step(f0.copy(), tolrsp=tol)
        assert ref.isor1 == want
        print(f"{name}: TOLRSP = R_{k} x (1 {'+' if sign > 0 else '-'} {margin:g}), oracle stops at {want}",
              flush=True)
        st, hist = gpu_step(f0.copy(), tol)
        assert st.isor1 == want, (name, sign, st.isor1, want)
        assert st.isor2 == ref.isor2
        print(f"  device: stops at {st.isor1}; R_{k} relative distance {abs(hist[k - 1] - rk) / rk:.2e}", flush=True)
