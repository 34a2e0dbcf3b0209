"""Batches of grids through one set of launches (uc_step_batch).

Each grid of a batch is its own SPGSOLV call -- its own DT, controls and global
stops -- so a batch must give exactly what stepping each grid alone gives (the
same per-cell arithmetic; per-grid residual sums over the grid's own tiles in the
same order), and each grid must match the oracle to the parity bar.  The tuning
(stiffness bins, tile order, persistent waves, block size) changes only the order
in which cells are processed, never a result.
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import assert_state_close
from oracle import pyoracle

pytestmark = pytest.mark.gpu

# (li, lj, fuel, seed, dt seconds, pre-ignition band, mode)
GRIDS = [(31, 21, "CH4", 11, 1e-5, 0.0, 0), (45, 23, "C7H16", 12, 3e-6, 0.2, 0),
         (23, 17, "C2H4", 13, 1e-6, 0.0, 0), (61, 41, "C7H16", 14, 6e-6, 0.1, 0),
         (27, 29, "CH4", 15, 2e-6, 0.0, 2), (38, 19, "C2H4", 16, 1e-5, 0.3, 0)]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def make(spec):
    from unicorn_amd.fields import flame_field
    li, lj, fuel, seed, dt, band, mode = spec
    return flame_field(li, lj, fuel, seed, dt_seconds=dt, preignition_band=band), mode


def run_batch(fields, modes, nsteps, tuning=None, every_step=False):
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    devs = [DeviceField(f) for f in fields]
    ctx = ChemStep(sum(f.n for f in fields), max_grids=len(fields))
    if tuning:
        ctx.set_tuning(**tuning)
    ctls = [Controls(dt=f.dt, mode=m) for f, m in zip(fields, modes)]
    stats, states = [], []
    for _ in range(nsteps):
        stats.append(ctx.step_batch(devs, ctls))
        if every_step:
            states.append([d.to_host() for d in devs])
    if every_step:
        return states, stats, devs
    return [d.to_host() for d in devs], stats, devs


def run_alone(f, mode, nsteps):
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    dev = DeviceField(f)
    ctx = ChemStep(f.n)
    stats = [ctx.step(dev, Controls(dt=f.dt, mode=mode)) for _ in range(nsteps)]
    return dev.to_host(), stats, dev


def same(a, b):
    return all(np.array_equal(getattr(a, k), getattr(b, k)) for k in ("rho", "y", "tk", "ht", "stmf", "stnd"))


def test_batch_equals_grids_alone_and_matches_oracle(torch_cuda):
    made = [make(s) for s in GRIDS]
    fields = [f for f, _ in made]
    modes = [m for _, m in made]
    got, bstats, bdevs = run_batch([f.copy() for f in fields], modes, 3)
    for g, (f, mode) in enumerate(made):
        print(f"grid {g}: {f.li}x{f.lj} {f.fuel} dt {f.dt:.3e} mode {mode}", flush=True)
        alone, astats, adev = run_alone(f.copy(), mode, 3)
        assert [s[g].isor1 for s in bstats] == [s.isor1 for s in astats]
        assert [s[g].isor2 for s in bstats] == [s.isor2 for s in astats]
        assert [s[g].cell_sweeps for s in bstats] == [s.cell_sweeps for s in astats]
        assert same(got[g], alone), f"grid {g}: batch differs from the grid alone"
        assert np.array_equal(bdevs[g].qdot.cpu().numpy(), adev.qdot.cpu().numpy())
        assert np.array_equal(bdevs[g].qdsp.cpu().numpy(), adev.qdsp.cpu().numpy())
        ref = f.copy()
        outs = [pyoracle.step(ref, mode=mode) for _ in range(3)]
        if mode == pyoracle.MODE_GLOBAL:
            assert [s[g].isor1 for s in bstats] == [o.isor1 for o in outs]
            assert [s[g].isor2 for s in bstats] == [o.isor2 for o in outs]
        e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
        assert_state_close(got[g], e, f.interior(), f"batch grid {g}")


@pytest.mark.parametrize("keep", [False, True])
def test_batch_keeps_rates_where_not_formed(torch_cuda, keep):
    """QDOT and QDSP of a batch's grids change only at the cells SOLVEH forms them
    (interior, ISKIP < 3, or < 10 with kept rate coefficients); every other cell keeps the
    caller's values, and the formed ones equal the grid stepped alone."""
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    made = [make(s) for s in GRIDS[:3]]
    fields = [f for f, _ in made]
    for f in fields:
        f.iskip[f.iskip.size // 3: f.iskip.size // 3 + 40] = 5  # a band SOLVESP skips
    devs = [DeviceField(f.copy()) for f in fields]
    for d in devs:
        d.qdot.fill_(7.0)
        d.qdsp.fill_(-3.0)
    ctx = ChemStep(sum(f.n for f in fields), max_grids=len(fields))
    ctx.keep_rates(keep)
    ctx.step_batch(devs, [Controls(dt=f.dt) for f in fields])
    for g, f in enumerate(fields):
        ij = np.arange(f.n)
        i, j = ij // f.lj, ij % f.lj
        formed = (i > 0) & (i < f.li - 1) & (j > 0) & (j < f.lj - 1) & (f.iskip < (10 if keep else 3))
        q, qs = devs[g].qdot.cpu().numpy(), devs[g].qdsp.cpu().numpy()
        print(f"grid {g}: {formed.sum()} cells formed of {f.n}, kept rates {keep}", flush=True)
        assert np.all(q[~formed] == 7.0) and np.all(qs[:, ~formed] == -3.0)
        assert np.all(q[formed] != 7.0)
        alone = DeviceField(f.copy())
        actx = ChemStep(f.n)
        actx.keep_rates(keep)
        actx.step(alone, Controls(dt=f.dt))
        assert np.array_equal(q[formed], alone.qdot.cpu().numpy()[formed])
        assert np.array_equal(qs[:, formed], alone.qdsp.cpu().numpy()[:, formed])


@pytest.mark.parametrize("tuning", [dict(bins=1, lpt=0, persistent=0, block=256, graph=1),
                                    dict(bins=16, lpt=1, persistent=1, block=64, graph=1),
                                    dict(bins=32, lpt=0, persistent=1, block=128, graph=1),
                                    dict(bins=4, lpt=1, persistent=1, block=512, graph=1)])
def test_tuning_changes_no_result(torch_cuda, tuning):
    """Every step (each replays the captured sweep-segment graph one or more times, with
    the step's other launches and host synchronisations in between) against the same
    steps launched directly without graphs."""
    made = [make(s) for s in GRIDS[:4]]
    fields = [f for f, _ in made]
    modes = [m for _, m in made]
    eager = dict(bins=1, lpt=0, persistent=1, block=256, graph=0)
    base, bst, _ = run_batch([f.copy() for f in fields], modes, 4, eager, every_step=True)
    got, gst, _ = run_batch([f.copy() for f in fields], modes, 4, tuning, every_step=True)
    for step in range(4):
        for g in range(len(fields)):
            assert bst[step][g].isor1 == gst[step][g].isor1 and bst[step][g].isor2 == gst[step][g].isor2
            assert same(base[step][g], got[step][g]), (tuning, step, g)


def test_autotune_leaves_state_and_picks_a_valid_tuning(torch_cuda):
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    made = [make(s) for s in GRIDS[:3]]
    devs = [DeviceField(f) for f, _ in made]
    ctls = [Controls(dt=f.dt, mode=m) for f, m in made]
    before = [d.to_host() for d in devs]
    ctx = ChemStep(sum(f.n for f, _ in made), max_grids=len(made))
    t = ctx.autotune(devs, ctls)
    assert 1 <= t["bins"] <= 32 and t["block"] in (64, 128, 256, 512) and t["lpt"] in (0, 1) and t["persistent"] in (0, 1)
    assert t["graph"] in (0, 1) and t["lanes"] in (1, 2, 4) and t["rate_parts"] in (0, 1, 2, 4, 8)
    assert not (t["lanes"] > 1 and t["persistent"])
    assert ctx.tuning() == t
    for b, d in zip(before, devs):
        assert same(b, d.to_host())
