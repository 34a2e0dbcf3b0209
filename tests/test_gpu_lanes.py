"""Lanes per cell of the species sweep (uc_tuning.lanes) and of the rate-coefficient
kernel (uc_tuning.rate_parts), through the C ABI.

With W lanes per cell the sweep runs in W waves of one workgroup, one lane of each per
cell: every species block's reaction terms are divided among the W waves, the partial
sums meet in LDS and every wave applies the same update.  The production and loss sums
are therefore added in another order than the one-thread-per-cell form adds them, so the
split forms are held to the parity bar -- against the reference's own outputs and against
the oracle, sweep counts exact -- and not bit for bit against the other form.

The rate kernel's parts each form whole slots by the one-lane expressions, so a step with
any rate_parts must give the one-lane step's bits.
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import CASES, assert_rates_close, assert_state_close, expected, load
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def ctx_with(n, max_grids=1, **tuning):
    from unicorn_amd.chem import ChemStep
    ctx = ChemStep(n, max_grids=max_grids)
    t = dict(bins=1, lpt=0, persistent=0, block=256, graph=1, lanes=1, rate_parts=1)
    t.update(tuning)
    ctx.set_tuning(**t)
    return ctx


@pytest.mark.parametrize("lanes", [2, 4])
@pytest.mark.parametrize("name", CASES)
def test_split_sweep_matches_reference_fixtures(torch_cuda, name, lanes):
    from unicorn_amd.chem import Controls, DeviceField
    f0, z = load(name)
    inter = f0.interior()
    for ns in [int(s) for s in z["steps"]]:
        print(f"case {name} steps {ns} lanes {lanes}", flush=True)
        dev = DeviceField(f0.copy())
        ctx = ctx_with(f0.n, lanes=lanes)
        stats = [ctx.step(dev, Controls(dt=f0.dt)) for _ in range(ns)]
        e = expected(z, ns)
        assert [s.isor1 for s in stats] == list(e["isor1"])
        assert [s.isor2 for s in stats] == list(e["isor2"])
        assert np.allclose([s.resdc for s in stats], e["resdc"], rtol=1e-9, atol=0)
        assert np.allclose([s.resdh for s in stats], e["resdh"], rtol=1e-9, atol=0)
        assert_state_close(dev.to_host(), e, inter, f"{name}/{ns} lanes {lanes}")
        assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, z["dump_cells"]],
                           e["qdot"], e["qdsp_dump"], inter, f"{name}/{ns} lanes {lanes}", per_value=ns == 1)


@pytest.mark.parametrize("lanes", [2, 4])
def test_split_sweep_rates_per_value_at_every_cell(torch_cuda, lanes):
    """One step of the soot fixture's field: QDOT and QDSP of every active cell against the
    oracle, per value."""
    from unicorn_amd.chem import Controls, DeviceField
    f0, _ = load("sooty")
    ref = f0.copy()
    out = pyoracle.step(ref)
    print(f"sooty, one step, lanes {lanes}: rates of every cell", flush=True)
    dev = DeviceField(f0.copy())
    st = ctx_with(f0.n, lanes=lanes).step(dev, Controls(dt=f0.dt))
    assert (st.isor1, st.isor2) == (out.isor1, out.isor2)
    act = f0.active()
    assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, act], out.qdot, out.qdsp[:, act], act,
                       f"sooty lanes {lanes}: every cell")


@pytest.mark.parametrize("lanes", [2, 4])
def test_split_sweep_batch_matches_oracle(torch_cuda, lanes):
    """Grids of different sizes, fuels, DT and stop modes in one batch, stiffness-binned,
    three chained steps: every grid against the oracle."""
    from unicorn_amd.chem import Controls, DeviceField
    from unicorn_amd.fields import flame_field
    specs = [(31, 21, "CH4", 11, 1e-5, 0.0, 0), (45, 23, "C7H16", 12, 3e-6, 0.2, 0),
             (27, 29, "CH4", 15, 2e-6, 0.0, 2), (38, 19, "C2H4", 16, 1e-5, 0.3, 0)]
    fields = [flame_field(li, lj, fu, sd, dt_seconds=dt, preignition_band=b) for li, lj, fu, sd, dt, b, _ in specs]
    modes = [s[-1] for s in specs]
    devs = [DeviceField(f.copy()) for f in fields]
    ctx = ctx_with(sum(f.n for f in fields), max_grids=len(fields), lanes=lanes, bins=16)
    ctls = [Controls(dt=f.dt, mode=m) for f, m in zip(fields, modes)]
    print(f"batch of {len(fields)} grids, lanes {lanes}", flush=True)
    stats = [ctx.step_batch(devs, ctls) for _ in range(3)]
    for g, (f, mode) in enumerate(zip(fields, modes)):
        ref = f.copy()
        outs = [pyoracle.step(ref, mode=mode) for _ in range(3)]
        if mode == pyoracle.MODE_GLOBAL:
            assert [s[g].isor1 for s in stats] == [o.isor1 for o in outs]
            assert [s[g].isor2 for s in stats] == [o.isor2 for o in outs]
        e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
        assert_state_close(devs[g].to_host(), e, f.interior(), f"batch grid {g} lanes {lanes}")


@pytest.mark.parametrize("parts", [2, 4, 8])
def test_rate_parts_same_bits(torch_cuda, parts):
    """Two chained steps of the soot fixture's field and a three-grid batch with the rate
    kernel in parts: every output array equal, bit for bit, to the one-lane run's."""
    from unicorn_amd.chem import Controls, DeviceField
    from unicorn_amd.fields import flame_field
    f0, _ = load("sooty")
    fields = [f0, flame_field(45, 23, "C7H16", 12, dt_seconds=3e-6, preignition_band=0.2),
              flame_field(27, 29, "CH4", 15, dt_seconds=2e-6)]
    runs = []
    for rp in (1, parts):
        print(f"rate_parts {rp}", flush=True)
        single = DeviceField(f0.copy())
        ctx = ctx_with(f0.n, persistent=1, rate_parts=rp)
        s1 = [ctx.step(single, Controls(dt=f0.dt)) for _ in range(2)]
        devs = [DeviceField(f.copy()) for f in fields]
        bctx = ctx_with(sum(f.n for f in fields), max_grids=len(fields), bins=16, rate_parts=rp)
        s2 = [bctx.step_batch(devs, [Controls(dt=f.dt) for f in fields]) for _ in range(2)]
        runs.append((single, s1, devs, s2))
    (a, sa, da, ba), (b, sb, db, bb) = runs
    assert [(s.isor1, s.isor2, s.resdc, s.resdh) for s in sa] == [(s.isor1, s.isor2, s.resdc, s.resdh) for s in sb]
    assert [[(g.isor1, g.isor2) for g in s] for s in ba] == [[(g.isor1, g.isor2) for g in s] for s in bb]
    for x, y in [(a, b)] + list(zip(da, db)):
        for k in ("rho", "y", "tk", "ht", "stmf", "stnd", "qdot", "qdsp"):
            u, v = getattr(x, k), getattr(y, k)
            assert torch_cuda.equal(u, v), (parts, k)


def test_lanes_refused_out_of_range(torch_cuda):
    from unicorn_amd.chem import ChemError, ChemStep
    ctx = ChemStep(1000)
    for bad in (dict(lanes=3, persistent=0), dict(lanes=2, persistent=1), dict(lanes=8, persistent=0),
                dict(rate_parts=3), dict(rate_parts=16), dict(rate_parts=-1)):
        t = dict(bins=1, lpt=0, persistent=0, block=256, graph=1, lanes=1, rate_parts=1)
        t.update(bad)
        with pytest.raises(ChemError):
            ctx.set_tuning(**t)
    assert ctx.tuning()["lanes"] == 1 and ctx.tuning()["rate_parts"] == 0
