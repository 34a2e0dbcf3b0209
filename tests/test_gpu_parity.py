"""GPU parity of the gfx950 chemistry step, through the C ABI.

* against the reference's own outputs (tests/golden): sweep counts exact,
  Y to 1e-8 relative + 1e-14 absolute, T/HT/rho to 1e-8 relative, production
  rates after one step per value to 1e-10 relative (plus 1e-13 of the field's largest
  rate of that species, which only cancellation-dominated values lean on), after chained
  steps to 1e-10 of that largest rate;
* against the oracle (bit for bit with the reference on the fixtures) at every active
  cell of each fixture's field, on larger fields and in the per-cell stop mode;
* at the full 1001 x 401 C7H16 size: closure, positivity and bounds on every
  cell, and a 2 % random sample of cells re-run by the oracle with the GPU's
  sweep counts.
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


def gpu_run(field, nsteps, mode=0, ctx=None):
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    dev = DeviceField(field)
    ctx = ctx or ChemStep(field.n)
    stats = [ctx.step(dev, Controls(dt=field.dt, mode=mode)) for _ in range(nsteps)]
    return dev, stats


@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_reference_fixtures(torch_cuda, name):
    f0, z = load(name)
    inter = f0.interior()
    for ns in [int(s) for s in z["steps"]]:
        print(f"case {name} steps {ns}", flush=True)
        dev, stats = gpu_run(f0.copy(), ns)
        e = expected(z, ns)
        assert [s.isor1 for s in stats] == list(e["isor1"])
        assert [s.isor2 for s in stats] == list(e["isor2"])
        assert np.allclose([s.resdc for s in stats], e["resdc"], rtol=1e-9, atol=0)
        assert np.allclose([s.resdh for s in stats], e["resdh"], rtol=1e-9, atol=0)
        got = dev.to_host()
        assert_state_close(got, e, inter, f"{name}/{ns}")
        assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, z["dump_cells"]],
                           e["qdot"], e["qdsp_dump"], inter, f"{name}/{ns}", per_value=ns == 1)


@pytest.mark.parametrize("name", CASES)
def test_gpu_rates_per_value_at_every_cell(torch_cuda, name):
    """QDOT and QDSP of every active cell after one step from the fixture's input, against
    the oracle (whose step equals the reference's on these fields bit for bit), per value."""
    f0, z = load(name)
    ref = f0.copy()
    out = pyoracle.step(ref)
    print(f"case {name}: one step, rates of every cell", flush=True)
    dev, stats = gpu_run(f0.copy(), 1)
    assert (stats[0].isor1, stats[0].isor2) == (out.isor1, out.isor2)
    act = f0.active()
    assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, act], out.qdot, out.qdsp[:, act], act,
                       f"{name}: every cell")


@pytest.mark.parametrize("mode", [pyoracle.MODE_GLOBAL, pyoracle.MODE_PERCELL])
def test_gpu_matches_oracle_c7_band(torch_cuda, mode):
    from unicorn_amd.fields import flame_field
    f0 = flame_field(161, 61, "C7H16", 21, preignition_band=0.1)
    ref = f0.copy()
    outs = [pyoracle.step(ref, mode=mode) for _ in range(2)]
    dev, stats = gpu_run(f0.copy(), 2, mode=mode)
    inter = f0.interior()
    if mode == pyoracle.MODE_GLOBAL:
        assert [s.isor1 for s in stats] == [o.isor1 for o in outs]
        assert [s.isor2 for s in stats] == [o.isor2 for o in outs]
    assert stats[-1].cell_sweeps == outs[-1].cell_sweeps or mode == pyoracle.MODE_GLOBAL
    e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    assert_state_close(dev.to_host(), e, inter, f"c7 mode {mode}")
    assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, inter], outs[-1].qdot,
                       outs[-1].qdsp[:, inter], inter, f"c7 mode {mode}", per_value=False)


def test_gpu_full_size_properties_and_sample(torch_cuda):
    from unicorn_amd.fields import config1
    f0 = config1()
    dev, stats = gpu_run(f0.copy(), 1)
    st = stats[0]
    got = dev.to_host()
    inter = f0.interior()
    y = got.y[:, inter]
    assert np.all(y >= 0.0)
    assert np.allclose(y.sum(axis=0) + got.stmf[inter], 1.0, rtol=0, atol=1e-13)
    T = got.tk[inter] * 294.0
    assert T.min() >= 293.0 - 1e-9 and T.max() <= 5000.0
    # a 2 % sample through the oracle with the sweep counts the GPU decided
    rng = np.random.default_rng(2)
    cells = np.sort(rng.choice(inter, size=len(inter) // 50, replace=False))
    sub = f0.take(cells)
    isor1 = min(st.isor1, 100)
    isor2 = min(st.isor2, 100)
    out = pyoracle.step(sub, isor1=isor1, tolrsp=-1.0, isor2=isor2, tolrh=-1.0)
    si = sub.active()
    assert len(si) == len(cells)
    # sub's interior lists the sample cells in order; got.take lays them out alike
    e = {k: getattr(sub, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    got_sub = got.take(cells)
    for k in ("y", "tk", "ht", "stmf", "stnd"):
        a, b = getattr(got_sub, k), e[k]
        if k == "y":
            bad = np.abs(a[:, si] - b[:, si]) > 1e-8 * np.abs(b[:, si]) + 1e-14
            assert not bad.any(), (k, bad.sum())
        else:
            assert np.all(np.abs(a[si] - b[si]) <= 1e-8 * np.abs(b[si]) + 1e-300), k
    # the sample's production rates, per value
    assert_rates_close(dev.qdot.cpu().numpy()[cells], dev.qdsp.cpu().numpy()[:, cells], out.qdot[si],
                       out.qdsp[:, si], np.arange(cells.size), "1001x401, 2% sample, one step")


def _holed(field, seed, frac=0.3):
    """The field with ISKIP in 1..10 at a seeded 30 % of its interior cells (>= 3 skips
    the cell, ckrates.f:93)."""
    rng = np.random.default_rng(seed)
    f = field.copy()
    c = f.interior()
    pick = c[rng.random(c.size) < frac]
    f.iskip[pick] = rng.integers(1, 11, pick.size)
    return f


def test_gpu_iskip_holes_match_oracle(torch_cuda):
    """Skipped cells scattered over a 61 x 41 C7H16 field: the active-cell list built on
    the device gives the oracle's two steps (sweep counts exact, the state to the bar on
    the active cells), and the skipped cells come out as the oracle leaves them."""
    from unicorn_amd.fields import flame_field
    f0 = _holed(flame_field(61, 41, "C7H16", 31, preignition_band=0.1), 5)
    act = f0.active()
    skip = np.setdiff1d(f0.interior(), act)
    assert act.size > 0 and skip.size > 0.15 * f0.interior().size
    ref = f0.copy()
    outs = [pyoracle.step(ref) for _ in range(2)]
    dev, stats = gpu_run(f0.copy(), 2)
    assert [s.isor1 for s in stats] == [o.isor1 for o in outs]
    assert [s.isor2 for s in stats] == [o.isor2 for o in outs]
    e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    got = dev.to_host()
    assert_state_close(got, e, act, "iskip holes, active")
    assert_state_close(got, e, skip, "iskip holes, skipped")
    for name in ("tk", "ht", "stmf", "stnd"):
        assert np.array_equal(getattr(got, name)[skip], getattr(f0, name)[skip]), name


def test_gpu_device_and_host_cell_lists_agree(torch_cuda):
    """The same holed field through uc_step (its cell list built on the device) and as
    the second grid of a two-grid uc_step_batch (the list built on the host): the same
    results, bit for bit, over two steps."""
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import flame_field
    f0 = _holed(flame_field(61, 41, "C7H16", 32, preignition_band=0.1), 6)
    g0 = flame_field(41, 31, "CH4", 7)
    one, _ = gpu_run(f0.copy(), 2)
    ctx = ChemStep(f0.n + g0.n, max_grids=2)
    devs = [DeviceField(g0), DeviceField(f0)]
    for _ in range(2):
        ctx.step_batch(devs, [Controls(dt=g0.dt), Controls(dt=f0.dt)])
    a, b = one.to_host(), devs[1].to_host()
    for name in ("rho", "y", "tk", "ht", "stmf", "stnd"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert np.array_equal(one.qdot.cpu().numpy()[f0.active()], devs[1].qdot.cpu().numpy()[f0.active()])
