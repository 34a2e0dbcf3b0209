"""GPU parity of the coupled chemistry step (uc_step_coupled), through the C ABI.

The coupled step runs the reference's lexicographic Gauss-Seidel with the transport
coefficients SPGSOLV assembles (E/W/P/Q of every species, STE..STS, ZE..ZS), the
species-diffusion enthalpy flux of RHSZ and the kinetic energy of T(h), as a
pipelined anti-diagonal wavefront.

* against the flang-built reference on the coupled fixture (tests/golden/coupled.npz,
  ISKIP holes, soot): sweep counts exact, the state to the parity bar, over 1 and 3
  chained steps;
* against the oracle (bit-exact with the reference) on larger fields with other seeds;
* with zero coefficients the wavefront path equals the chemistry-only step;
* velocities alone (no coefficients) run the chemistry-only sweeps with the kinetic
  energy in T(h).
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import (COUPLED_CASES, assert_rates_close, assert_state_close, expected, load,
                       load_coupling)
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def coupled_run(field, cp, nsteps, keep=None, ctx=None):
    from unicorn_amd.chem import ChemStep, Controls, DeviceCoupling, DeviceField
    dev = DeviceField(field)
    dcp = DeviceCoupling(cp, keep=keep)
    ctx = ctx or ChemStep(field.n)
    stats = [ctx.step_coupled(dev, Controls(dt=field.dt), dcp) for _ in range(nsteps)]
    return dev, stats


@pytest.mark.parametrize("name", COUPLED_CASES)
def test_gpu_coupled_matches_reference_fixture(torch_cuda, name):
    f0, z = load(name)
    cp = load_coupling(z)
    inter = f0.interior()
    for ns in [int(s) for s in z["steps"]]:
        print(f"case {name} steps {ns}", flush=True)
        dev, stats = coupled_run(f0.copy(), cp, ns)
        e = expected(z, ns)
        assert [s.isor1 for s in stats] == list(e["isor1"])
        assert [s.isor2 for s in stats] == list(e["isor2"])
        assert np.allclose([s.resdc for s in stats], e["resdc"], rtol=1e-9, atol=0)
        assert np.allclose([s.resdh for s in stats], e["resdh"], rtol=1e-9, atol=0)
        got = dev.to_host()
        assert_state_close(got, e, inter, f"{name}/{ns}")
        assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, z["dump_cells"]],
                           e["qdot"], e["qdsp_dump"], inter, f"{name}/{ns}", per_value=ns == 1)


@pytest.mark.parametrize("name", COUPLED_CASES)
def test_gpu_coupled_rates_per_value_at_every_cell(torch_cuda, name):
    """One coupled step from the fixture's input: QDOT and QDSP of every active cell against
    the oracle (bit for bit with the reference on this fixture), per value.  In the coupled
    sweeps each cell reads its neighbours' current iterates, so last-bit differences travel
    through the grid within the step: up to one value in 10^3 (measured: 11 of 17,472) lean
    on the cancellation floor, against one in 10^4 for the chemistry-only step."""
    f0, z = load(name)
    cp = load_coupling(z)
    ref = f0.copy()
    out = pyoracle.step(ref, coupling=cp)
    print(f"coupled {name}: one step, rates of every cell", flush=True)
    dev, stats = coupled_run(f0.copy(), cp, 1)
    assert (stats[0].isor1, stats[0].isor2) == (out.isor1, out.isor2)
    act = f0.active()
    assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, act], out.qdot, out.qdsp[:, act], act,
                       f"coupled {name}: every cell", max_over=1e-3)


@pytest.mark.parametrize("fuel,li,lj,seed", [("CH4", 61, 41, 31), ("C2H4", 47, 29, 32)])
def test_gpu_coupled_matches_oracle(torch_cuda, fuel, li, lj, seed):
    from unicorn_amd.fields import flame_field, synthetic_coupling
    f0 = flame_field(li, lj, fuel, seed)
    rng = np.random.default_rng(seed)
    holes = rng.choice(f0.interior(), size=f0.n // 40, replace=False)
    f0.iskip[holes] = 10
    cp = synthetic_coupling(f0, seed=seed + 1)
    ref = f0.copy()
    outs = [pyoracle.step(ref, coupling=cp) for _ in range(2)]
    print(f"coupled {fuel} {li}x{lj} seed {seed}", flush=True)
    dev, stats = coupled_run(f0.copy(), cp, 2)
    assert [s.isor1 for s in stats] == [o.isor1 for o in outs]
    assert [s.isor2 for s in stats] == [o.isor2 for o in outs]
    got = dev.to_host()
    inter = f0.interior()
    e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    assert_state_close(got, e, inter, f"{fuel}")
    cells = f0.active()
    assert_rates_close(dev.qdot.cpu().numpy(), dev.qdsp.cpu().numpy()[:, cells], outs[-1].qdot,
                       outs[-1].qdsp[:, cells], inter, f"{fuel}", per_value=False)
    # skipped cells keep their state
    assert np.array_equal(got.y[:, holes], ref.y[:, holes])


def test_gpu_coupled_zero_coefficients_equal_the_chemistry_only_step(torch_cuda):
    """Zero coefficient arrays take the wavefront path; the result is the chemistry-only
    step's (the neighbour terms vanish exactly), with the same sweep counts."""
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import flame_field, synthetic_coupling
    f0 = flame_field(41, 27, "C2H4", 5)
    cp = synthetic_coupling(f0, seed=6).zero_like()
    print("zero coefficients, wavefront vs chemistry-only", flush=True)
    dev, stats = coupled_run(f0.copy(), cp, 2, keep=("coef", "stc", "zc"))
    ref = DeviceField(f0.copy())
    ctx = ChemStep(f0.n)
    rs = [ctx.step(ref, Controls(dt=f0.dt)) for _ in range(2)]
    assert [s.isor1 for s in stats] == [s.isor1 for s in rs]
    assert [s.isor2 for s in stats] == [s.isor2 for s in rs]
    a, b = dev.to_host(), ref.to_host()
    inter = f0.interior()
    e = {k: getattr(b, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    assert_state_close(a, e, inter, "zero coefficients")
    same = all(np.array_equal(getattr(a, k)[..., inter], getattr(b, k)[..., inter])
               for k in ("y", "tk", "ht", "stmf", "stnd"))
    print("bitwise equal:", same)


def test_gpu_velocities_enter_the_temperature(torch_cuda):
    """Velocities alone: chemistry-only sweeps, ENTH = (FPZ - (U*U+V*V+W*W)/2)*HSTR + ...
    in the Newton T(h) (solveh.f:1176-1177), against the oracle."""
    from unicorn_amd.fields import Coupling, flame_field, synthetic_coupling
    f0 = flame_field(41, 27, "CH4", 8)
    full = synthetic_coupling(f0, seed=9, velocity=0.8)
    cp = full.zero_like()
    ref = f0.copy()
    o = pyoracle.step(ref, coupling=Coupling(**{**cp.__dict__, "coef": None, "stc": None, "zc": None,
                                                "ekt": None, "tmu": None, "vdf": None}))
    print("velocities only", flush=True)
    dev, stats = coupled_run(f0.copy(), cp, 1, keep=("u", "v", "w"))
    assert (stats[0].isor1, stats[0].isor2) == (o.isor1, o.isor2)
    got = dev.to_host()
    inter = f0.interior()
    e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    assert_state_close(got, e, inter, "velocities")
    # and the kinetic energy matters: without it T differs by far more than the parity bar
    plain = f0.copy()
    pyoracle.step(plain)
    assert np.abs(plain.tk[inter] - got.tk[inter]).max() > 100 * 1e-8 * np.abs(plain.tk[inter]).max()


@pytest.mark.parametrize("controls", [dict(relxsp=0.6, relxh=0.3), dict(isor1=7, isor2=5),
                                      dict(relxh=0.2, isor2=200)],
                         ids=["two-segment-stops", "sweep-limits", "long-enthalpy"])
def test_gpu_coupled_segments_and_limits(torch_cuda, controls):
    """The coupled step's segment logic against the oracle over two chained steps: global
    stops past one 48-sweep segment (87 then 42 species sweeps, 84 enthalpy sweeps), a
    predicted first segment that falls short, the sweep limits reached (ISOR+1 returned),
    and an enthalpy relaxation of 130 sweeps (three segments)."""
    from unicorn_amd.chem import ChemStep, Controls, DeviceCoupling, DeviceField
    from unicorn_amd.fields import flame_field, synthetic_coupling
    f0 = flame_field(31, 23, "CH4", 12)
    cp = synthetic_coupling(f0, seed=13)
    ref = f0.copy()
    outs = [pyoracle.step(ref, coupling=cp, **controls) for _ in range(2)]
    print(f"coupled segments {controls}", flush=True)
    dev, dcp, ctx = DeviceField(f0.copy()), DeviceCoupling(cp), ChemStep(f0.n)
    stats = [ctx.step_coupled(dev, Controls(dt=f0.dt, **controls), dcp) for _ in range(2)]
    assert [(s.isor1, s.isor2) for s in stats] == [(o.isor1, o.isor2) for o in outs]
    got = dev.to_host()
    inter = f0.interior()
    e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    assert_state_close(got, e, inter, f"segments {controls}")


def test_gpu_coupled_five_chained_steps_match_oracle(torch_cuda):
    """Five chained coupled steps (51 x 37 C2H4 field, soot, holes) stay within the parity
    bar of the oracle, with the same sweep counts at every step."""
    from unicorn_amd.fields import flame_field, synthetic_coupling
    f0 = flame_field(51, 37, "C2H4", 41)
    rng = np.random.default_rng(41)
    sel = rng.random(f0.n) < 0.3
    f0.stmf[sel] = 10.0 ** rng.uniform(-7, -4, sel.sum())
    f0.stnd[sel] = 10.0 ** rng.uniform(12, 16, sel.sum())
    f0.iskip[rng.choice(f0.interior(), size=30, replace=False)] = 10
    cp = synthetic_coupling(f0, seed=42)
    ref = f0.copy()
    outs = [pyoracle.step(ref, coupling=cp) for _ in range(5)]
    print("coupled five chained steps", [(o.isor1, o.isor2) for o in outs], flush=True)
    dev, stats = coupled_run(f0.copy(), cp, 5)
    assert [(s.isor1, s.isor2) for s in stats] == [(o.isor1, o.isor2) for o in outs]
    got = dev.to_host()
    inter = f0.interior()
    e = {k: getattr(ref, k) for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    assert_state_close(got, e, inter, "five chained coupled steps")


def test_gpu_coupled_full_size_invariants(torch_cuda):
    """The 1001 x 401 C7H16 field with coupling (the bench's recipe), two steps: every
    active cell closed (sum Y + STMF = 1 to rounding), Y >= 0, T within the Newton clamp,
    the run repeated gives the same bits, and the coupling changes the result against the
    chemistry-only step.  (The serial oracle takes minutes per step at this size.)"""
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import config1, synthetic_coupling
    f0 = config1()
    cp = synthetic_coupling(f0, seed=3, strength=0.02, flux=2e-4, velocity=0.3)
    print("coupled full size", flush=True)
    runs = [coupled_run(f0.copy(), cp, 2) for _ in range(2)]
    (dev, stats), (dev2, stats2) = runs
    assert [(s.isor1, s.isor2) for s in stats] == [(s.isor1, s.isor2) for s in stats2]
    a, b = dev.to_host(), dev2.to_host()
    assert np.array_equal(a.y, b.y) and np.array_equal(a.tk, b.tk) and np.array_equal(a.ht, b.ht)
    cells = f0.active()
    s = a.y[:, cells].sum(axis=0) + a.stmf[cells]
    assert np.abs(s - 1.0).max() <= 1e-12
    assert (a.y[:, cells] >= 0.0).all()
    T = a.tk[cells] * 294.0
    assert T.min() >= 293.0 - 1e-9 and T.max() <= 5000.0 + 1e-9
    plain = DeviceField(f0.copy())
    ctx = ChemStep(f0.n)
    for _ in range(2):
        ctx.step(plain, Controls(dt=f0.dt))
    dT = np.abs(plain.to_host().tk[cells] - a.tk[cells]) * 294.0
    assert dT.max() > 1.0, dT.max()
