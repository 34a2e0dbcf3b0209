"""The seeded synthetic inputs (unicorn_amd.fields): the coupling generator's
properties that the coupled tests and the bench rely on.  CPU only."""
from __future__ import annotations

import numpy as np

from unicorn_amd.fields import flame_field, synthetic_coupling


def test_synthetic_coupling_is_a_diffusion_scheme():
    f = flame_field(23, 17, "CH4", 5)
    cp = synthetic_coupling(f, seed=9, strength=0.1)
    li, lj = f.li, f.lj
    coef = cp.coef.reshape(51, 4, li, lj)
    # every face coefficient is non-positive, so PPP = RHO - (E+W+P+Q) >= RHO: each cell's
    # equation is diagonally dominant and the Gauss-Seidel sweeps converge
    for a in (coef, cp.stc, cp.zc):
        assert np.all(a <= 0.0)
    e, w, p, q = coef[:, 0], coef[:, 1], coef[:, 2], coef[:, 3]
    rho = f.rho.reshape(li, lj)
    # a face is shared: E of cell i and W of cell i+1 come from one conductance, scaled by
    # each cell's own density (strength * rho * g)
    ratio_x = (e[:, :-1, :] / rho[None, :-1, :]) / (w[:, 1:, :] / rho[None, 1:, :])
    ratio_y = (p[:, :, :-1] / rho[None, :, :-1]) / (q[:, :, 1:] / rho[None, :, 1:])
    assert np.allclose(ratio_x, 1.0) and np.allclose(ratio_y, 1.0)
    # the strength sets the off-diagonal weight: |E+W+P+Q| / RHO in [0.1*4*0.5*0.5, 0.1*4*1.5*2]
    off = -(e + w + p + q) / rho[None]
    assert off.min() >= 0.1 * 4 * 0.25 - 1e-12 and off.max() <= 0.1 * 4 * 3.0 + 1e-12
    # geometry: uniform spacing on the unit square, radius J*dy, velocities bounded
    assert np.allclose(cp.xxc, 1.0 / (li - 1)) and np.allclose(cp.yys, 1.0 / (lj - 1))
    assert np.allclose(cp.ygeo.reshape(li, lj)[:, 3], 3.0 / (lj - 1))
    assert np.abs(cp.u).max() <= 0.3 and np.all(cp.w == 0.0)


def test_synthetic_coupling_flux_scale_is_linear():
    """The flux inputs are scaled so the median |DT/2 x flux divergence| is `flux` times the
    median |RHO*HT|: doubling `flux` doubles VDF, TMU and EKT and leaves the rest."""
    f = flame_field(21, 15, "C2H4", 4)
    a = synthetic_coupling(f, seed=2, flux=1e-3)
    b = synthetic_coupling(f, seed=2, flux=2e-3)
    for k in ("vdf", "tmu", "ekt"):
        assert np.allclose(getattr(b, k), 2.0 * getattr(a, k))
    for k in ("coef", "stc", "zc", "u", "v"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    assert np.all(a.vdf > 0.0) and np.all(a.tmu >= 0.0) and np.all(a.ekt >= 0.0)


def _convection_share(f, tr):
    """Share of (E face, species) pairs where |rho u| exceeds the face conductance
    (spgsolv.f:233-349), i.e. where the hybrid scheme upwinds."""
    from unicorn_amd.scales import Scales
    sc = Scales()
    lj = f.lj
    emu = sc.rei * sc.beta3 * f.rho
    inter = f.interior()
    ce = inter + lj
    i = inter // lj
    rhoue = 0.5 * (f.rho[ce] + f.rho[inter]) * tr.u[ce]
    tur = sc.rei * (tr.tmu[ce] + tr.tmu[inter]) / tr.sigsp
    d = (emu[ce] * tr.vdf[:51, ce] + emu[inter] * tr.vdf[:51, inter] + tur) / tr.xxs[i + 1]
    return float(np.mean(np.abs(rhoue)[None, :] > d))


def test_synthetic_transport_regimes():
    """The transport generator's inputs are positive and shaped for the field; at velocity
    0.3 every face is diffusion-dominated, at velocity 5 a good share is convection-
    dominated (the two regimes the assembly fixtures pin)."""
    from unicorn_amd.fields import flame_field, synthetic_transport
    f = flame_field(25, 17, "CH4", 13)
    slow = synthetic_transport(f, seed=19, velocity=0.3)
    fast = synthetic_transport(f, seed=21, velocity=5.0)
    for tr in (slow, fast):
        assert tr.vdf.shape == (52, f.n) and tr.xxc.shape == (f.li,) and tr.yys.shape == (f.lj,)
        assert (tr.vdf > 0).all() and (tr.tmu >= 0).all() and (tr.vtc >= 0).all() and (tr.stdf >= 0).all()
        assert np.abs(tr.u).max() > 0 and np.abs(tr.v).max() > 0
    assert _convection_share(f, slow) == 0.0
    assert 0.1 < _convection_share(f, fast) < 0.9
