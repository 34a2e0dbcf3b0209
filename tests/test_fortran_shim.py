"""The Fortran drop-in: libunicorn_fshim.so exports CKRATES/SOLVESP/SOLVEH
(ckrates_, solvesp_, solveh_) over the reference's COMMON blocks.

oracle/_ref/ref_driver_shim is the reference driver (oracle/ref_driver.c) linked
against that library instead of the flang-built routines, so these tests run
the shipped Fortran ABI end to end: COMMON layout, argument lists, ISOR1/ISOR2
in/out, the SPGSOLV epilogue and density update left to the caller.  On the GPU
its outputs must match the reference's own (the golden fixtures) to the same
bar as the device API, with and without the transport coefficients, enthalpy flux
inputs and velocities SPGSOLV leaves in the COMMON blocks.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from golden_io import (CASES, COUPLED_CASES, assert_rates_close, assert_state_close, expected, load,
                       load_coupling)
from oracle import refrun
from unicorn_amd.fields import ChemField

ROOT = Path(__file__).resolve().parents[1]
SHIM_LIB = ROOT / "unicorn_amd" / "lib" / "libunicorn_fshim.so"
COMMONS = ("cb01_", "cb02_", "cb03_", "cb04_", "cb06_", "hform_", "hrls_", "dumd_", "soot_", "dummy_",
           "dum0_", "dumx_", "dumy_", "dums_")


def _nm(args):
    return subprocess.run(["nm", "-D", *args, str(SHIM_LIB)], check=True, capture_output=True,
                          text=True).stdout


def test_shim_exports_fortran_names_and_binds_commons():
    if not SHIM_LIB.exists():
        pytest.skip("libunicorn_fshim.so not built")
    defined = {l.split()[-1] for l in _nm(["--defined-only"]).splitlines() if l.strip()}
    undefined = {l.split()[-1] for l in _nm(["--undefined-only"]).splitlines() if l.strip()}
    assert {"ckrates_", "solvesp_", "solveh_"} <= defined
    # the COMMON blocks belong to the Fortran program, not to the library
    assert set(COMMONS) <= undefined
    assert not (set(COMMONS) & defined)


def test_shim_without_gpu_stops_with_a_message():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    if not refrun.available("shim"):
        pytest.skip("ref_driver_shim not built")
    f0, _ = load("sooty")
    with pytest.raises(subprocess.CalledProcessError) as e:
        refrun.run(f0, nsteps=1, driver="shim")
    assert e.value.returncode == 3
    assert "unicorn_fshim: ckrates_:" in e.value.stderr


def _compare(r, f0, e, dump, label):
    inter = f0.interior()
    assert list(r.isor1) == list(e["isor1"]), label
    assert list(r.isor2) == list(e["isor2"]), label
    assert np.allclose(r.resdc, e["resdc"], rtol=1e-9, atol=0)
    assert np.allclose(r.resdh, e["resdh"], rtol=1e-9, atol=0)
    got = ChemField(f0.li, f0.lj, r.rho, r.y, r.tk, r.ht, r.stmf, r.stnd, f0.iskip, f0.dt, f0.fuel)
    assert_state_close(got, e, inter, label)
    assert_rates_close(r.qdot, r.qdsp[:, dump], e["qdot"], e["qdsp_dump"], inter, label,
                       per_value=len(r.isor1) == 1)


@pytest.mark.gpu
@pytest.mark.parametrize("probe", [refrun.FLAG_PROBE_FLUX, refrun.FLAG_PROBE_COEF])
def test_shim_single_coefficient_probes_match_the_reference(probe):
    """One non-zero VDF (with SPGSOLV's EMU: an enthalpy flux around one cell) or one
    non-zero /DUMMY/ coefficient: the drop-in takes the coupled path and matches the
    flang-built routines run on the same COMMON blocks."""
    if not (refrun.available("shim") and refrun.available("ref")):
        pytest.skip("reference and shim drivers not built")
    f0, z = load("sooty")
    dump = z["dump_cells"]
    ref = refrun.run(f0.copy(), nsteps=1, probe=probe)
    e = dict(isor1=ref.isor1, isor2=ref.isor2, resdc=ref.resdc, resdh=ref.resdh, rho=ref.rho, y=ref.y,
             tk=ref.tk, ht=ref.ht, stmf=ref.stmf, stnd=ref.stnd, qdot=ref.qdot, qdsp_dump=ref.qdsp[:, dump])
    print(f"shim probe {probe}", flush=True)
    r = refrun.run(f0.copy(), nsteps=1, driver="shim", probe=probe)
    _compare(r, f0, e, dump, f"probe {probe}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", COUPLED_CASES)
def test_shim_matches_coupled_fixture(name):
    """The coupled fixture (every coefficient, the flux inputs, geometry and velocities
    in the COMMON blocks) through the drop-in, against the reference's outputs."""
    if not refrun.available("shim"):
        pytest.skip("ref_driver_shim not built")
    f0, z = load(name)
    cp = load_coupling(z)
    for ns in [int(s) for s in z["steps"]]:
        print(f"shim coupled {name} steps {ns}", flush=True)
        r = refrun.run(f0.copy(), nsteps=ns, driver="shim", coupling=cp)
        _compare(r, f0, expected(z, ns), z["dump_cells"], f"shim {name}/{ns}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_shim_matches_reference_fixtures(name):
    if not refrun.available("shim"):
        pytest.skip("ref_driver_shim not built")
    f0, z = load(name)
    inter = f0.interior()
    for ns in [int(s) for s in z["steps"]]:
        print(f"shim case {name} steps {ns}", flush=True)
        r = refrun.run(f0.copy(), nsteps=ns, driver="shim")
        e = expected(z, ns)
        assert list(r.isor1) == list(e["isor1"])
        assert list(r.isor2) == list(e["isor2"])
        assert np.allclose(r.resdc, e["resdc"], rtol=1e-9, atol=0)
        assert np.allclose(r.resdh, e["resdh"], rtol=1e-9, atol=0)
        got = ChemField(f0.li, f0.lj, r.rho, r.y, r.tk, r.ht, r.stmf, r.stnd, f0.iskip, f0.dt, f0.fuel)
        assert_state_close(got, e, inter, f"shim {name}/{ns}")
        dump = z["dump_cells"]
        assert_rates_close(r.qdot, r.qdsp[:, dump], e["qdot"], e["qdsp_dump"], inter, f"shim {name}/{ns}",
                           per_value=ns == 1)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["assemble", "assemble_conv"])
def test_shim_spgsolv_matches_reference(name):
    """The drop-in's spgsolv_ (coefficient assembly on the device, then the coupled
    chemistry, then SPGSOLV's epilogue) against the flang-built spgsolv.f on the same
    COMMON blocks (tests/golden/assemble.npz's field and transport inputs), three chained
    calls: sweep counts exact, RESDC/RESDH, the state and QDOT to the parity bar.  Both normalise the global stops
    by the Fortran build's 711 x 131 grid (spgsolv.f:100-101)."""
    from golden_io import load_transport
    if not (refrun.available("shim") and refrun.available("asm")):
        pytest.skip("reference SPGSOLV and shim drivers not built")
    f0, z = load(name)
    tr = load_transport(z)
    # three chained calls: the drop-in keeps QDOT/QDSP resident after the first upload,
    # predicts the first sweep segment from the previous call's count and rebuilds its cell
    # list and coefficient buffers every call
    ref = refrun.spgsolv(f0.copy(), tr, driver="asm", ncalls=3)
    print("shim spgsolv_", ref["isor1"], ref["isor2"], flush=True)
    got = refrun.spgsolv(f0.copy(), tr, driver="shim", ncalls=3)
    inter = f0.interior()
    print("shim", got["isor1"], got["isor2"], "max |dY|", np.abs(got["y"][:, inter] - ref["y"][:, inter]).max(),
          "max |dT|/T", np.abs(got["tk"][inter] - ref["tk"][inter]).max() / np.abs(ref["tk"][inter]).max(),
          "max |dHT|", np.abs(got["ht"][inter] - ref["ht"][inter]).max(), flush=True)
    assert (got["isor1"], got["isor2"]) == (ref["isor1"], ref["isor2"])
    assert np.isclose(got["resdc"], ref["resdc"], rtol=1e-9, atol=0)
    assert np.isclose(got["resdh"], ref["resdh"], rtol=1e-9, atol=0)
    g = ChemField(f0.li, f0.lj, f0.rho, got["y"], got["tk"], got["ht"], got["stmf"], got["stnd"], f0.iskip, f0.dt,
                  f0.fuel)
    assert_state_close(g, {k: ref[k] for k in ("y", "tk", "ht", "stmf", "stnd")} | {"rho": f0.rho}, inter,
                       "spgsolv_")
    sc = np.abs(ref["qdot"][inter]).max()
    assert np.all(np.abs(got["qdot"][inter] - ref["qdot"][inter]) <= 1e-10 * sc), "spgsolv_ QDOT"


@pytest.mark.gpu
def test_shim_stale_rate_cells_match_reference():
    """Cells with 3 <= ISKIP < 10: CKRATES and SOLVESP skip them (ckrates.f:94,
    solvesp.f:89), but SOLVEH still forms their QDOT and QDSP (solveh.f:119-1142), from the
    species SPGSOLV's prologue copies and from whatever /DUMR/ RC holds -- the coefficients
    of the last call that advanced the cell, zero if none did.  Three chained spgsolv_ calls
    on the assemble fixture: a twentieth of the active cells are advanced by the first call
    and skipped (ISKIP 5) by the later ones, another twentieth are skipped (ISKIP 4) from the
    start.  QDOT and QDSP of both sets against the flang-built SPGSOLV, per value; zero
    values bit for bit, sign included."""
    from golden_io import assert_rates_close, load_transport
    if not (refrun.available("shim") and refrun.available("asm")):
        pytest.skip("reference SPGSOLV and shim drivers not built")
    f0, z = load("assemble")
    tr = load_transport(z)
    rng = np.random.default_rng(4)
    act = f0.active()
    pick = rng.choice(act, size=2 * (len(act) // 20), replace=False)
    later, never = np.sort(pick[: len(pick) // 2]), np.sort(pick[len(pick) // 2:])
    f1 = f0.copy()
    f1.iskip[never] = 4
    isk2 = f1.iskip.copy()
    isk2[later] = 5
    ref = refrun.spgsolv(f1.copy(), tr, driver="asm", ncalls=3, iskip_after=isk2)
    print(f"stale cells: {len(later)} advanced once, {len(never)} never; reference sweeps", ref["isor1"],
          ref["isor2"], flush=True)
    got = refrun.spgsolv(f1.copy(), tr, driver="shim", ncalls=3, iskip_after=isk2)
    assert (got["isor1"], got["isor2"]) == (ref["isor1"], ref["isor2"])
    assert np.abs(ref["qdsp"][:, later]).max() > 0.0, "the kept coefficients must matter"
    assert_rates_close(got["qdot"], got["qdsp"][:, later], ref["qdot"], ref["qdsp"][:, later], later,
                       "cells advanced once, then skipped")
    for k in ("qdot", "qdsp"):
        a, b = got[k][..., never], ref[k][..., never]
        assert np.array_equal(a, b) and np.array_equal(np.signbit(a), np.signbit(b)), k
