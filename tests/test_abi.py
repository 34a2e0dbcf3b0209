"""The C-ABI libraries: they load and export every function include/unicorn_chem.h
and include/unicorn_surrogate.h declare for them, and the chemistry library's
host-only entry points agree with the other restatements of the reference's
scales.  No GPU call is made here."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from unicorn_amd import chem
from unicorn_amd.scales import Scales

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "unicorn_chem.h"
SUR_HEADER = ROOT / "include" / "unicorn_surrogate.h"
SUR_LIB = ROOT / "unicorn_amd" / "lib" / "libunicorn_surrogate.so"

pytestmark = pytest.mark.skipif(not chem.LIB_PATH.exists(), reason="library not built")


def declared_functions(header=HEADER):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(\w+)\s*\(", text, flags=re.M)))


def exported(lib: Path):
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


def test_library_exports_header():
    names = declared_functions()
    core = [n for n in names if n.startswith("uc_")]
    assert set(core) == set(chem.EXPORTS), (core, chem.EXPORTS)
    missing = set(core) - exported(chem.LIB_PATH)
    assert not missing, missing


def test_surrogate_library_exports_header():
    if not SUR_LIB.exists():
        pytest.skip("surrogate library not built")
    names = declared_functions(SUR_HEADER)
    assert names and all(n.startswith(("us_", "um_")) for n in names), names
    missing = set(names) - exported(SUR_LIB)
    assert not missing, missing
    # and nothing exported that the header leaves out
    extra = {n for n in exported(SUR_LIB) if n.startswith(("us_", "um_"))} - set(names)
    assert not extra, extra


def test_fortran_entry_points_declared():
    names = declared_functions()
    assert {"ckrates_", "solvesp_", "solveh_"} <= set(names)


def test_abi_version_and_errors():
    lib = chem.library()
    assert lib.uc_abi_version() == 4
    with pytest.raises(chem.ChemError):
        chem.scales_from_card(vref=-1.0)


def test_scales_match_python_restatement():
    s = chem.scales_from_card()
    p = Scales()
    assert s.wmstr == p.wmstr and s.rstr == p.rstr and s.beta1 == p.beta1 and s.pback == p.pback
    assert np.array_equal(np.array(s.wm), p.wm)
    assert np.allclose(np.array(s.hfo), p.hfo, rtol=2e-16, atol=0)


def test_device_field_refuses_wrong_shapes_before_any_allocation():
    """A field whose arrays do not have the (li*lj,) / (52, li*lj) shapes the kernels
    index is refused on the host; float32 arrays are widened to float64."""
    from unicorn_amd.fields import flame_field
    f = flame_field(9, 7, "CH4", seed=1)
    ok = chem.checked_arrays(f)
    assert all(a.flags.c_contiguous for a in ok.values())
    assert ok["y"].dtype == np.float64 and ok["iskip"].dtype == np.int32
    g = f.copy()
    g.rho = g.rho.astype(np.float32)
    assert chem.checked_arrays(g)["rho"].dtype == np.float64
    for name, bad in (("y", f.y.T.copy()), ("y", f.y[:51]), ("rho", f.rho[:-1]),
                      ("tk", f.tk.reshape(f.li, f.lj)), ("iskip", f.iskip.astype(np.float64))):
        g = f.copy()
        setattr(g, name, bad)
        with pytest.raises(ValueError, match=name):
            chem.checked_arrays(g)
        with pytest.raises(ValueError, match=name):
            chem.DeviceField(g, device="cpu")   # refused before torch allocates anything
    g = f.copy()
    g.li = 2
    with pytest.raises(ValueError, match="at least 3"):
        chem.checked_arrays(g)
