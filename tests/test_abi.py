"""The C-ABI library: it loads, exports every function include/unicorn_chem.h
declares for it, and its host-only entry points agree with the other
restatements of the reference's scales.  No GPU call is made here."""
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

pytestmark = pytest.mark.skipif(not chem.LIB_PATH.exists(), reason="library not built")


def declared_functions():
    text = HEADER.read_text()
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


def test_fortran_entry_points_declared():
    names = declared_functions()
    assert {"ckrates_", "solvesp_", "solveh_"} <= set(names)


def test_abi_version_and_errors():
    lib = chem.library()
    assert lib.uc_abi_version() == 1
    with pytest.raises(chem.ChemError):
        chem.scales_from_card(vref=-1.0)


def test_scales_match_python_restatement():
    s = chem.scales_from_card()
    p = Scales()
    assert s.wmstr == p.wmstr and s.rstr == p.rstr and s.beta1 == p.beta1 and s.pback == p.pback
    assert np.array_equal(np.array(s.wm), p.wm)
    assert np.allclose(np.array(s.hfo), p.hfo, rtol=2e-16, atol=0)
