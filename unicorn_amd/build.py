"""Build the gfx950 chemistry library in-tree (unicorn_amd/lib/libunicorn_chem.so).

Steps: generate the mechanism-specific device code (csrc/gen_chem.py ->
csrc/chem_gen.inc), then compile csrc/chem_kernels.hip with hipcc for gfx950.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed-FP32 VALU forms are disabled on purpose (see the kernel sources' notes)
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd, **kw):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, **kw)


def build_chem(verbose=False) -> Path:
    sys.path.insert(0, str(CSRC))
    try:
        import gen_chem  # noqa: E402
        gen_chem.main()
    finally:
        sys.path.pop(0)
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / "libunicorn_chem.so"
    tmp = LIBDIR / "libunicorn_chem.so.tmp"
    run([HIPCC, *HIP_FLAGS, CSRC / "chem_kernels.hip", "-o", tmp], cwd=CSRC)
    tmp.replace(out)
    return out


def build_fshim(li: int = 711, lj: int = 131) -> Path:
    """The Fortran drop-in (ckrates_/solvesp_/solveh_ over the COMMON blocks) for a
    Fortran build with PARAMETER (LI=li, LJ=lj); its COMMON symbols bind to the program."""
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / "libunicorn_fshim.so"
    tmp = LIBDIR / "libunicorn_fshim.so.tmp"
    run([HIPCC, *HIP_FLAGS, f"-DUC_LI={li}", f"-DUC_LJ={lj}", CSRC / "fortran_shim.hip", "-o", tmp], cwd=CSRC)
    tmp.replace(out)
    return out


def build_surrogate() -> Path:
    """The surrogate's kernels (MLP forward/backward/Adam on BF16 MFMA, features, gate)."""
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / "libunicorn_surrogate.so"
    tmp = LIBDIR / "libunicorn_surrogate.so.tmp"
    run([HIPCC, *HIP_FLAGS, CSRC / "surrogate_kernels.hip", "-o", tmp], cwd=CSRC)
    tmp.replace(out)
    return out


def main():
    build_chem()
    build_fshim()
    build_surrogate()


if __name__ == "__main__":
    main()
