"""Test infrastructure (oracle): NOT part of the shipped product.

Runs the reference program itself (PROGRAM UNICORN, src/unicornd-mild-sd-5.f) as
built by ``make -C oracle prog``:

* ``unicorn_ref``  all flang, as the reference ships it;
* ``unicorn_spy``  CKRATES/SOLVESP/SOLVEH from oracle/spy.c, which dumps the COMMON
  state the program holds at each call and then runs the flang-built routines;
* ``unicorn_shim`` the same program linked against the gfx950 drop-in
  (unicorn_amd/lib/libunicorn_fshim.so).

Each run happens in its own directory holding ``cfg/input.uni`` (the reference's
sample deck ``docs/input.uni``, copied by the Makefile to oracle/_ref/input.uni and
edited here: time steps, restart, fuel, physics switches) and, for a restart, the
``FLAME.DATA`` the deck names.  Only tests use this module.
"""
from __future__ import annotations

import os
import re
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = HERE / "_ref"
DECK = REF / "input.uni"
LI, LJ = 711, 131      # the program's PARAMETER grid (main:3)
LE = LI * LJ
LSP = 52


def program(name: str) -> Path:
    return REF / name


def available(*names: str) -> bool:
    return DECK.exists() and all(program(n).exists() for n in names)


def write_deck(rundir: Path, *, itend: int, istore: int = 1, istb: int = 0, iread: int = 0,
               infuel: int | None = None, ithrm: int | None = None, ichem: int | None = None) -> Path:
    """cfg/input.uni in rundir: the sample deck with ITEND/ISTORE/ISTB (docs/input.uni:22),
    IREAD (:2), INFUEL (:5) and ITHRM/ICHEM (:6) replaced.  The program stores its
    flame and detail data after steps ITR >= ISTB with MOD(ITR-ISTB, ISTORE) = 0
    (main:1228), ISTORE = 1 meaning the last step only (main:240)."""
    lines = DECK.read_text().splitlines()

    def fields(line):
        data, _, tail = line.partition("/")
        return [f.strip() for f in data.split(",")], tail

    f, t = fields(lines[1])
    f[1] = str(iread)
    lines[1] = ", ".join(f) + " /" + t
    if infuel is not None:
        f, t = fields(lines[4])
        f[-1] = "%02d" % infuel
        lines[4] = ",".join(f) + "/" + t
    f, t = fields(lines[5])
    if ithrm is not None:
        f[2] = str(ithrm)
    if ichem is not None:
        f[3] = str(ichem)
    lines[5] = ", ".join(f) + "/" + t
    f, t = fields(lines[21])
    f[0], f[3], f[4] = str(itend), str(istore), str(istb)
    lines[21] = ", ".join(f) + "/" + t
    (rundir / "cfg").mkdir(parents=True, exist_ok=True)
    p = rundir / "cfg" / "input.uni"
    p.write_text("\n".join(lines) + "\n")
    return p


def run(name: str, rundir: Path, env: dict | None = None, timeout: int = 900) -> str:
    """Run one of the programs in rundir; returns its standard output.  Raises with the
    output's tail if it fails."""
    e = dict(os.environ)
    e.update(env or {})
    p = subprocess.run([str(program(name))], cwd=rundir, env=e, capture_output=True, text=True,
                       timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f"{name} exited {p.returncode}:\n{p.stdout[-2000:]}\n{p.stderr[-2000:]}")
    return p.stdout


@dataclass
class StepLine:
    """One line of the program's per-step residual table (main:1215-1220)."""
    resid: float
    sweeps_uvw: tuple
    resdm: float
    resdh: float
    isor2: int
    resdc: float
    isor1: int


# sweep counts are I2 fields: a count above 99 prints as ** (read here as -1)
_STEP = re.compile(r"^\s*(\S+)\(\s*(\d+),\s*(\d+),\s*(\d+)\)\s+(\S+)\s+(\S+)\(\s*(\d+|\*\*)\)\s+(\S+)\("
                   r"\s*(\d+|\*\*)\)")


def _count(v: str) -> int:
    return -1 if v == "**" else int(v)


def step_lines(stdout: str) -> list:
    out = []
    for line in stdout.splitlines():
        m = _STEP.match(line)
        if m:
            g = m.groups()
            out.append(StepLine(float(g[0]), (int(g[1]), int(g[2]), int(g[3])), float(g[4]), float(g[5]),
                                _count(g[6]), float(g[7]), _count(g[8])))
    return out


class Dump:
    """The spy's files in one directory: const0_*, rates<n>_*, hin<n>_*, hout<n>_*,
    sp<n>_ctl, h<n>_ctl."""

    def __init__(self, d: Path):
        self.d = Path(d)

    def arr(self, tag: str, name: str, dtype=np.float64) -> np.ndarray:
        a = np.fromfile(self.d / f"{tag}_{name}.bin", dtype=dtype)
        if a.size == LSP * LE:
            a = a.reshape(LSP, LE)
        return a

    def has(self, tag: str) -> bool:
        return (self.d / f"{tag}_flags.bin").exists()

    def consts(self) -> dict:
        cb06 = np.fromfile(self.d / "const0_cb06.bin")
        cb02 = np.fromfile(self.d / "const0_cb02.bin")
        cb08 = np.fromfile(self.d / "const0_cb08.bin")
        raw = (self.d / "const0_reac.bin").read_bytes()
        i32 = np.frombuffer(raw[:8 + 4 * 544], dtype=np.int32)
        f = np.frombuffer(raw[8 + 4 * 544:], dtype=np.float64).reshape(10, 544)
        return {"cb06": cb06, "rref_vref_tref_pref": cb02[:4], "wm": cb02[4:],
                "hfo": np.fromfile(self.d / "const0_hfo.bin"), "tpol": cb08[:2],
                "polsp": cb08[2:2 + 14 * LSP].reshape(LSP, 14), "cisp": cb08[2 + 14 * LSP:].reshape(LSP, 12),
                "lrev": int(i32[0]), "nrlind": int(i32[1]), "itbef": i32[2:].copy(),
                "alf": f[0], "akf": f[1], "exaf": f[2], "allow": f[3], "aklow": f[4], "ealow": f[5],
                "troe": f[6:10], "reac_raw": raw}

    def ctl(self, tag: str) -> np.ndarray:
        return np.fromfile(self.d / f"{tag}_ctl.bin")


def flame_restart(rundir: Path, field, scales, infuel: int = 11, itr: int = 0):
    """Write FLAME.DATA (the deck's restart file, docs/input.uni:35) for a chemistry
    field on the program's grid, in PRNTDK's format."""
    from unicorn_amd import formats
    assert field.li == LI and field.lj == LJ, "the program's grid is 711 x 131"
    z = np.zeros(LE)
    fd = formats.from_chem_field(field, formats.scales_row(scales), x=z, y=z, itr=itr, infuel=infuel)
    formats.write_flame_data(rundir / "FLAME.DATA", fd)
    return fd
