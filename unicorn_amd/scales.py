"""Reference scales, nondimensional groups and mechanism constants of the
UNICORN chemistry step, restated from the program's set-up code.

Follows src/unicornd-mild-sd-5.f (``main``):
  main:331-353   physical constants (GASC=8.3144 and friends are default REAL,
                 i.e. float32 literals widened to REAL*8)
  main:421-453   freestream properties and the /CB06/ scales
  main:461-464   WM(k) = W_k/1000/WMSTR, HFO(k) = HCONV*a6_low(k)/WM(k)
  docs/input.uni:4-6,25-27  the sample case's reference values and solver controls

The mechanism table (``mech/sd52.json``) carries the NASA-7 polynomials exactly
as the reference's BLOCK DATA holds them.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from functools import lru_cache
from pathlib import Path

import numpy as np

LSP = 52
LRX = 544


def f32(x: float) -> float:
    """A default-REAL Fortran literal widened to REAL*8."""
    return float(np.float32(x))


# main:76-86, g/mol
WSP = np.array([
    2.016, 32.00, 1.008, 16.00, 17.01, 18.02, 33.01, 34.01, 28.01, 44.01,
    29.02, 30.03, 16.04, 15.04, 14.03, 14.03, 28.05, 31.03, 29.06, 30.07,
    13.02, 26.04, 27.05, 43.05, 44.05, 42.04, 41.03, 25.03, 31.03, 32.04,
    46.07, 44.05, 45.06, 45.06, 43.05, 45.06, 40.07, 39.06, 41.07, 42.08,
    44.10, 43.09, 43.09, 54.09, 29.02, 68.12, 100.2, 56.11, 70.14, 39.95,
    4.003, 28.01], dtype=np.float64)

GASC = f32(8.3144)     # main:333
RUCAL = f32(1.9872)    # main:335
GFORCE = f32(9.80665)  # main:332
SIGSP = f32(0.9)       # main:346

# fuel option -> species index (main:307-317), 1-based species numbers
FUEL_INDEX = {"H2": 1, "CH4": 13, "CH3OH": 30, "C2H2": 22, "C2H4": 17, "C2H6": 20,
              "C3H8": 41, "C3H6": 40, "CH2O": 12, "CO": 9, "C7H16": 47}


@lru_cache(maxsize=None)
def mechanism() -> dict:
    return json.loads((Path(__file__).parent / "mech" / "sd52.json").read_text())


def nasa7() -> np.ndarray:
    """POLSP(14,52) as a (52,14) array: [0:7] T<=TPOL1 fit, [7:14] T>TPOL1 fit."""
    m = mechanism()
    return np.array([[float.fromhex(v) for v in s["nasa7"]] for s in m["species"]])


@dataclass
class Scales:
    """/CB06/, /CB02/ and /HFORM/ contents for a given reference-values card."""
    vref: float = 10.0
    tref: float = 294.0
    pref: float = 1.0133e5
    rtin: float = 0.0
    rtot: float = 0.100
    fo2in: float = 0.233
    igrav: int = 1
    d: dict = field(default_factory=dict, repr=False)

    def __post_init__(self):
        vref, tref, pref = self.vref, self.tref, self.pref
        wmref = 1.0 / (self.fo2in / WSP[1] + (1.0 - self.fo2in) / WSP[LSP - 1]) / 1000.0  # main:421
        amref = 5.69428e-07 * tref ** f32(0.62245)                                       # main:422
        acref = 2.78763e-04 * tref ** f32(0.796783)                                      # main:423
        dref = 1.2e-04
        rref = pref / (GASC / wmref) / tref                                              # main:427
        d = self.d
        d["alstr"] = self.rtot - self.rtin
        d["vstr"] = vref
        d["rstr"] = rref
        d["pstr"] = rref * vref * vref
        d["tstr"] = tref
        d["hstr"] = vref * vref
        d["acstr"] = acref
        d["tmstr"] = d["alstr"] / vref
        d["amstr"] = amref
        d["dstr"] = dref
        d["wmstr"] = wmref
        d["rei"] = amref / rref / vref / d["alstr"]
        d["beta1"] = GASC * d["rstr"] * d["tstr"] / d["pstr"] / d["wmstr"]
        d["beta2"] = d["tstr"] * d["acstr"] / d["amstr"] / d["vstr"] / d["vstr"]
        d["beta3"] = d["rstr"] * d["dstr"] / d["amstr"]
        d["pback"] = pref / d["pstr"]
        d["hconv"] = GASC / d["wmstr"] / d["hstr"]
        d["gasc"] = GASC
        self.wm = WSP / 1000.0 / wmref                                                   # main:462
        pol = nasa7()
        self.hfo = d["hconv"] * pol[:, 5] / self.wm                                      # main:463

    def __getattr__(self, k):
        d = self.__dict__.get("d")
        if d is not None and k in d:
            return d[k]
        raise AttributeError(k)

    def dt_nondim(self, dt_seconds: float) -> float:
        """Physical time step -> the reference's nondimensional DT (time / TMSTR)."""
        return dt_seconds / self.tmstr

    def as_array(self) -> np.ndarray:
        """Packed constant block handed to the C-ABI (order fixed by include/unicorn_chem.h)."""
        d = self.d
        return np.array([d["alstr"], d["vstr"], d["rstr"], d["tstr"], d["hstr"], d["wmstr"],
                         d["gasc"], d["rei"], d["beta1"], d["beta3"], d["pback"], d["hconv"]],
                        dtype=np.float64)


def sensible_enthalpy(T: np.ndarray, Y: np.ndarray, sc: Scales) -> np.ndarray:
    """Nondimensional HT from T (K) and Y (52, n), as src/lib/readff.f:353-373 forms it
    (with U=V=W=0).  Used to build synthetic states."""
    pol = nasa7()
    T = np.asarray(T, dtype=np.float64)
    hi = T > 1000.0
    T2, T3, T4, T5 = T * T, T * T * T, T ** 4, T ** 5
    h = np.zeros_like(T)
    for k in range(LSP):
        a = np.where(hi[None, :], pol[k, 7:14, None], pol[k, 0:7, None])
        hk = (a[0] * T + a[1] * T2 / 2.0 + a[2] * T3 / 3.0 + a[3] * T4 / 4.0 + a[4] * T5 / 5.0 + a[5])
        h += Y[k] * (hk * sc.gasc / sc.wmstr / sc.hstr / sc.wm[k] - sc.hfo[k])
    return h


def density(T: np.ndarray, Y: np.ndarray, sc: Scales) -> np.ndarray:
    """Nondimensional rho = PBACK/(BETA1*TK*sum(Y/WM)) (main:1056-1061)."""
    rbar = np.zeros(np.shape(T))
    for k in range(LSP):
        rbar += Y[k] / sc.wm[k]
    return sc.pback / (sc.beta1 * (np.asarray(T) / sc.tstr) * rbar)
