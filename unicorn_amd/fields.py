"""Seeded synthetic reacting fields for the chemistry step.

A field is an LI x LJ axisymmetric grid (I axial, J radial), stored exactly as
the reference's COMMON arrays are: one array per variable, (LJ, LI) Fortran
order, i.e. J fastest (src/lib/ckrates.f:8-20).  Here that is a flat index
``c = i*lj + j`` and species arrays of shape (52, li*lj).

The recipe (mixture fraction, temperature profile, Burke-Schumann majors,
log-uniform minors, renormalisation, ideal-gas density, sensible enthalpy):

  Z(r, z)   = exp(-(r / r_j(z))^2),  r_j(z) = R0 + SLOPE*z,  times (1 + 0.02*U(-1,1)), clipped to [0,1]
  T(Z)      = piecewise linear: 300 K at Z=0 and Z=1, 2000 K at Z=0.055
  majors    = Burke-Schumann split of fuel/air (O2 0.233, N2 0.767 by mass) for the fuel
              C_xH_y: CO2, H2O products; 10 % of product carbon as CO, 10 % of product
              hydrogen as H2
  minors    = every other species log-uniform in [1e-8, 1e-4] x 10^(-30 theta),
              theta = clip((1500 K - T)/1200 K, 0, 1)
  then all Y_k are divided by their sum; rho = PBACK/(BETA1 T sum Y/W) (main:1056-1061);
  HT from T and Y as src/lib/readff.f:353-373 forms it.

The temperature factor on the minors keeps the state inside the range where the
reference step is well conditioned.  The mechanism's reverse rate coefficients
grow without bound as T falls (RC of slot 496, HO2+C2H3+C3H4 -> O2+C5H8, is
~1e18 at 1000 K, ~1e29 at 700 K and ~1e76 at 300 K), so a radical pool of
1e-8..1e-4 in a cold cell makes the reference's own sweeps diverge (residual
~1e113, half of a 61 x 41 field turned to noise after clipping and
renormalisation) and the result chaotic under 1-ulp input changes.  At
T >= 1500 K the minors keep the full [1e-8, 1e-4] range.

A pre-ignition band (a fraction of the axial rows centred on mid-length) holds
the unburnt mixing-line composition at T ~ U(700, 1100) K: these are the stiff
cells.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .scales import FUEL_INDEX, LSP, WSP, Scales, density, sensible_enthalpy

R_LEN = 0.02      # radial extent, m
Z_LEN = 0.10      # axial extent, m
R0 = 0.002        # jet half-width at z=0, m
SLOPE = 0.08      # jet spreading rate
Z_PEAK = 0.055
T_PEAK = 2000.0
T_COLD = 300.0
MINOR_T_FULL = 1500.0      # K: minors span the full [1e-8, 1e-4] at and above this
MINOR_COLD_DECADES = 30.0  # decades the minor range is lowered by at T_COLD
Y_O2_AIR = 0.233
Y_N2_AIR = 0.767

FUEL_CH = {"CH4": (1, 4), "C2H4": (2, 4), "C7H16": (7, 16), "C2H6": (2, 6), "C3H8": (3, 8),
           "C2H2": (2, 2), "C3H6": (3, 6), "H2": (0, 2)}

IDX = {"H2": 0, "O2": 1, "H2O": 5, "CO": 8, "CO2": 9, "N2": 51}


@dataclass
class ChemField:
    """State of one grid in the reference layout (flat c = i*lj + j)."""
    li: int
    lj: int
    rho: np.ndarray      # (n,)   nondimensional density
    y: np.ndarray        # (52,n) mass fractions
    tk: np.ndarray       # (n,)   T / TSTR
    ht: np.ndarray       # (n,)   nondimensional sensible enthalpy
    stmf: np.ndarray     # (n,)   soot mass fraction
    stnd: np.ndarray     # (n,)   soot number density
    iskip: np.ndarray    # (n,)   int32 mask (>=3 skipped by the chemistry)
    dt: float            # nondimensional time step
    fuel: str = "CH4"

    @property
    def n(self) -> int:
        return self.li * self.lj

    def interior(self) -> np.ndarray:
        """Flat indices of the interior cells (I=2..LI-1, J=2..LJ-1)."""
        i, j = np.meshgrid(np.arange(1, self.li - 1), np.arange(1, self.lj - 1), indexing="ij")
        return (i * self.lj + j).ravel()

    def active(self) -> np.ndarray:
        """Interior cells the chemistry advances (ISKIP < 3, ckrates.f:93)."""
        c = self.interior()
        return c[self.iskip[c] < 3]

    def copy(self) -> "ChemField":
        return ChemField(self.li, self.lj, self.rho.copy(), self.y.copy(), self.tk.copy(),
                         self.ht.copy(), self.stmf.copy(), self.stnd.copy(), self.iskip.copy(),
                         self.dt, self.fuel)

    def take(self, cells: np.ndarray, lj: int | None = None) -> "ChemField":
        """A new field made of the given cells, laid out as an (li', lj') grid
        whose interior holds them (its frame cells are copies, masked)."""
        cells = np.asarray(cells)
        lj2 = lj or min(129, max(1, len(cells)))
        li2 = -(-len(cells) // lj2)
        LI2, LJ2 = li2 + 2, lj2 + 2
        src = np.full(LI2 * LJ2, cells[0], dtype=np.int64)
        mask = np.full(LI2 * LJ2, 10, dtype=np.int32)
        for q, c in enumerate(cells):
            i, j = divmod(q, lj2)
            d = (i + 1) * LJ2 + (j + 1)
            src[d] = c
            mask[d] = self.iskip[c]
        f = ChemField(LI2, LJ2, self.rho[src].copy(), self.y[:, src].copy(), self.tk[src].copy(),
                      self.ht[src].copy(), self.stmf[src].copy(), self.stnd[src].copy(), mask,
                      self.dt, self.fuel)
        return f


def burke_schumann(Z: np.ndarray, fuel: str) -> np.ndarray:
    """Major-species mass fractions (52, n) for a fuel/air Burke-Schumann flame sheet."""
    x, yh = FUEL_CH[fuel]
    kf = FUEL_INDEX[fuel] - 1
    wf = WSP[kf]
    nu_o2 = x + yh / 4.0
    s = nu_o2 * WSP[IDX["O2"]] / wf
    zst = Y_O2_AIR / (s + Y_O2_AIR)
    Y = np.zeros((LSP, Z.size))
    lean = Z <= zst
    burnt = np.where(lean, Z, zst * (1.0 - Z) / (1.0 - zst))      # fuel mass burnt
    Y[kf] = np.where(lean, 0.0, (Z - zst) / (1.0 - zst))
    Y[IDX["O2"]] = np.where(lean, Y_O2_AIR * (1.0 - Z / zst), 0.0)
    Y[IDX["N2"]] = Y_N2_AIR * (1.0 - Z)
    y_co2 = burnt * x * WSP[IDX["CO2"]] / wf
    y_h2o = burnt * (yh / 2.0) * WSP[IDX["H2O"]] / wf
    Y[IDX["CO"]] = 0.1 * y_co2 * WSP[IDX["CO"]] / WSP[IDX["CO2"]]
    Y[IDX["CO2"]] = 0.9 * y_co2
    Y[IDX["H2"]] = 0.1 * y_h2o * WSP[IDX["H2"]] / WSP[IDX["H2O"]]
    Y[IDX["H2O"]] = 0.9 * y_h2o
    return Y


def mixing_line(Z: np.ndarray, fuel: str) -> np.ndarray:
    Y = np.zeros((LSP, Z.size))
    Y[FUEL_INDEX[fuel] - 1] = Z
    Y[IDX["O2"]] = Y_O2_AIR * (1.0 - Z)
    Y[IDX["N2"]] = Y_N2_AIR * (1.0 - Z)
    return Y


def flame_field(li: int, lj: int, fuel: str = "CH4", seed: int = 0, dt_seconds: float = 1e-5,
                preignition_band: float = 0.0, scales: Scales | None = None,
                t_peak: float = T_PEAK) -> ChemField:
    sc = scales or Scales()
    rng = np.random.Generator(np.random.PCG64(seed))
    z = np.linspace(0.0, Z_LEN, li)
    r = np.linspace(0.0, R_LEN, lj)
    rj = R0 + SLOPE * z
    Z = np.exp(-(r[None, :] / rj[:, None]) ** 2).ravel()          # (li*lj,) i-major, j fastest
    Z = np.clip(Z * (1.0 + 0.02 * rng.uniform(-1.0, 1.0, Z.size)), 0.0, 1.0)
    T = np.where(Z <= Z_PEAK, T_COLD + (t_peak - T_COLD) * Z / Z_PEAK,
                 t_peak + (T_COLD - t_peak) * (Z - Z_PEAK) / (1.0 - Z_PEAK))
    Y = burke_schumann(Z, fuel)
    if preignition_band > 0.0:
        rows = np.arange(li)
        half = 0.5 * preignition_band * li
        band_rows = np.abs(rows - 0.5 * (li - 1)) < half
        band = np.repeat(band_rows, lj)
        Yb = mixing_line(Z, fuel)
        Y[:, band] = Yb[:, band]
        T[band] = rng.uniform(700.0, 1100.0, int(band.sum()))
    majors = {FUEL_INDEX[fuel] - 1, IDX["O2"], IDX["N2"], IDX["CO2"], IDX["H2O"], IDX["CO"], IDX["H2"]}
    theta = np.clip((MINOR_T_FULL - T) / (MINOR_T_FULL - T_COLD), 0.0, 1.0)
    for k in range(LSP):
        if k not in majors:
            Y[k] = 10.0 ** (rng.uniform(-8.0, -4.0, Z.size) - MINOR_COLD_DECADES * theta)
    Y /= Y.sum(axis=0, keepdims=True)
    n = li * lj
    f = ChemField(li, lj, density(T, Y, sc), Y, T / sc.tstr, sensible_enthalpy(T, Y, sc),
                  np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32), sc.dt_nondim(dt_seconds),
                  fuel)
    return f


def config0(seed: int = 0) -> ChemField:
    """61 x 41 CH4 field at 1 atm, dt = 1e-5 s."""
    return flame_field(61, 41, "CH4", seed)


def config1(seed: int = 1) -> ChemField:
    """1001 x 401 C7H16 field, 10 % pre-ignition band at 700-1100 K, dt = 1e-5 s."""
    return flame_field(1001, 401, "C7H16", seed, preignition_band=0.10)


def batch_grids(ngrids: int = 256, seed: int = 2) -> list[ChemField]:
    """A batch of independent grids: LI log-uniform in [31, 301], LJ log-uniform in
    [21, 151] (about 2.1 M cells for 256 grids), fuel per grid drawn from CH4, C2H4
    and C7H16, dt per grid log-uniform in [1e-6, 1e-5] s; each state built as the
    61 x 41 field's (flame_field, its own seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for g in range(ngrids):
        li = int(round(np.exp(rng.uniform(np.log(31), np.log(301)))))
        lj = int(round(np.exp(rng.uniform(np.log(21), np.log(151)))))
        fuel = ("CH4", "C2H4", "C7H16")[int(rng.integers(3))]
        dt = float(np.exp(rng.uniform(np.log(1e-6), np.log(1e-5))))
        out.append(flame_field(li, lj, fuel, seed=1000 + g, dt_seconds=dt))
    return out
