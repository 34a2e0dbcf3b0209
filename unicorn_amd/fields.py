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

from .scales import FUEL_INDEX, LSP, WSP, Scales, density, nasa7, sensible_enthalpy

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


@dataclass
class Coupling:
    """What SPGSOLV hands the chemistry beyond the state (spgsolv.f:98-1290), in the
    field layout: the hybrid-scheme coefficients of the species (E, W, P, Q of species
    1..51 in the order of /DUMMY/ and /DUM0/, solvesp.f:24-58), of the soot scalars
    (STE..STS) and of the enthalpy (ZE..ZS); the inputs of RDHSP (solveh.f:52-53: EMU,
    VDF, TMU, EKT, SIGSP); the geometry of the RHSZ flux (solveh.f:113-115: XXC, XXS,
    YYC, YYS, Y, ISYM); and the velocities of T(h) (solveh.f:1176-1177).  P couples
    J+1, Q J-1, E I+1, W I-1."""
    coef: np.ndarray     # (204, n)
    stc: np.ndarray      # (4, n)  STE, STW, STN, STS
    zc: np.ndarray       # (4, n)  ZE, ZW, ZN, ZS
    ekt: np.ndarray      # (n,)
    tmu: np.ndarray      # (n,)
    vdf: np.ndarray      # (52, n)
    ygeo: np.ndarray     # (n,)  radius Y(J,I)
    xxc: np.ndarray      # (li,)
    xxs: np.ndarray      # (li,)
    yyc: np.ndarray      # (lj,)
    yys: np.ndarray      # (lj,)
    u: np.ndarray        # (n,)
    v: np.ndarray        # (n,)
    w: np.ndarray        # (n,)
    emu: np.ndarray | None = None   # (n,); None: REI*BETA3*RHO as SPGSOLV sets it (spgsolv.f:175)
    sigsp: float = float(np.float32(0.9))   # main:346
    isym: int = 1

    def zero_like(self) -> "Coupling":
        """The same geometry and velocities with every coefficient and flux input zero."""
        z = np.zeros_like
        return Coupling(z(self.coef), z(self.stc), z(self.zc), z(self.ekt), z(self.tmu), z(self.vdf),
                        self.ygeo, self.xxc, self.xxs, self.yyc, self.yys, self.u, self.v, self.w,
                        None, self.sigsp, self.isym)


def synthetic_coupling(f: ChemField, seed: int = 0, strength: float = 0.1, flux: float = 1e-3,
                       velocity: float = 0.3, scales: Scales | None = None) -> Coupling:
    """Seeded coupling inputs for a field, shaped like a diffusion-dominated hybrid
    scheme: every face gets a conductance g >= 0 (uniform in [0.5, 1.5], times a
    per-species factor in [0.5, 2]) and the two cells sharing it see -strength*rho*g
    (E of cell i equals W of cell i+1), so each cell's equation stays diagonally
    dominant (PPP = RHO - (E+W+P+Q) > RHO).  Soot faces are ten times weaker.  VDF,
    TMU and EKT (positive, seeded) are scaled together so that the median |DT/2 x flux
    divergence| is `flux` times the median |RHO*HT|.  Uniform spacing on the unit
    square, radius Y = J*dy, ISYM = 1; velocities uniform in [-velocity, velocity]."""
    sc = scales or Scales()
    rng = np.random.Generator(np.random.PCG64(seed))
    li, lj, n = f.li, f.lj, f.n
    rho = f.rho.reshape(li, lj)

    def faces(scale):
        gx = rng.uniform(0.5, 1.5, (li + 1, lj)) * scale   # face between i-1 and i at index i
        gy = rng.uniform(0.5, 1.5, (li, lj + 1)) * scale   # face between j-1 and j at index j
        e = -strength * rho * gx[1:, :]
        w = -strength * rho * gx[:-1, :]
        p = -strength * rho * gy[:, 1:]
        q = -strength * rho * gy[:, :-1]
        return [a.ravel() for a in (e, w, p, q)]

    coef = np.empty((204, n))
    for m in range(51):
        coef[4 * m:4 * m + 4] = faces(rng.uniform(0.5, 2.0))
    stc = np.array(faces(0.1))
    zc = np.array(faces(1.0))
    dx, dy = 1.0 / (li - 1), 1.0 / (lj - 1)
    xxc = np.full(li, dx); xxs = np.full(li, dx)
    yyc = np.full(lj, dy); yys = np.full(lj, dy)
    ygeo = np.tile(np.arange(lj) * dy, li)
    u = rng.uniform(-velocity, velocity, n)
    v = rng.uniform(-velocity, velocity, n)
    w = np.zeros(n)
    vdf = rng.uniform(0.5, 2.0, (LSP, n))
    tmu = rng.uniform(0.0, 1.0, n)
    ekt = rng.uniform(0.0, 1.0, n)
    cp = Coupling(coef, stc, zc, ekt, tmu, vdf, ygeo, xxc, xxs, yyc, yys, u, v, w)
    # calibrate the RDHSP inputs: the flux is linear in them (numpy, approximate)
    pol = nasa7()
    T = f.tk * sc.tstr
    hi = T > 1000.0
    hssp = np.empty((LSP, n))
    for k in range(LSP):
        a = np.where(hi[None, :], pol[k, 7:12, None], pol[k, 0:5, None])
        hssp[k] = (a[0] * T + a[1] * T ** 2 / 2 + a[2] * T ** 3 / 3 + a[3] * T ** 4 / 4
                   + a[4] * T ** 5 / 5) / sc.wm[k] * sc.gasc / sc.wmstr / sc.hstr
    emu = sc.rei * sc.beta3 * f.rho
    a0 = emu[None, :] * vdf + (sc.rei * tmu / cp.sigsp - 0.1 * ekt)[None, :]
    rd = (a0 * hssp).reshape(LSP, li, lj)
    y = f.y.reshape(LSP, li, lj)
    div = (((rd[:, 2:, 1:-1] + rd[:, 1:-1, 1:-1]) * (y[:, 2:, 1:-1] - y[:, 1:-1, 1:-1])
            - (rd[:, 1:-1, 1:-1] + rd[:, :-2, 1:-1]) * (y[:, 1:-1, 1:-1] - y[:, :-2, 1:-1])) / dx / dx
           + ((rd[:, 1:-1, 2:] + rd[:, 1:-1, 1:-1]) * (y[:, 1:-1, 2:] - y[:, 1:-1, 1:-1])
              - (rd[:, 1:-1, 1:-1] + rd[:, 1:-1, :-2]) * (y[:, 1:-1, 1:-1] - y[:, 1:-1, :-2])) / dy / dy)
    x = np.abs(f.dt * 0.5 * div.sum(axis=0))
    target = flux * np.median(np.abs(f.rho * f.ht))
    med = float(np.median(x))
    s = target / med if med > 0 else 1.0
    cp.vdf = vdf * s
    cp.tmu = tmu * s
    cp.ekt = 0.1 * ekt * s
    return cp


@dataclass
class Transport:
    """SPGSOLV's inputs to its coefficient assembly beyond the chemistry state
    (spgsolv.f:98-349), in the field layout: thermal conductivity VTC, turbulent
    viscosity TMU, species diffusivities VDF (51 used), soot diffusivity STDF, the face
    velocities U (U(J,I) on the face between I-1 and I) and V (between J-1 and J), and
    the geometry (radius Y, XXC/XXS, YYC/YYS, ISYM); SIGSP and SIGH as SPGSOLV receives
    them (main:346-347, both 0.9)."""
    vtc: np.ndarray      # (n,)
    tmu: np.ndarray      # (n,)
    vdf: np.ndarray      # (52, n)
    stdf: np.ndarray     # (n,)
    u: np.ndarray        # (n,)
    v: np.ndarray        # (n,)
    ygeo: np.ndarray     # (n,)
    xxc: np.ndarray      # (li,)
    xxs: np.ndarray      # (li,)
    yyc: np.ndarray      # (lj,)
    yys: np.ndarray      # (lj,)
    sigsp: float = float(np.float32(0.9))
    sigh: float = float(np.float32(0.9))
    isym: int = 1

    def coupling(self, coef, stc, zc, ekt, w=None) -> Coupling:
        """The coupling inputs of the chemistry step from assembled coefficients (numpy
        arrays): SOLVESP's and SOLVEH's view of what SPGSOLV has set."""
        n = self.tmu.size
        return Coupling(np.asarray(coef), np.asarray(stc), np.asarray(zc), np.asarray(ekt), self.tmu, self.vdf,
                        self.ygeo, self.xxc, self.xxs, self.yyc, self.yys, self.u, self.v,
                        np.zeros(n) if w is None else w, sigsp=self.sigsp, isym=self.isym)


def synthetic_transport(f: ChemField, seed: int = 0, strength: float = 0.02, velocity: float = 0.3,
                        scales: Scales | None = None) -> Transport:
    """Seeded transport inputs for a field: uniform spacing on the unit square, radius
    Y = J*dy, ISYM = 1; VDF uniform in [0.5, 2], TMU in [0, 1] and STDF in [0, 0.1]
    times scales chosen so that a face's diffusive coefficient DT/DX * conductance is
    about `strength` times RHO, VTC so that EKT is of the same order; velocities uniform
    in [-velocity, velocity]."""
    sc = scales or Scales()
    rng = np.random.Generator(np.random.PCG64(seed))
    li, lj, n = f.li, f.lj, f.n
    dx, dy = 1.0 / (li - 1), 1.0 / (lj - 1)
    rho = float(np.median(f.rho))
    dscale = strength * min(dx, dy) ** 2 / (f.dt * sc.rei * sc.beta3)
    vdf = rng.uniform(0.5, 2.0, (LSP, n)) * dscale
    tmu = rng.uniform(0.0, 1.0, n) * dscale * sc.beta3 * rho
    stdf = rng.uniform(0.0, 0.1, n) * dscale
    # EKT = REI (BETA2 VTC / CP + TMU / SIGH) of about the species' conductance scale
    pol = nasa7()
    T = f.tk * sc.tstr
    hi = T > 1000.0
    cp = np.zeros(n)
    for k in range(LSP):
        a = np.where(hi[None, :], pol[k, 7:12, None], pol[k, 0:5, None])
        cp += (a[0] + a[1] * T + a[2] * T ** 2 + a[3] * T ** 3 + a[4] * T ** 4) / sc.wm[k] * f.y[k]
    cp = cp / sc.wmstr * sc.gasc * sc.tstr / sc.hstr
    ekt = rng.uniform(0.5, 1.5, n) * dscale * sc.beta3 * rho
    vtc = np.maximum(ekt - tmu / float(np.float32(0.9)), 0.0) * cp / sc.beta2
    return Transport(vtc, tmu, vdf, stdf, rng.uniform(-velocity, velocity, n), rng.uniform(-velocity, velocity, n),
                     np.tile(np.arange(lj) * dy, li), np.full(li, dx), np.full(li, dx), np.full(lj, dy),
                     np.full(lj, dy))
