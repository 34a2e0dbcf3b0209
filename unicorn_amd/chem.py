"""Python host layer over the C ABI of the gfx950 chemistry step.

``libunicorn_chem.so`` (built by ``unicorn_amd.build``) exports the functions
declared in ``include/unicorn_chem.h``; this module binds them with ctypes and
uses torch only for device memory and streams.  There is no CPU fallback: if
the library or a GPU is missing, the calls raise.

Host-side mirror of the reference's chemistry entry point
(``SPGSOLV(ISOR1,RELXSP,TOLRSP,RESDC,ISOR2,RELXH,TOLRH,RESDH,SIGH,SIGSP)``,
src/lib/spgsolv.f:6-7): ``ChemStep.step`` takes the same controls (sweep limits,
relaxation factors, tolerances) and returns the same outputs (sweeps used,
RESDC, RESDH), advancing the state in place.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .fields import ChemField

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libunicorn_chem.so"
MODE_GLOBAL, MODE_PERCELL = 0, 2


class UcScales(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "alstr", "vstr", "rstr", "pstr", "tstr", "tmstr", "amstr", "acstr", "dstr", "hstr",
        "cpstr", "akstr", "epstr", "wmstr", "gasc", "beta1", "beta2", "beta3", "beta4", "rei",
        "pback")] + [("wm", C.c_double * 52), ("hfo", C.c_double * 52)]


class UcControls(C.Structure):
    _fields_ = [("dt", C.c_double), ("relxsp", C.c_double), ("tolrsp", C.c_double),
                ("relxh", C.c_double), ("tolrh", C.c_double), ("isor1", C.c_int32),
                ("isor2", C.c_int32), ("mode", C.c_int32), ("predict", C.c_int32)]


class UcStats(C.Structure):
    _fields_ = [("isor1", C.c_int32), ("isor2", C.c_int32), ("resdc", C.c_double),
                ("resdh", C.c_double), ("cell_sweeps", C.c_int64), ("redo_cells", C.c_int64),
                ("ms_rates", C.c_double), ("ms_sweeps", C.c_double), ("ms_sources", C.c_double),
                ("ms_total", C.c_double), ("sweep_launches", C.c_int32), ("first_segment", C.c_int32),
                ("t_out_of_limit", C.c_int32), ("newton_failed", C.c_int32), ("first_failed", C.c_int32)]


class UcTuning(C.Structure):
    _fields_ = [("bins", C.c_int32), ("lpt", C.c_int32), ("persistent", C.c_int32), ("block", C.c_int32),
                ("graph", C.c_int32), ("lanes", C.c_int32), ("rate_parts", C.c_int32)]


class UcField(C.Structure):
    _fields_ = [("li", C.c_int32), ("lj", C.c_int32)] + [
        (n, C.c_void_p) for n in ("rho", "y", "tk", "ht", "stmf", "stnd", "iskip", "qdot", "qdsp")]


class UcCoupling(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "coef", "stc", "zc", "emu", "ekt", "tmu", "vdf", "ygeo", "xxc", "xxs", "yyc", "yys", "u", "v", "w")] + [
        ("sigsp", C.c_double), ("isym", C.c_int32)]


class UcTransport(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "vtc", "tmu", "vdf", "stdf", "u", "v", "ygeo", "xxc", "xxs", "yyc", "yys")] + [
        ("sigsp", C.c_double), ("sigh", C.c_double), ("isym", C.c_int32)]


EXPORTS = ("uc_abi_version", "uc_last_error", "uc_scales_from_card", "uc_create", "uc_create_batch",
           "uc_destroy", "uc_step", "uc_step_host", "uc_step_batch", "uc_autotune", "uc_set_tuning",
           "uc_get_tuning", "uc_cells_changed", "uc_clip", "uc_set_clip", "uc_step_coupled", "uc_assemble",
           "uc_transport_props", "uc_transport_props_noalloc", "uc_th_flags", "uc_sweep_residuals",
           "uc_keep_rates")

_lib = None


class ChemError(RuntimeError):
    pass


def library() -> C.CDLL:
    """Load the in-tree library (after torch, so both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (binds libamdhip64.so.7 first)
    except ImportError:
        pass
    if not LIB_PATH.exists():
        raise ChemError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(LIB_PATH))
    lib.uc_last_error.restype = C.c_char_p
    for name in EXPORTS[2:]:
        getattr(lib, name).restype = C.c_int
    lib.uc_scales_from_card.argtypes = [C.c_double] * 6 + [C.c_int32, C.POINTER(UcScales)]
    lib.uc_create.argtypes = [C.POINTER(UcScales), C.c_int64, C.POINTER(C.c_void_p)]
    lib.uc_create_batch.argtypes = [C.POINTER(UcScales), C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]
    lib.uc_step_batch.argtypes = [C.c_void_p, C.POINTER(UcField), C.POINTER(UcControls), C.c_int32, C.c_void_p,
                                  C.POINTER(UcStats)]
    lib.uc_autotune.argtypes = [C.c_void_p, C.POINTER(UcField), C.POINTER(UcControls), C.c_int32, C.c_void_p,
                                C.POINTER(UcTuning)]
    lib.uc_set_tuning.argtypes = [C.c_void_p, C.POINTER(UcTuning)]
    lib.uc_get_tuning.argtypes = [C.c_void_p, C.POINTER(UcTuning)]
    lib.uc_th_flags.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    lib.uc_sweep_residuals.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
    lib.uc_cells_changed.argtypes = [C.c_void_p]
    lib.uc_clip.argtypes = [C.POINTER(UcField), C.c_void_p]
    lib.uc_set_clip.argtypes = [C.c_void_p, C.c_int32]
    lib.uc_keep_rates.argtypes = [C.c_void_p, C.c_int32]
    lib.uc_destroy.argtypes = [C.c_void_p]
    lib.uc_step.argtypes = [C.c_void_p, C.POINTER(UcField), C.POINTER(UcControls), C.c_void_p,
                            C.POINTER(UcStats)]
    lib.uc_step_host.argtypes = [C.c_void_p, C.POINTER(UcField), C.POINTER(UcControls),
                                 C.POINTER(UcStats)]
    lib.uc_step_coupled.argtypes = [C.c_void_p, C.POINTER(UcField), C.POINTER(UcControls), C.POINTER(UcCoupling),
                                    C.c_void_p, C.POINTER(UcStats)]
    lib.uc_assemble.argtypes = [C.c_void_p, C.POINTER(UcField), C.POINTER(UcControls), C.POINTER(UcTransport),
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.uc_transport_props.argtypes = [C.c_void_p, C.POINTER(UcField)] + [C.c_void_p] * 5
    lib.uc_transport_props_noalloc.argtypes = [C.POINTER(UcScales), C.POINTER(UcField)] + [C.c_void_p] * 5
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise ChemError(f"unicorn_chem error {rc}: {library().uc_last_error().decode()}")


def scales_from_card(vref=10.0, tref=294.0, pref=1.0133e5, rtin=0.0, rtot=0.100, fo2in=0.233,
                     igrav=1) -> UcScales:
    s = UcScales()
    _check(library().uc_scales_from_card(vref, tref, pref, rtin, rtot, fo2in, igrav, C.byref(s)))
    return s


@dataclass
class Controls:
    """SPGSOLV's solver controls (docs/input.uni:25-27 defaults)."""
    dt: float
    isor1: int = 100
    relxsp: float = 0.99
    tolrsp: float = 1e-8
    isor2: int = 100
    relxh: float = 0.99
    tolrh: float = 1e-8
    mode: int = MODE_GLOBAL

    def c(self) -> UcControls:
        return UcControls(self.dt, self.relxsp, self.tolrsp, self.relxh, self.tolrh,
                          self.isor1, self.isor2, self.mode, 0)


@dataclass
class StepStats:
    isor1: int
    isor2: int
    resdc: float
    resdh: float
    cell_sweeps: int
    redo_cells: int
    ms_rates: float
    ms_sweeps: float
    ms_sources: float
    ms_total: float
    sweep_launches: int
    first_segment: int
    t_out_of_limit: int     # cells above the high-range enthalpy fit at 5000 K (solveh.f:1200-1204)
    newton_failed: int      # cells whose T(h) Newton did not converge (solveh.f:1206-1219)
    first_failed: int       # the first of those in I-then-J order (flat index), or -1

    @classmethod
    def of(cls, s: UcStats) -> "StepStats":
        return cls(*(getattr(s, n) for n, _ in UcStats._fields_))


def checked_arrays(f: ChemField) -> dict:
    """The field's arrays as contiguous float64 (int32 for ISKIP) host arrays, after
    checking that every one has the shape the kernels index: (li*lj,) per variable and
    (52, li*lj) for the species.  Raises ValueError before anything is allocated on the
    device, since the kernels read li*lj (x 52) elements from each pointer."""
    if int(f.li) < 3 or int(f.lj) < 3:
        raise ValueError(f"field {f.li}x{f.lj}: LI and LJ must be at least 3")
    n = int(f.li) * int(f.lj)
    out = {}
    for name, shape in (("rho", (n,)), ("y", (52, n)), ("tk", (n,)), ("ht", (n,)),
                        ("stmf", (n,)), ("stnd", (n,)), ("iskip", (n,))):
        a = np.asarray(getattr(f, name))
        if a.shape != shape:
            raise ValueError(f"field {f.li}x{f.lj}: {name} has shape {a.shape}, expected {shape}")
        if name == "iskip":
            if not np.issubdtype(a.dtype, np.integer):
                raise ValueError(f"field: iskip has dtype {a.dtype}, expected an integer type")
            out[name] = np.ascontiguousarray(a, dtype=np.int32)
        else:
            if not (np.issubdtype(a.dtype, np.floating) or np.issubdtype(a.dtype, np.integer)):
                raise ValueError(f"field: {name} has dtype {a.dtype}, expected a real type")
            out[name] = np.ascontiguousarray(a, dtype=np.float64)
    return out


class DeviceField:
    """A ChemField resident in device memory (torch tensors on the current GPU)."""

    def __init__(self, f: ChemField, device="cuda"):
        host = checked_arrays(f)
        import torch
        t = lambda a: torch.from_numpy(a).to(device)  # noqa: E731
        self.li, self.lj, self.dt, self.fuel = f.li, f.lj, f.dt, f.fuel
        self.rho, self.y, self.tk, self.ht = t(host["rho"]), t(host["y"]), t(host["tk"]), t(host["ht"])
        self.stmf, self.stnd = t(host["stmf"]), t(host["stnd"])
        self.iskip = t(host["iskip"])
        self.qdot = torch.zeros(f.n, dtype=torch.float64, device=device)
        self.qdsp = torch.zeros((52, f.n), dtype=torch.float64, device=device)

    def c(self) -> UcField:
        return UcField(self.li, self.lj, self.rho.data_ptr(), self.y.data_ptr(), self.tk.data_ptr(),
                       self.ht.data_ptr(), self.stmf.data_ptr(), self.stnd.data_ptr(),
                       self.iskip.data_ptr(), self.qdot.data_ptr(), self.qdsp.data_ptr())

    def to_host(self) -> ChemField:
        g = lambda a: a.detach().cpu().numpy().copy()  # noqa: E731
        return ChemField(self.li, self.lj, g(self.rho), g(self.y), g(self.tk), g(self.ht), g(self.stmf),
                         g(self.stnd), g(self.iskip), self.dt, self.fuel)


class DeviceCoupling:
    """A unicorn_amd.fields.Coupling resident in device memory (torch tensors); arrays
    that are None stay null (zero)."""

    NAMES = ("coef", "stc", "zc", "emu", "ekt", "tmu", "vdf", "ygeo", "xxc", "xxs", "yyc", "yys", "u", "v", "w")

    def __init__(self, cp, device="cuda", keep=None):
        """keep: names to upload (default all that are set); the others stay null."""
        import torch
        self.sigsp, self.isym = cp.sigsp, cp.isym
        for n in self.NAMES:
            a = getattr(cp, n, None)
            if a is not None and (keep is None or n in keep):
                setattr(self, n, torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device))
            else:
                setattr(self, n, None)

    def c(self) -> UcCoupling:
        s = UcCoupling()
        for n in self.NAMES:
            t = getattr(self, n)
            setattr(s, n, t.data_ptr() if t is not None else None)
        s.sigsp, s.isym = self.sigsp, self.isym
        return s


class DeviceTransport:
    """A unicorn_amd.fields.Transport resident in device memory (torch tensors)."""

    NAMES = ("vtc", "tmu", "vdf", "stdf", "u", "v", "ygeo", "xxc", "xxs", "yyc", "yys")

    def __init__(self, tr, device="cuda"):
        import torch
        self.sigsp, self.sigh, self.isym = tr.sigsp, tr.sigh, tr.isym
        for n in self.NAMES:
            setattr(self, n, torch.from_numpy(np.ascontiguousarray(getattr(tr, n), dtype=np.float64)).to(device))

    def c(self) -> UcTransport:
        s = UcTransport()
        for n in self.NAMES:
            setattr(s, n, getattr(self, n).data_ptr())
        s.sigsp, s.sigh, s.isym = self.sigsp, self.sigh, self.isym
        return s


def clip(dev: "DeviceField"):
    """main:821-828 on a device field (Y<0 -> 0, N2 by closure, every cell), on the
    current stream: the first phase of a step, on its own."""
    import torch
    f = dev.c()
    _check(library().uc_clip(C.byref(f), C.c_void_p(torch.cuda.current_stream().cuda_stream)))


class ChemStep:
    """A chemistry-step context on the current GPU (scales + scratch)."""

    def __init__(self, max_cells: int, scales: UcScales | None = None, max_grids: int = 1):
        self.lib = library()
        self.scales = scales or scales_from_card()
        h = C.c_void_p()
        _check(self.lib.uc_create_batch(C.byref(self.scales), int(max_cells), int(max_grids), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.uc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, f: DeviceField, ctl: Controls, stream=None) -> StepStats:
        """Advance a device-resident field by one chemistry step."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        st = UcStats()
        fc = f.c()
        cc = ctl.c()
        _check(self.lib.uc_step(self.h, C.byref(fc), C.byref(cc), C.c_void_p(stream), C.byref(st)))
        return StepStats.of(st)

    def step_coupled(self, f: DeviceField, ctl: Controls, cp: "DeviceCoupling", stream=None) -> StepStats:
        """Advance a device-resident field by one chemistry step with the coupling inputs
        SPGSOLV assembles (transport coefficients, enthalpy flux inputs, velocities)."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        st = UcStats()
        fc, cc, pc = f.c(), ctl.c(), cp.c()
        _check(self.lib.uc_step_coupled(self.h, C.byref(fc), C.byref(cc), C.byref(pc), C.c_void_p(stream),
                                        C.byref(st)))
        return StepStats.of(st)

    def assemble(self, f: DeviceField, ctl: Controls, tr: "DeviceTransport", stream=None) -> "DeviceCoupling":
        """SPGSOLV's coefficient assembly on the device (uc_assemble): the coupling inputs of
        step_coupled (coefficients, EKT; TMU, VDF, geometry and velocities from tr), in
        new device arrays.  Cells the assembly does not write (boundary, ISKIP >= 10)
        hold zeros here; the reference leaves face conductances there (spgsolv.f:233-348),
        which no relaxed cell reads."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        n = f.rho.numel()
        cp = DeviceCoupling.__new__(DeviceCoupling)
        for name in DeviceCoupling.NAMES:
            setattr(cp, name, None)
        z = lambda *shape: torch.zeros(shape, dtype=torch.float64, device=f.rho.device)
        cp.coef, cp.stc, cp.zc, cp.ekt = z(204, n), z(4, n), z(4, n), z(n)
        for name in ("tmu", "vdf", "ygeo", "xxc", "xxs", "yyc", "yys", "u", "v"):
            setattr(cp, name, getattr(tr, name))
        cp.w = z(n)
        cp.sigsp, cp.isym = tr.sigsp, tr.isym
        fc, cc, tc = f.c(), ctl.c(), tr.c()
        _check(self.lib.uc_assemble(self.h, C.byref(fc), C.byref(cc), C.byref(tc), cp.coef.data_ptr(),
                                    cp.stc.data_ptr(), cp.zc.data_ptr(), cp.ekt.data_ptr(), C.c_void_p(stream)))
        return cp

    def transport_props(self, f: DeviceField, stream=None) -> dict:
        """The mixture-averaged transport properties of every cell (uc_transport_props,
        main:843-1005): new device tensors vmu (n,), vtc (n,), vdf (52, n), stdf (n,)."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        n = f.rho.numel()
        z = lambda *shape: torch.empty(shape, dtype=torch.float64, device=f.rho.device)  # noqa: E731
        out = {"vmu": z(n), "vtc": z(n), "vdf": z(52, n), "stdf": z(n)}
        fc = f.c()
        _check(self.lib.uc_transport_props(self.h, C.byref(fc), out["vmu"].data_ptr(), out["vtc"].data_ptr(),
                                           out["vdf"].data_ptr(), out["stdf"].data_ptr(), C.c_void_p(stream)))
        return out

    def step_batch(self, fields: list[DeviceField], ctls: list[Controls], stream=None) -> list[StepStats]:
        """Advance several device-resident grids by one chemistry step each, in one batch."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        n = len(fields)
        fa = (UcField * n)(*[f.c() for f in fields])
        ca = (UcControls * n)(*[c.c() for c in ctls])
        sa = (UcStats * n)()
        _check(self.lib.uc_step_batch(self.h, fa, ca, n, C.c_void_p(stream), sa))
        return [StepStats.of(x) for x in sa]

    def autotune(self, fields: list[DeviceField], ctls: list[Controls], stream=None) -> dict:
        """Pick the tuning for these grids (their state is left unchanged)."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        n = len(fields)
        fa = (UcField * n)(*[f.c() for f in fields])
        ca = (UcControls * n)(*[c.c() for c in ctls])
        t = UcTuning()
        _check(self.lib.uc_autotune(self.h, fa, ca, n, C.c_void_p(stream), C.byref(t)))
        return {n_: getattr(t, n_) for n_, _ in UcTuning._fields_}

    def set_tuning(self, bins=1, lpt=0, persistent=1, block=256, graph=0, lanes=1, rate_parts=0):
        t = UcTuning(bins, lpt, persistent, block, graph, lanes, rate_parts)
        _check(self.lib.uc_set_tuning(self.h, C.byref(t)))

    def cells_changed(self):
        """The fields' ISKIP arrays changed in place: rebuild the cell list next step."""
        _check(self.lib.uc_cells_changed(self.h))

    def set_clip(self, on: bool):
        """Whether this context's steps start with main:821-828 (default on)."""
        _check(self.lib.uc_set_clip(self.h, 1 if on else 0))

    def keep_rates(self, on: bool):
        """Keep each cell's RC of the last step that advanced it and form QDOT/QDSP of the
        cells with 3 <= ISKIP < 10 from it, as SOLVEH does from /DUMR/ (default off)."""
        _check(self.lib.uc_keep_rates(self.h, 1 if on else 0))

    def th_flags(self, dev, grid: int = 0):
        """The last step's per-cell T(h) flags of grid `grid` (dev: its DeviceField), a
        uint8 device tensor in the field layout: 1 above the high-range fit at 5000 K,
        2 no Newton root in 100 steps (solveh.f:1200-1227), 0 elsewhere."""
        import torch
        out = torch.empty(dev.rho.numel(), dtype=torch.uint8, device=dev.rho.device)
        _check(self.lib.uc_th_flags(self.h, grid, C.c_void_p(out.data_ptr()),
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out

    def sweep_residuals(self, grid: int = 0) -> np.ndarray:
        """RSD/SPNORM after each species sweep of the last step of grid `grid`, up to the
        decided count (the quantity solvesp.f:2834 tests against TOLRSP)."""
        n = self.lib.uc_sweep_residuals(self.h, grid, None, 0)
        if n < 0:
            _check(n)
        out = np.zeros(n)
        if n:
            self.lib.uc_sweep_residuals(self.h, grid, out.ctypes.data_as(C.c_void_p), n)
        return out

    def tuning(self) -> dict:
        t = UcTuning()
        _check(self.lib.uc_get_tuning(self.h, C.byref(t)))
        return {n_: getattr(t, n_) for n_, _ in UcTuning._fields_}

    def step_host(self, f: ChemField, ctl: Controls) -> tuple[StepStats, np.ndarray, np.ndarray]:
        """Advance a host field (copied to and from the GPU); returns stats, QDOT, QDSP."""
        h = checked_arrays(f)
        for name in ("rho", "y", "tk", "ht", "stmf", "stnd"):
            if h[name] is not getattr(f, name):
                setattr(f, name, h[name])   # the step writes the caller's field in place
        isk = h["iskip"]
        qdot = np.zeros(f.n)
        qdsp = np.zeros((52, f.n))
        p = lambda a: a.ctypes.data  # noqa: E731
        fc = UcField(f.li, f.lj, p(f.rho), p(f.y), p(f.tk), p(f.ht), p(f.stmf), p(f.stnd), p(isk),
                     p(qdot), p(qdsp))
        st = UcStats()
        cc = ctl.c()
        _check(self.lib.uc_step_host(self.h, C.byref(fc), C.byref(cc), C.byref(st)))
        return StepStats.of(st), qdot, qdsp
