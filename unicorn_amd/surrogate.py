"""Learned surrogate of the chemistry step, with a gate and FP64 fallback.

A 54 -> 4 x 256 (GELU) -> 53 MLP predicts one chemistry step from (T, ln dt,
Box-Cox(Y_k)).  Its outputs are dT and, per species, the change of Box-Cox(Y_k)
(lambda = 0.1), each standardised.  Weights and activations are BF16, accumulation
FP32.  Forward, backward and Adam run in this project's own kernels:
- ``csrc/mlp_kernels.hip``: MFMA 32x32x16 BF16 GEMMs;
- ``csrc/surrogate_kernels.hip``: features, training pairs, the fused inference
  kernel and the gate.
torch only allocates device memory and shuffles indices.

The training pairs come from this project's FP64 step itself (``ChemStep``), run
on seeded fields.  At inference the predicted change is corrected over species
1..49, with the least abundance-weighted norm, so that it conserves C, H and O.  A
per-cell gate then accepts it only when:
- every input lies in the training range (and STMF, STND in theirs), every output
  in the range of the training targets, and the new T in SOLVEH's clamp range;
- the applied change, after clipping and the N2 closure, conserves C, H and O to
  ``tol_el``;
- |sum Y + STMF - 1| <= ``tol_sum``;
- the chemistry is slow: |dT| <= ``tol_dt`` (K), every |dY_k| <= ``tol_dy`` and every
  predicted change of Box-Cox(Y_k) <= ``tol_bc`` (relative: it flags the radical
  build-up before an ignition), so that the network's relative error stays a small
  absolute one (fast cells go to FP64).
Rejected cells are advanced by the FP64 step (every accepted cell masked out) with the
caller's controls: the global stop over the rejected cells, normalised by the whole grid
as the reference normalises it, lands close to the full field's own stop because the
stiff cells the gate rejects carry most of the field's residual.
Accepted cells keep their soot scalars, which the network does not predict.  No
counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .scales import LSP, Scales, WSP, mechanism

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libunicorn_surrogate.so"
NF, NT, NPAD, HID = 54, 53, 64, 256
# (in, out) of the five layers, padded; used: 54 -> 256 -> 256 -> 256 -> 256 -> 53
LAYERS = [(NPAD, HID), (HID, HID), (HID, HID), (HID, HID), (HID, NPAD)]
HULL_MARGIN = 1e-3
MAJORS = ("H2", "O2", "H2O", "CO", "CO2", "CH4", "C2H4", "C7H16", "N2")

_lib = None


class SurrogateError(RuntimeError):
    pass


class Norm(C.Structure):
    _fields_ = [("xmean", C.c_float * NF), ("xsd", C.c_float * NF), ("tmean", C.c_float * NT),
                ("tsd", C.c_float * NT), ("tscale", C.c_float * NT)]


class Thermo(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("tstr", "hstr", "gasc", "wmstr", "pback", "beta1")] + [
        ("rwm", C.c_double * LSP), ("hfo", C.c_double * LSP), ("wm", C.c_double * LSP),
        ("nasah", (C.c_double * 14) * LSP),
        ("mel", (C.c_double * LSP) * 3), ("xlo", C.c_double * NF),
        ("xhi", C.c_double * NF), ("tlo", C.c_double * NT), ("thi", C.c_double * NT), ("tmin", C.c_double),
        ("tmax", C.c_double), ("stmf_max", C.c_double), ("stnd_max", C.c_double)]


def library() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (one HIP runtime)
    if not LIB_PATH.exists():
        raise SurrogateError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(LIB_PATH))
    lib.um_last_error.restype = C.c_char_p
    v, i, i64, d, f = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_float
    sig = {
        "um_gemm_nt": [v, i, v, i, v, i, i, i, i, i, v, v],
        "um_bias_act": [v, i, v, i, i, i, v, i, v, i, v, i, v],
        "um_act_backward": [v, i, v, i, v, i, i, v, i, v, i, v],
        "um_mse": [v, i, v, i, i, i, i, v, i, v, i, v, v],
        "um_rowsum_bf16": [v, i, i, i, v, v],
        "um_adam": [v, v, v, v, i64, f, f, f, f, i, v],
        "um_pack_w": [v, i, i, i, v, i, v, i, v],
        "um_cast_bf16": [v, i64, v, v],
        "um_pack_infer": [v, i, i, i, v, v],
        "us_collect": [v, v, v, v, i64, v, C.c_int32, d, d, v, v, i64, v],
        "us_colstats": [v, i64, i, i, v, i64, v, v],
        "us_batch": [v, v, v, i, v, v, v, v, v],
        "us_features": [v, v, i64, v, C.c_int32, C.c_int32, d, d, v, v, v],
        "us_gate": [v, v, v, v, v, v, v, v, i64, v, C.c_int32, d, d, d, d, d, d, v, v, v],
        "us_infer": [v, v, v, v, v, v, v, v, v, v, v, v, v, i64, i64, v, C.c_int32, d, d, d, d, d, d, d, v, v],
        "us_density": [v, v, v, v, i64, i64, v],
        "us_sizes": [v],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    sizes = (C.c_int64 * 2)()
    lib.us_sizes(sizes)
    if sizes[0] != C.sizeof(Norm) or sizes[1] != C.sizeof(Thermo):
        raise SurrogateError(f"parameter block sizes differ: {list(sizes)} vs {C.sizeof(Norm)}, {C.sizeof(Thermo)}")
    _lib = lib
    return lib


def _ok(rc: int):
    if rc < 0:
        raise SurrogateError(f"surrogate kernel error {rc}: {library().um_last_error().decode()}")
    return rc


def _p(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# ------------------------------------------------------------------ the network
class MLP:
    """The surrogate network on the current GPU: FP32 master weights, BF16 copies in both
    orientations, Adam moments, and the activation buffers of a batch."""

    def __init__(self, seed: int = 0, device: str = "cuda"):
        import torch
        self.lib = library()
        g = np.random.default_rng(seed)
        self.W, self.b = [], []
        for li, (fin, fout) in enumerate(LAYERS):
            used_in = NF if li == 0 else HID
            used_out = NT if li == len(LAYERS) - 1 else HID
            w = np.zeros((fout, fin), np.float32)
            w[:used_out, :used_in] = g.normal(0.0, math.sqrt(2.0 / used_in), (used_out, used_in))  # He init
            self.W.append(torch.from_numpy(w).to(device))
            self.b.append(torch.zeros(fout, dtype=torch.float32, device=device))
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.Wb = [torch.empty(w.shape, dtype=torch.int16, device=device) for w in self.W]
        self.WT = [torch.empty((w.shape[1], w.shape[0]), dtype=torch.int16, device=device) for w in self.W]
        self.t = 0
        self.device = device
        self.pack()
        self._bufs = {}

    def pack(self):
        for w, wb, wt in zip(self.W, self.Wb, self.WT):
            o, i = w.shape
            _ok(self.lib.um_pack_w(_p(w), o, i, i, _p(wb), i, _p(wt), o, _stream()))

    def pack_infer(self) -> tuple:
        """bf16 weights in the fused inference kernel's order (the input layer as is, the
        others with their columns in result-fragment order) and pointer arrays of them
        and of the fp32 biases, for us_infer."""
        import torch
        wp = []
        for li, w in enumerate(self.W):
            o, i = w.shape
            t = torch.empty((o, i), dtype=torch.int16, device=self.device)
            _ok(self.lib.um_pack_infer(_p(w), o, i, 0 if li == 0 else 1, _p(t), _stream()))
            wp.append(t)
        wptr = (C.c_void_p * 5)(*[t.data_ptr() for t in wp])
        bptr = (C.c_void_p * 5)(*[b.data_ptr() for b in self.b])
        return wp, wptr, bptr

    def buffers(self, B: int, train: bool):
        import torch
        key = (B, train)
        if key not in self._bufs:
            d = self.device
            bf = lambda *s: torch.empty(s, dtype=torch.int16, device=d)   # noqa: E731
            f32 = lambda *s: torch.empty(s, dtype=torch.float32, device=d)  # noqa: E731
            z = {"Z": [f32(B, fo) for _, fo in LAYERS], "H": [bf(B, HID) for _ in range(4)],
                 "out": f32(B, NPAD)}
            if train:
                z.update({"HT": [bf(HID, B) for _ in range(4)], "dH": f32(B, HID), "dZ": bf(B, HID),
                          "dZT": bf(HID, B), "dO": bf(B, NPAD), "dOT": bf(NPAD, B),
                          "gW": [torch.zeros_like(w) for w in self.W], "gb": [torch.zeros_like(b) for b in self.b],
                          "part": f32(32 * HID * HID), "loss": f32((B * NPAD + 255) // 256)})
            self._bufs[key] = z
        return self._bufs[key]

    def forward(self, X, B: int, train: bool = False, XT=None):
        """X: bf16 [B, 64] (int16 tensor).  Returns the fp32 output buffer [B, 64]."""
        buf = self.buffers(B, train)
        s = _stream()
        h = X
        for li, (fin, fout) in enumerate(LAYERS):
            Z = buf["Z"][li]
            _ok(self.lib.um_gemm_nt(_p(h), fin, _p(self.Wb[li]), fin, _p(Z), fout, B, fout, fin, 1, None, s))
            if li < 4:
                H = buf["H"][li]
                HT = buf["HT"][li] if train else None
                _ok(self.lib.um_bias_act(_p(Z), fout, _p(self.b[li]), B, fout, 1, _p(H), fout,
                                         _p(HT) if HT is not None else None, B, None, 0, s))
                h = H
            else:
                _ok(self.lib.um_bias_act(_p(Z), fout, _p(self.b[li]), B, fout, 0, None, 0, None, 0,
                                         _p(buf["out"]), NPAD, s))
        return buf["out"]

    def backward(self, X, XT, target, B: int) -> "torch.Tensor":
        """After forward(train=True): gradients of the MSE over the 53 used outputs into
        gW/gb; returns the per-block loss partials (sum of squared errors)."""
        buf = self.buffers(B, True)
        s = _stream()
        nl = _ok(self.lib.um_mse(_p(buf["out"]), NPAD, _p(target), NPAD, B, NPAD, NT, _p(buf["dO"]), NPAD,
                                 _p(buf["dOT"]), B, _p(buf["loss"]), s))
        splits = max(1, min(32, B // 256))
        dZ, dZT, width = buf["dO"], buf["dOT"], NPAD
        for li in range(len(LAYERS) - 1, -1, -1):
            fin, fout = LAYERS[li]
            inT = XT if li == 0 else buf["HT"][li - 1]
            # dW = dZ^T (fout x B) . in^T (fin x B)^T ; db = row sums of dZ^T
            _ok(self.lib.um_gemm_nt(_p(dZT), B, _p(inT), B, _p(buf["gW"][li]), fin, fout, fin, B, splits,
                                    _p(buf["part"]), s))
            _ok(self.lib.um_rowsum_bf16(_p(dZT), B, fout, B, _p(buf["gb"][li]), s))
            if li == 0:
                break
            # dH = dZ (B x fout) . W (fout x fin) = dZ . (W^T)^T with W^T [fin][fout]
            _ok(self.lib.um_gemm_nt(_p(dZ), width, _p(self.WT[li]), fout, _p(buf["dH"]), HID, B, fin, fout, 1,
                                    None, s))
            _ok(self.lib.um_act_backward(_p(buf["dH"]), HID, _p(buf["Z"][li - 1]), HID, _p(self.b[li - 1]), B, HID,
                                         _p(buf["dZ"]), HID, _p(buf["dZT"]), B, s))
            dZ, dZT, width = buf["dZ"], buf["dZT"], HID
        return buf["loss"][:nl]

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        buf = self.buffers(next(iter(k for k in self._bufs if k[1]))[0], True)
        self.t += 1
        s = _stream()
        for li in range(len(LAYERS)):
            _ok(self.lib.um_adam(_p(self.W[li]), _p(buf["gW"][li]), _p(self.mW[li]), _p(self.vW[li]),
                                 self.W[li].numel(), lr, b1, b2, eps, self.t, s))
            _ok(self.lib.um_adam(_p(self.b[li]), _p(buf["gb"][li]), _p(self.mb[li]), _p(self.vb[li]),
                                 self.b[li].numel(), lr, b1, b2, eps, self.t, s))
        self.pack()

    def state(self) -> dict:
        return {f"W{i}": w.cpu().numpy() for i, w in enumerate(self.W)} | {
            f"b{i}": b.cpu().numpy() for i, b in enumerate(self.b)}

    def load(self, st: dict):
        import torch
        for i in range(len(LAYERS)):
            self.W[i].copy_(torch.from_numpy(np.asarray(st[f"W{i}"], np.float32)))
            self.b[i].copy_(torch.from_numpy(np.asarray(st[f"b{i}"], np.float32)))
        self.pack()


# ------------------------------------------------------------------ data
@dataclass
class Dataset:
    xraw: "torch.Tensor"   # [P, 54] fp32: T (K), ln dt (s), Y
    traw: "torch.Tensor"   # [P, 53] fp32: dT (K), change of Box-Cox(Y)
    rows: int
    stmf_max: float = 0.0  # largest soot mass fraction / number density among the states
    stnd_max: float = 0.0


def training_fields(seed: int) -> list:
    """Seeded states as the 1001 x 401 C7H16 field and the batched grids are built:
    C7H16 fields with pre-ignition bands and CH4 / C2H4 / C7H16 flames, peak T in
    [1600, 2400] K (with the 300 K cold ends, T spans 300..2400 K).  dt is log-uniform
    in [1e-6, 1e-5] s, and every fourth field runs at exactly 1e-5 s, the production
    step, so that its ln dt lies inside the training range rather than on its edge."""
    from .fields import flame_field
    rng = np.random.default_rng(seed)
    out = []
    for g in range(24):
        fuel = "C7H16" if g % 3 == 0 else ("CH4", "C2H4", "C7H16")[int(rng.integers(3))]
        li = int(rng.integers(121, 302))
        lj = int(rng.integers(61, 152))
        dt = float(np.exp(rng.uniform(np.log(1e-6), np.log(1e-5))))
        if g % 4 == 0:
            dt = 1e-5
        band = float(rng.uniform(0.05, 0.3)) if fuel == "C7H16" else 0.0
        out.append(flame_field(li, lj, fuel, seed=10_000 * seed + g, dt_seconds=dt, preignition_band=band,
                               t_peak=float(rng.uniform(1600.0, 2400.0))))
    return out


def generate_pairs(npairs: int, seed: int = 7, steps_per_field: int = 200, rounds: int = 4,
                   early_steps: int = 10, early_weight: float = 5.0, log=print) -> Dataset:
    """(state, increment) pairs from chained FP64 steps of the training fields.  Each
    round advances 24 seeded fields by steps_per_field steps and, at every step, takes a
    seeded random subset of each field's active cells, sized so that ``rounds`` rounds
    fill npairs rows; the states therefore cover whole trajectories, from the seeded
    state's first relaxation to late relaxation.  The first ``early_steps`` steps, where
    the seeded minor species relax by orders of magnitude, are sampled
    ``early_weight`` times as densely."""
    import torch

    from .chem import ChemStep, Controls, DeviceField
    lib = library()
    sc = Scales()
    xraw = torch.empty((npairs, NF), dtype=torch.float32, device="cuda")
    traw = torch.empty((npairs, NT), dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(seed)
    weights = [early_weight if s < early_steps else 1.0 for s in range(steps_per_field)]
    rows, rnd = 0, 0
    smax = torch.zeros(2, dtype=torch.float64, device="cuda")
    while rows < npairs:
        fields = training_fields(seed + rnd)
        devs = [DeviceField(f) for f in fields]
        ctls = [Controls(dt=f.dt) for f in fields]
        ctx = ChemStep(sum(f.n for f in fields), max_grids=len(fields))
        cells = [torch.from_numpy(f.active().astype(np.int32)).cuda() for f in fields]
        nint = sum(int(c.numel()) for c in cells)
        frac = npairs / (rounds * sum(weights) * nint)
        logdt = [math.log(f.dt * sc.tmstr) for f in fields]
        for step in range(steps_per_field):
            before = [(d.y.clone(), d.tk.clone()) for d in devs]
            for d, cl in zip(devs, cells):
                smax[0] = torch.maximum(smax[0], d.stmf[cl.long()].max())
                smax[1] = torch.maximum(smax[1], d.stnd[cl.long()].max())
            ctx.step_batch(devs, ctls)
            for d, (y0, t0), cl, ld in zip(devs, before, cells, logdt):
                want = int(math.ceil(min(1.0, frac * weights[step]) * cl.numel()))
                take = min(max(1, want), npairs - rows)
                if take <= 0:
                    break
                sel = cl[torch.randperm(cl.numel(), device="cuda", generator=gen)[:take]].contiguous()
                _ok(lib.us_collect(_p(y0), _p(t0), _p(d.y), _p(d.tk), d.y.shape[1], _p(sel), take, sc.tstr, ld,
                                   _p(xraw), _p(traw), rows, _stream()))
                rows += take
            if rows >= npairs:
                break
        log(f"pairs: {rows} after round {rnd} ({step + 1} steps, {frac:.3f} of {nint} cells per step)")
        ctx.close()
        rnd += 1
    torch.cuda.synchronize()
    sm = smax.cpu().numpy()
    return Dataset(xraw, traw, rows, float(sm[0]), float(sm[1]))


def target_scales(ds: Dataset, sample: int = 1 << 18, seed: int = 0, quantile: float = 0.5) -> np.ndarray:
    """Per-target compression scale: the given quantile of |change| over a seeded sample
    of rows (the median by default; floored at 1e-6 of the largest |change| in the
    sample, and at 1e-30).  A lower quantile resolves small changes logarithmically."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    idx = torch.randint(0, ds.rows, (min(sample, ds.rows),), device="cuda", generator=gen)
    a = ds.traw[idx].abs().double().cpu().numpy()
    return np.maximum(np.maximum(np.quantile(a, quantile, axis=0), 1e-6 * a.max(axis=0)), 1e-30)


def column_stats(ds: Dataset, scale_quantile: float = 0.5) -> dict:
    """mean, sd, min, max of the transformed features and of the compressed targets
    asinh(d / scale), and the targets' scales."""
    import torch
    lib = library()
    out = {}
    scale = target_scales(ds, quantile=scale_quantile)
    sc_dev = torch.from_numpy(scale.astype(np.float32)).cuda()
    for name, raw, ncol, kind in (("x", ds.xraw, NF, 0), ("t", ds.traw, NT, 1)):
        chunk = 1 << 16
        nchunk = (ds.rows + chunk - 1) // chunk
        part = torch.empty(nchunk * ncol * 4, dtype=torch.float64, device="cuda")
        _ok(lib.us_colstats(_p(raw), ds.rows, ncol, kind, _p(sc_dev) if kind else None, chunk, _p(part), _stream()))
        p = part.view(nchunk, ncol, 4).cpu().numpy()
        s1, s2 = p[:, :, 0].sum(0), p[:, :, 1].sum(0)
        mean = s1 / ds.rows
        sd = np.sqrt(np.maximum(s2 / ds.rows - mean * mean, 0.0))
        sd = np.where(sd > 1e-12 * np.maximum(1.0, np.abs(mean)), sd, 1.0)
        out[name] = dict(mean=mean, sd=sd, lo=p[:, :, 2].min(0), hi=p[:, :, 3].max(0))
    out["t"]["scale"] = scale.astype(np.float32).astype(np.float64)
    return out


def norm_block(stats: dict) -> Norm:
    nm = Norm()
    for i in range(NF):
        nm.xmean[i], nm.xsd[i] = stats["x"]["mean"][i], stats["x"]["sd"][i]
    for i in range(NT):
        nm.tmean[i], nm.tsd[i] = stats["t"]["mean"][i], stats["t"]["sd"][i]
        nm.tscale[i] = stats["t"]["scale"][i]
    return nm


def _to_device(struct) -> "torch.Tensor":
    import torch
    raw = np.frombuffer(bytes(struct), dtype=np.uint8).copy()
    return torch.from_numpy(raw).cuda()


def train(ds: Dataset, stats: dict, epochs: int = 30, batch: int = 8192, lr: float = 1e-3, seed: int = 0,
          val_frac: float = 0.05, log=print) -> tuple[MLP, dict]:
    """Adam with a cosine learning rate, MSE on the standardised outputs; the last
    val_frac of a seeded permutation is held out."""
    import torch
    lib = library()
    net = MLP(seed)
    nm_dev = _to_device(norm_block(stats))
    gen = torch.Generator(device="cuda").manual_seed(seed)
    perm = torch.randperm(ds.rows, device="cuda", generator=gen)
    nval = int(ds.rows * val_frac)
    tr_idx, va_idx = perm[: ds.rows - nval], perm[ds.rows - nval:]
    ntr = tr_idx.numel() // batch * batch
    X = torch.empty((batch, NPAD), dtype=torch.int16, device="cuda")
    XT = torch.empty((NPAD, batch), dtype=torch.int16, device="cuda")
    T = torch.empty((batch, NPAD), dtype=torch.float32, device="cuda")
    steps_total = epochs * (ntr // batch)
    hist = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = tr_idx[torch.randperm(tr_idx.numel(), device="cuda", generator=gen)][:ntr].contiguous()
        loss_sum = torch.zeros((), dtype=torch.float64, device="cuda")
        for bi in range(ntr // batch):
            idx = order[bi * batch:(bi + 1) * batch]
            _ok(lib.us_batch(_p(ds.xraw), _p(ds.traw), _p(idx), batch, _p(nm_dev), _p(X), _p(XT), _p(T), _stream()))
            net.forward(X, batch, train=True)
            lp = net.backward(X, XT, T, batch)
            loss_sum += lp.double().sum()
            lr_t = 0.5 * lr * (1.0 + math.cos(math.pi * step / steps_total))
            net.adam(lr_t)
            step += 1
        tr_mse = float(loss_sum) / (ntr * NT)
        va_mse = evaluate(net, ds, va_idx, nm_dev, batch)
        hist.append((ep, tr_mse, va_mse))
        log(f"epoch {ep}: train mse {tr_mse:.4e}  val mse {va_mse:.4e}  ({time.perf_counter() - t0:.1f} s)")
    return net, {"history": hist, "val_rows": nval, "train_rows": ntr, "seconds": time.perf_counter() - t0}


def evaluate(net: MLP, ds: Dataset, idx, nm_dev, batch: int = 8192) -> float:
    """mean squared error (standardised outputs) over rows idx."""
    import torch
    lib = library()
    X = torch.empty((batch, NPAD), dtype=torch.int16, device="cuda")
    XT = torch.empty((NPAD, batch), dtype=torch.int16, device="cuda")
    T = torch.empty((batch, NPAD), dtype=torch.float32, device="cuda")
    total, n = 0.0, 0
    for bi in range(idx.numel() // batch):
        sel = idx[bi * batch:(bi + 1) * batch].contiguous()
        _ok(lib.us_batch(_p(ds.xraw), _p(ds.traw), _p(sel), batch, _p(nm_dev), _p(X), _p(XT), _p(T), _stream()))
        out = net.forward(X, batch, train=False)
        total += float(((out[:, :NT] - T[:, :NT]) ** 2).double().sum())
        n += batch * NT
    return total / max(n, 1)


# ------------------------------------------------------------------ inference
def element_matrix() -> np.ndarray:
    """Mass of C, H and O per unit mass of each species (formulas of the mechanism)."""
    mech = mechanism()
    aw = {"C": 12.011, "H": 1.008, "O": 16.00}
    m = np.zeros((3, LSP))
    for k, s in enumerate(mech["species"]):
        for e, n in s["formula"].items():
            if e in aw:
                m["CHO".index(e), k] = n * aw[e] / WSP[k]
    return m


def thermo_block(stats: dict, sc: Scales | None = None) -> Thermo:
    from .scales import nasa7
    sc = sc or Scales()
    th = Thermo()
    th.tstr, th.hstr, th.gasc, th.wmstr, th.pback, th.beta1 = (sc.tstr, sc.hstr, sc.gasc, sc.wmstr, sc.pback,
                                                               sc.beta1)
    pol = nasa7()
    mel = element_matrix()
    for k in range(LSP):
        th.rwm[k] = 1.0 / sc.wm[k]
        th.wm[k] = sc.wm[k]
        th.hfo[k] = sc.hfo[k]
        for r in (0, 7):  # per range: a_i / i for i = 1..5, a_6, a_7 (unused)
            for i in range(7):
                th.nasah[k][r + i] = pol[k, r + i] / (i + 1) if i < 5 else pol[k, r + i]
        for e in range(3):
            th.mel[e][k] = mel[e, k]
    # the hull, widened by HULL_MARGIN of each feature's range (the stored bounds are the
    # float32 features' extremes; a state exactly on the edge must pass)
    for i in range(NF):
        lo, hi = float(stats["x"]["lo"][i]), float(stats["x"]["hi"][i])
        m = HULL_MARGIN * (hi - lo) + 1e-6 * max(1.0, abs(lo), abs(hi))
        th.xlo[i], th.xhi[i] = lo - m, hi + m
    for i in range(NT):
        lo, hi = float(stats["t"]["lo"][i]), float(stats["t"]["hi"][i])
        m = HULL_MARGIN * (hi - lo)
        th.tlo[i], th.thi[i] = lo - m, hi + m
    th.tmin, th.tmax = sc.tstr - 1.0, 5000.0  # solveh.f:38,1229-1230
    th.stmf_max = float(stats.get("stmf_max", 0.0)) * (1.0 + HULL_MARGIN)
    th.stnd_max = float(stats.get("stnd_max", 0.0)) * (1.0 + HULL_MARGIN)
    return th


class SurrogateStep:
    """Gated surrogate step with FP64 fallback on one device-resident field.

    fused=True runs the features and the whole network in one kernel (activations in
    registers, weights shared through LDS), then the gate (us_infer); fused=False runs the features, each layer's GEMM and activation as
    separate kernels (us_features, the training forward pass), then the same gate
    (us_gate); the tests use it to cross-check the fused network."""

    def __init__(self, net: MLP, stats: dict, field, tol_el: float = 1e-6, tol_sum: float = 1e-6,
                 tol_dt: float = 5.0, tol_dy: float = 2e-3, tol_bc: float = 0.03, fused: bool = True):
        import torch

        from .chem import ChemStep
        self.lib = library()
        self.net = net
        self.sc = Scales()
        self.nm = _to_device(norm_block(stats))
        self.th = _to_device(thermo_block(stats, self.sc))
        self.tol_el, self.tol_sum, self.tol_dt, self.tol_dy, self.tol_bc = tol_el, tol_sum, tol_dt, tol_dy, tol_bc
        self.fused = fused
        self.li, self.lj = field.li, field.lj
        self.cells = torch.from_numpy(field.active().astype(np.int32)).cuda()  # ISKIP < 3 (ckrates.f:93)
        self.ncells = int(self.cells.numel())
        self.bpad = (self.ncells + 63) // 64 * 64
        self.mask = torch.zeros(field.n, dtype=torch.int32, device="cuda")
        self.stats = torch.zeros(8, dtype=torch.int64, device="cuda")
        if fused:
            self.wp, self.wptr, self.bptr = net.pack_infer()
            ld = max(self.ncells, 1)
            self.out = torch.empty((NPAD, ld), dtype=torch.float32, device="cuda")  # outputs, species-major
            self.pbuf = torch.empty((LSP, ld), dtype=torch.float64, device="cuda")  # Y^0.1 of the features
        else:
            self.X = torch.zeros((self.bpad, NPAD), dtype=torch.int16, device="cuda")
            self.reject = torch.zeros(field.n, dtype=torch.int32, device="cuda")
        self.fp64 = ChemStep(field.n)
        self.fp64.set_clip(False)  # step() clips every cell before the prediction

    def predict(self, dev, ctl, out=None):
        """The network and the gate only (no fallback): accepted cells of dev get their new
        Y, T and HT (rho is left to the finish of the step), self.mask marks them (10)
        and keeps ISKIP elsewhere.  out: optional float32
        [ncells, 64] tensor for the standardised network outputs (fused path)."""
        s = _stream()
        log_dt = math.log(ctl.dt * self.sc.tmstr)
        self.stats.zero_()
        n = dev.y.shape[1]
        if self.fused:
            self.mask.copy_(dev.iskip)
            _ok(self.lib.us_infer(self.wptr, self.bptr, _p(self.nm), _p(self.th), _p(dev.y), _p(dev.tk), _p(dev.ht),
                                  _p(dev.stmf), _p(dev.stnd), _p(dev.iskip), _p(self.mask), _p(self.out),
                                  _p(self.pbuf), self.out.shape[1], n, _p(self.cells), self.ncells, self.sc.tstr,
                                  log_dt, self.tol_el, self.tol_sum, self.tol_dt, self.tol_dy, self.tol_bc,
                                  _p(self.stats), s))
            if out is not None:
                out.copy_(self.out[:, : self.ncells].t())
        else:
            _ok(self.lib.us_features(_p(dev.y), _p(dev.tk), n, _p(self.cells), self.ncells, self.bpad,
                                     self.sc.tstr, log_dt, _p(self.nm), _p(self.X), s))
            o = self.net.forward(self.X, self.bpad)
            if out is not None:
                out.copy_(o[: self.ncells])
            _ok(self.lib.us_gate(_p(o), _p(self.nm), _p(self.th), _p(dev.y), _p(dev.tk), _p(dev.ht),
                                 _p(dev.stmf), _p(dev.stnd), n, _p(self.cells), self.ncells, log_dt,
                                 self.tol_el, self.tol_sum, self.tol_dt, self.tol_dy, self.tol_bc, _p(self.reject),
                                 _p(self.stats), s))
            self.mask.copy_(dev.iskip)
            acc = self.cells[self.reject[self.cells.long()] == 0].long()
            self.mask[acc] = 10
        return self.stats

    def step(self, dev, ctl) -> dict:
        """Advance the device field by one step as the FP64 step would: main:821-828 on
        every cell, then the surrogate where the gate accepts and the FP64 step (without
        its own clip) on the rejected cells, every accepted cell masked out.  Returns
        counts and timings."""
        import torch

        from .chem import clip
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t2 = torch.cuda.Event(enable_timing=True)
        t0.record()
        clip(dev)
        self.predict(dev, ctl)
        t1.record()
        st = self.stats.cpu().numpy()
        nrej = int(st[1:5].sum())
        fp = None
        if nrej:
            saved = dev.iskip
            dev.iskip = self.mask
            self.fp64.cells_changed()
            try:
                fp = self.fp64.step(dev, ctl)
            finally:
                dev.iskip = saved
        else:  # no FP64 finish ran: the density of every cell and the edges
            _ok(self.lib.us_density(_p(self.th), _p(dev.y), _p(dev.tk), _p(dev.rho), self.li, self.lj, _stream()))
        t2.record()
        torch.cuda.synchronize()
        raw = st[5:8].astype(np.int64).view(np.float64)
        return {"accepted": int(st[0]), "rejected": nrej, "out_of_range": int(st[1]), "residual": int(st[2]),
                "sum": int(st[3]), "active": int(st[4]), "raw_element_residual_max": raw.tolist(),
                "ms_surrogate": t0.elapsed_time(t1), "ms_total": t0.elapsed_time(t2),
                "fp64_sweeps": None if fp is None else fp.isor1}


def save(path, net: MLP, stats: dict, meta: dict):
    d = {f"stat_{g}_{k}": np.asarray(v) for g in ("x", "t") for k, v in stats[g].items()}
    d.update(net.state())
    d["stmf_max"] = np.asarray(stats.get("stmf_max", 0.0))
    d["stnd_max"] = np.asarray(stats.get("stnd_max", 0.0))
    d["meta"] = np.frombuffer(repr(meta).encode(), dtype=np.uint8)
    np.savez(path, **d)


def load(path) -> tuple[MLP, dict]:
    z = np.load(path, allow_pickle=False)
    stats = {g: {k: z[f"stat_{g}_{k}"] for k in ("mean", "sd", "lo", "hi")} for g in ("x", "t")}
    stats["t"]["scale"] = z["stat_t_scale"]
    stats["stmf_max"] = float(z["stmf_max"])
    stats["stnd_max"] = float(z["stnd_max"])
    net = MLP(0)
    net.load({k: z[k] for k in z.files if k[0] in "Wb" and k[1:].isdigit()})
    return net, stats
