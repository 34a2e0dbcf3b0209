"""The surrogate network: the CPU oracle pinned by finite differences, the GPU's own
BF16-MFMA forward/backward checked against it, Adam against its formula, and the
trained network (unicorn_amd/surrogate_weights.npz) against the method without it.
"""
from __future__ import annotations

import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import mlp_oracle as mo

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "unicorn_amd" / "lib" / "libunicorn_surrogate.so"
WEIGHTS = ROOT / "unicorn_amd" / "surrogate_weights.npz"


def small_net(rng, dims=(8, 16, 16, 8)):
    Ws = [rng.normal(0, 1 / np.sqrt(i), (o, i)) for i, o in zip(dims[:-1], dims[1:])]
    bs = [rng.normal(0, 0.1, o) for o in dims[1:]]
    return Ws, bs


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    Ws, bs = small_net(rng)
    X = rng.normal(size=(5, 8))
    T = rng.normal(size=(5, 8))
    loss, gW, gb = mo.loss_and_grads(Ws, bs, X, T, nused=7, round_bf16=False)
    eps = 1e-6
    for li in range(len(Ws)):
        for _ in range(6):
            i, j = rng.integers(Ws[li].shape[0]), rng.integers(Ws[li].shape[1])
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[li][i, j] += eps
            Wm[li][i, j] -= eps
            fd = (mo.loss_and_grads(Wp, bs, X, T, 7, False)[0] - mo.loss_and_grads(Wm, bs, X, T, 7, False)[0]) / (2 * eps)
            assert abs(fd - gW[li][i, j]) <= 1e-6 * max(1.0, abs(fd)), (li, i, j, fd, gW[li][i, j])
        k = rng.integers(bs[li].shape[0])
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[li][k] += eps
        bm[li][k] -= eps
        fd = (mo.loss_and_grads(Ws, bp, X, T, 7, False)[0] - mo.loss_and_grads(Ws, bm, X, T, 7, False)[0]) / (2 * eps)
        assert abs(fd - gb[li][k]) <= 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -3.0, 1e-40], np.float32)
    got = mo.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0); 1 + 3*2^-8 -> 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.015625 and got[3] == -3.0


def test_library_exports():
    if not LIB.exists():
        pytest.skip("libunicorn_surrogate.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], check=True, capture_output=True, text=True).stdout
    names = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for n in ("um_gemm_nt", "um_bias_act", "um_act_backward", "um_mse", "um_rowsum_bf16", "um_adam", "um_pack_w",
              "um_pack_infer", "um_cast_bf16", "us_collect", "us_colstats", "us_batch", "us_features", "us_gate",
              "us_infer", "us_density", "us_sizes"):
        assert n in names, n


def test_element_matrix_mass_closure():
    """The C, H and O mass per unit mass of each species, with N, Ar and He added, sums
    to 1: the mechanism's formulas and molecular weights agree (a wrong atom count or
    weight shows as a species off by far more than the weights' rounding)."""
    from unicorn_amd.scales import LSP, WSP, mechanism
    from unicorn_amd.surrogate import element_matrix
    m = element_matrix()
    aw = {"N": 14.007, "AR": 39.948, "HE": 4.0026}
    rest = np.zeros(LSP)
    for k, sp in enumerate(mechanism()["species"]):
        for e, cnt in sp["formula"].items():
            if e.upper() in aw:
                rest[k] += cnt * aw[e.upper()] / WSP[k]
    tot = m.sum(0) + rest
    assert np.abs(tot - 1.0).max() < 2e-3, np.abs(tot - 1.0).max()


def test_inference_weight_order_is_a_permutation():
    """k_pack_infer's column order (position 8g+j <- input 16(j/4)+4g+j%4 in each block of
    32) is the order in which a 16x16x32 MFMA result fragment holds the neurons: lane
    group g, register r of tile t carries neuron 16t+4g+r, and k-step s takes tile 2s
    (j = 0..3) and tile 2s+1 (j = 4..7)."""
    def src(pos):
        q = pos % 32
        g, j = q // 8, q % 8
        return pos - q + 16 * (j // 4) + 4 * g + j % 4
    assert sorted(src(p) for p in range(256)) == list(range(256))
    for s in range(8):
        for g in range(4):
            for j in range(8):
                t, r = 2 * s + j // 4, j % 4
                assert src(32 * s + 8 * g + j) == 16 * t + 4 * g + r


@pytest.fixture(scope="module")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.gpu
def test_gpu_forward_backward_match_oracle(cuda):
    torch = cuda
    from unicorn_amd.surrogate import LAYERS, MLP, NPAD, NT, _p, _stream, library
    lib = library()
    net = MLP(seed=3)
    rng = np.random.default_rng(4)
    for b in net.b:  # non-zero biases
        b.copy_(torch.from_numpy(rng.normal(0, 0.1, b.numel()).astype(np.float32)))
    B = 256
    Xf = np.zeros((B, NPAD), np.float32)
    Xf[:, :54] = rng.normal(size=(B, 54))
    Tf = np.zeros((B, NPAD), np.float32)
    Tf[:, :NT] = rng.normal(size=(B, NT))
    X = torch.empty((B, NPAD), dtype=torch.int16, device="cuda")
    XT = torch.empty((NPAD, B), dtype=torch.int16, device="cuda")
    xf = torch.from_numpy(Xf).cuda()
    xfT = xf.t().contiguous()
    lib.um_cast_bf16(_p(xf), B * NPAD, _p(X), _stream())
    lib.um_cast_bf16(_p(xfT), B * NPAD, _p(XT), _stream())
    T = torch.from_numpy(Tf).cuda()
    out = net.forward(X, B, train=True).cpu().numpy().astype(np.float64)
    lp = net.backward(X, XT, T, B)
    loss_gpu = float(lp.double().sum()) / (B * NT)
    buf = net.buffers(B, True)
    Ws = [w.cpu().numpy().astype(np.float64) for w in net.W]
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b]
    ref_out, _ = mo.forward(Ws, bs, Xf.astype(np.float64), True)
    loss, gW, gb = mo.loss_and_grads(Ws, bs, Xf.astype(np.float64), Tf.astype(np.float64), NT, True)
    # forward: FP32 accumulation vs exact, on BF16-rounded operands
    assert np.abs(out[:, :NT] - ref_out[:, :NT]).max() <= 2e-2 * np.abs(ref_out[:, :NT]).max()
    assert abs(loss_gpu - loss) <= 1e-3 * loss
    for li in range(len(LAYERS)):
        g = buf["gW"][li].cpu().numpy().astype(np.float64)
        rel = np.linalg.norm(g - gW[li]) / np.linalg.norm(gW[li])
        cos = float((g * gW[li]).sum() / (np.linalg.norm(g) * np.linalg.norm(gW[li])))
        assert rel <= 2e-2 and cos >= 0.9998, (li, rel, cos)
        gbl = buf["gb"][li].cpu().numpy().astype(np.float64)
        assert np.linalg.norm(gbl - gb[li]) <= 2e-2 * np.linalg.norm(gb[li]), li


@pytest.mark.gpu
def test_gpu_gelu_matches_exact(cuda):
    """The device GELU (erf by Abramowitz & Stegun 7.1.28, float arithmetic) against the
    exact GELU in float64, over [-12, 12] and at its special values."""
    torch = cuda
    from unicorn_amd.surrogate import _p, _stream, library
    x = np.r_[np.linspace(-12, 12, 64 * 1000), 0.0, -0.0, 1e-30, -1e-30].astype(np.float32)
    n = x.size // 64 * 64
    x = x[:n]
    z = torch.from_numpy(x).cuda().view(-1, 64).contiguous()
    b = torch.zeros(64, dtype=torch.float32, device="cuda")
    out = torch.empty_like(z)
    assert library().um_bias_act(_p(z), 64, _p(b), z.shape[0], 64, 1, None, 0, None, 0, _p(out), 64, _stream()) == 0
    got = out.cpu().numpy().ravel().astype(np.float64)
    ref = mo.gelu(x.astype(np.float64))
    assert np.abs(got - ref).max() <= 4e-6 * np.maximum(1.0, np.abs(ref)).max(), np.abs(got - ref).max()


@pytest.mark.gpu
def test_gpu_adam_matches_formula(cuda):
    torch = cuda
    import ctypes as C

    from unicorn_amd.surrogate import _p, _stream, library
    lib = library()
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=1000).astype(np.float32)
    g = rng.normal(size=1000).astype(np.float32)
    p = torch.from_numpy(p0.copy()).cuda()
    gt = torch.from_numpy(g).cuda()
    m = torch.zeros(1000, device="cuda")
    v = torch.zeros(1000, device="cuda")
    for t in (1, 2):
        lib.um_adam(_p(p), _p(gt), _p(m), _p(v), 1000, C.c_float(1e-3), C.c_float(0.9), C.c_float(0.999),
                    C.c_float(1e-8), t, _stream())
    # two steps of Adam with the same gradient, by the formula
    m1, v1 = 0.1 * g, 0.001 * g * g
    q = p0 - 1e-3 * (m1 / 0.1) / (np.sqrt(v1 / 0.001) + 1e-8)
    m2, v2 = 0.9 * m1 + 0.1 * g, 0.999 * v1 + 0.001 * g * g
    q = q - 1e-3 * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + 1e-8)
    assert np.allclose(p.cpu().numpy(), q, rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
def test_trained_network_beats_frozen_chemistry(cuda):
    """On a field none of its training states came from (201 x 101 C7H16, its own seed),
    past the seeded state's first relaxation (4 FP64 steps), one gated surrogate step
    (FP64 fallback included) against the FP64 step from the same state: its RMS error
    in T and in the major species is far below that of the method without it (frozen
    chemistry, dT = dY = 0), its largest errors stay within 5 K and 2e-3, and the gate
    accepts most cells."""
    torch = cuda
    if not WEIGHTS.exists():
        pytest.skip("no trained weights")
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import flame_field
    from unicorn_amd.scales import Scales, mechanism
    from unicorn_amd.surrogate import MAJORS, SurrogateStep, load
    net, stats = load(WEIGHTS)
    f = flame_field(201, 101, "C7H16", seed=424242, preignition_band=0.1)
    ctl = Controls(dt=f.dt)
    ctx = ChemStep(f.n)
    ref = DeviceField(f)
    for _ in range(4):
        ctx.step(ref, ctl)
    s0 = ref.to_host()
    dev = DeviceField(s0)
    ctx.step(ref, ctl)
    st = SurrogateStep(net, stats, s0).step(dev, ctl)
    c = f.active()
    names = [sp["name"] for sp in mechanism()["species"]]
    mj = [names.index(m) for m in MAJORS]
    r, d = ref.to_host(), dev.to_host()
    tstr = Scales().tstr
    eT, fT = (d.tk[c] - r.tk[c]) * tstr, (s0.tk[c] - r.tk[c]) * tstr
    eY, fY = d.y[mj][:, c] - r.y[mj][:, c], s0.y[mj][:, c] - r.y[mj][:, c]
    rms = lambda a: float(np.sqrt(np.mean(a * a)))  # noqa: E731
    print(st, "T rms", rms(eT), rms(fT), "max", np.abs(eT).max(), "Y rms", rms(eY), rms(fY), "max", np.abs(eY).max())
    assert st["accepted"] >= 0.4 * c.size
    assert rms(eT) <= 0.3 * rms(fT) and rms(eY) <= 0.3 * rms(fY)
    assert np.abs(eT).max() <= 5.0 and np.abs(eY).max() <= 2e-3


def _field_stats(f, t_sd=(5.0, 1e-10)):
    """Normalisation and hull from the field's own features (so every cell is inside),
    targets scaled so that the toy network's changes stay far below the smallest mass
    fractions (a clipped negative Y breaks C/H/O conservation and rejects the cell)."""
    from unicorn_amd.scales import Scales
    sc = Scales()
    c = f.active()
    X = np.vstack([f.tk[c] * sc.tstr, np.full(c.size, math.log(f.dt * sc.tmstr)),
                   (np.maximum(f.y[:, c], 0.0) ** 0.1 - 1.0) / 0.1])
    sd = X.std(1)
    sd[sd <= 1e-12] = 1.0
    t_scale = np.r_[t_sd[0], np.full(52, t_sd[1])]  # change = scale * sinh(output)
    return {"x": {"mean": X.mean(1), "sd": sd, "lo": X.min(1), "hi": X.max(1)},
            "t": {"mean": np.zeros(53), "sd": np.ones(53), "lo": np.full(53, -1e3), "hi": np.full(53, 1e3),
                  "scale": t_scale}, "stmf_max": 0.0, "stnd_max": 0.0}


NO_ACTIVITY = {"tol_dt": 1e9, "tol_dy": 1.0, "tol_bc": 1e9}  # toy networks: the slow-chemistry rule off


def _toy_net(torch, seed=3, out_scale=0.05):
    from unicorn_amd.surrogate import MLP
    net = MLP(seed=seed)
    rng = np.random.default_rng(seed + 1)
    for b in net.b:
        b.copy_(torch.from_numpy(rng.normal(0, 0.1, b.numel()).astype(np.float32)))
    net.W[-1].mul_(out_scale)  # small predicted changes: most cells pass the gate
    net.pack()
    return net


def _bf16_bits_to_float(a):
    return (a.astype(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.gpu
def test_gpu_fused_inference_matches_oracle_and_unfused(cuda):
    """The fused kernel's network outputs against the CPU oracle on the same bf16
    features, and its gate against the separate-kernel path."""
    torch = cuda
    from unicorn_amd.chem import Controls, DeviceField
    from unicorn_amd.fields import flame_field
    from unicorn_amd.surrogate import NT, SurrogateStep
    f = flame_field(61, 41, "C7H16", seed=5, preignition_band=0.1)
    stats = _field_stats(f)
    net = _toy_net(torch)
    m = f.active().size
    res = {}
    for fused in (True, False):
        dev = DeviceField(f)
        ss = SurrogateStep(net, stats, f, fused=fused, **NO_ACTIVITY)
        o = torch.zeros((m, 64), dtype=torch.float32, device="cuda")
        st = ss.predict(dev, Controls(dt=f.dt), out=o).cpu().numpy()
        res[fused] = (o.cpu().numpy().astype(np.float64), dev.to_host(), ss.mask.cpu().numpy(), st)
        if not fused:
            X = _bf16_bits_to_float(ss.X[:m].cpu().numpy())
    Ws = [w.cpu().numpy().astype(np.float64) for w in net.W]
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b]
    ref, _ = mo.forward(Ws, bs, X, True)
    scale = np.abs(ref[:, :NT]).max()
    of, ou = res[True][0][:, :NT], res[False][0][:, :NT]
    print("fused-oracle", np.abs(of - ref[:, :NT]).max() / scale, "fused-unfused", np.abs(of - ou).max() / scale)
    assert np.abs(of - ref[:, :NT]).max() <= 2e-2 * scale
    assert np.abs(of - ou).max() <= 1e-2 * scale
    # the gate: k_gate (the separate-kernel path) on the fused kernel's own outputs
    # makes the same decisions and writes the same Y, T and HT (one gate_cell in both)
    from unicorn_amd.scales import Scales
    from unicorn_amd.surrogate import _p, _stream, library
    c = f.active()
    ss = SurrogateStep(net, stats, f, fused=False, **NO_ACTIVITY)
    dev = DeviceField(f)
    out = torch.from_numpy(res[True][0].astype(np.float32)).cuda()
    st = torch.zeros(8, dtype=torch.int64, device="cuda")
    log_dt = math.log(f.dt * Scales().tmstr)
    assert library().us_gate(_p(out), _p(ss.nm), _p(ss.th), _p(dev.y), _p(dev.tk), _p(dev.ht), _p(dev.stmf),
                             _p(dev.stnd), f.n, _p(ss.cells), ss.ncells, log_dt, ss.tol_el, ss.tol_sum,
                             ss.tol_dt, ss.tol_dy, ss.tol_bc, _p(ss.reject), _p(st), _stream()) == 0
    hu, hf_ = dev.to_host(), res[True][1]
    rej_u = ss.reject.cpu().numpy()[c] != 0
    rej_f = res[True][2][c] != 10
    print("stats fused", res[True][3][:4], "k_gate on its outputs", st.cpu().numpy()[:4])
    assert np.array_equal(rej_u, rej_f)
    assert (~rej_f).sum() >= 0.5 * c.size
    for name in ("y", "tk", "ht"):
        a, b = getattr(hf_, name), getattr(hu, name)
        assert np.allclose(a, b, rtol=1e-13, atol=1e-300), (name, np.abs(a - b).max())


@pytest.mark.gpu
def test_gpu_fused_gate_invariants(cuda):
    """Accepted cells: Y >= 0, sum Y + STMF = 1, C/H/O conserved, and the HT written
    gives back the applied T through SOLVEH's Newton (the oracle's); rejected cells
    bitwise untouched with ISKIP kept; cells above a narrowed T hull rejected as out of
    range, and the counters add up."""
    torch = cuda
    from oracle import pyoracle as po
    from unicorn_amd.chem import Controls, DeviceField
    from unicorn_amd.fields import flame_field
    from unicorn_amd.scales import Scales
    from unicorn_amd.surrogate import SurrogateStep, element_matrix, thermo_block
    sc = Scales()
    f = flame_field(61, 41, "C7H16", seed=6, preignition_band=0.1)
    stats = _field_stats(f)
    c = f.active()
    T0 = f.tk[c] * sc.tstr
    stats["x"]["hi"][0] = float(np.median(T0))  # the hotter half is outside the hull
    xhi = thermo_block(stats, sc).xhi[0]
    net = _toy_net(torch, seed=8)
    dev = DeviceField(f)
    ss = SurrogateStep(net, stats, f, **NO_ACTIVITY)
    st = ss.predict(dev, Controls(dt=f.dt)).cpu().numpy()
    g = dev.to_host()
    mask = ss.mask.cpu().numpy()
    acc, rej = c[mask[c] == 10], c[mask[c] != 10]
    print(st[:5], acc.size, rej.size)
    assert st[0] == acc.size and st[:5].sum() == c.size
    assert st[1] == int((T0 > xhi).sum())
    assert acc.size > 0.2 * c.size
    assert (mask[rej] == f.iskip[rej]).all()
    for name in ("y", "tk", "ht"):
        a, b = getattr(g, name), getattr(f, name)
        assert np.array_equal(a[..., rej], b[..., rej]), name
    assert (g.y[:, acc] >= 0.0).all()
    assert np.abs(g.y[:, acc].sum(0) + g.stmf[acc] - 1.0).max() <= 1e-6
    el = element_matrix()
    assert np.abs(el @ (g.y[:, acc] - f.y[:, acc])).max() <= 1e-6
    assert np.abs(g.tk[acc] - f.tk[acc]).max() > 0.0  # a change was applied
    for cc in acc[:: max(1, acc.size // 150)]:
        t = po.temperature(g.ht[cc], g.y[:, cc], f.tk[cc])
        assert abs(t - g.tk[cc]) * sc.tstr <= 0.01, (cc, t * sc.tstr, g.tk[cc] * sc.tstr)


@pytest.mark.gpu
def test_gpu_activity_rule(cuda):
    """A cell is accepted only when |dT| <= tol_dt and every |dY| <= tol_dy: with tol_dt
    set to the median |dT| that the permissive gate applies, the cells above it (and
    only those, among the ones that gate accepts) are rejected as active."""
    torch = cuda
    from unicorn_amd.chem import Controls, DeviceField
    from unicorn_amd.fields import flame_field
    from unicorn_amd.scales import Scales
    from unicorn_amd.surrogate import SurrogateStep
    sc = Scales()
    f = flame_field(61, 41, "C7H16", seed=11, preignition_band=0.1)
    stats = _field_stats(f)
    net = _toy_net(torch, seed=12)
    c = f.active()
    loose = DeviceField(f)
    ls = SurrogateStep(net, stats, f, **NO_ACTIVITY)
    ls.predict(loose, Controls(dt=f.dt))
    acc0 = ls.mask.cpu().numpy()[c] == 10
    dT = np.abs(loose.tk.cpu().numpy()[c] - f.tk[c]) * sc.tstr
    tol = float(np.median(dT[acc0]))
    dev = DeviceField(f)
    ss = SurrogateStep(net, stats, f, tol_dt=tol, tol_dy=1.0, tol_bc=1e9)
    st = ss.predict(dev, Controls(dt=f.dt)).cpu().numpy()
    acc = ss.mask.cpu().numpy()[c] == 10
    print(st[:5], tol, acc0.sum())
    assert st[4] > 0 and st[0] == acc.sum()
    # T0 + dT rounds once more than dT itself: compare away from the threshold
    assert acc[acc0 & (dT < 0.999 * tol)].all()
    assert not acc[dT > 1.001 * tol].any()


@pytest.mark.gpu
def test_gpu_surrogate_step_with_fallback(cuda):
    """step() = main:821-828 then predict() on the accepted cells, bitwise, plus, on the
    rejected ones, bitwise what the FP64 step gives for the field with the accepted cells
    masked out (ISKIP = 10); the field's own ISKIP is restored."""
    torch = cuda
    from unicorn_amd.chem import ChemStep, Controls, DeviceField, clip
    from unicorn_amd.fields import flame_field
    from unicorn_amd.scales import Scales
    from unicorn_amd.surrogate import SurrogateStep
    sc = Scales()
    f = flame_field(61, 41, "C7H16", seed=7, preignition_band=0.1)
    stats = _field_stats(f)
    c = f.active()
    stats["x"]["hi"][0] = float(np.median(f.tk[c] * sc.tstr))
    net = _toy_net(torch, seed=9)
    ctl = Controls(dt=f.dt)
    ss = SurrogateStep(net, stats, f, **NO_ACTIVITY)
    pa = DeviceField(f)
    clip(pa)
    ss.predict(pa, ctl)
    pa = pa.to_host()
    sb = DeviceField(f)
    st = ss.step(sb, ctl)
    sb = sb.to_host()
    mask = ss.mask.cpu().numpy()
    acc, rej = c[mask[c] == 10], c[mask[c] != 10]
    assert st["rejected"] == rej.size and rej.size > 0 and acc.size > 0
    for name in ("y", "tk", "ht"):
        assert np.array_equal(getattr(sb, name)[..., acc], getattr(pa, name)[..., acc]), name
    ref = DeviceField(f)
    ref.iskip = torch.from_numpy(mask).cuda()
    ChemStep(f.n).step(ref, ctl)
    ref = ref.to_host()
    for name in ("y", "tk", "ht", "stmf", "stnd"):
        assert np.array_equal(getattr(sb, name)[..., rej], getattr(ref, name)[..., rej]), name
    assert np.array_equal(sb.iskip, f.iskip)


@pytest.mark.gpu
def test_gpu_density_matches_fp64_finish(cuda):
    """us_density (used when the gate accepts every cell) gives the FP64 finish's
    density on the state that finish saw: every cell and the edges."""
    torch = cuda
    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import flame_field
    from unicorn_amd.surrogate import _p, _stream, library, thermo_block, _to_device
    f = flame_field(61, 41, "CH4", seed=3)
    dev = DeviceField(f)
    ChemStep(f.n).step(dev, Controls(dt=f.dt))
    rho_fp64 = dev.rho.cpu().numpy()
    dev.rho.zero_()
    th = _to_device(thermo_block(_field_stats(f)))
    assert library().us_density(_p(th), _p(dev.y), _p(dev.tk), _p(dev.rho), f.li, f.lj, _stream()) == 0
    rho = dev.rho.cpu().numpy()
    assert np.allclose(rho, rho_fp64, rtol=1e-13, atol=0.0), np.abs(rho / rho_fp64 - 1).max()
