"""Train the chemistry surrogate and evaluate it with FP64 fallback.

    python -m unicorn_amd.surrogate_run train  [--pairs 20000000 --epochs 30 --out PATH]
    python -m unicorn_amd.surrogate_run infer  [--weights PATH --steps 200]

train: (state, increment) pairs from this project's FP64 step on seeded fields
(unicorn_amd.surrogate.training_fields), Adam at lr 1e-3 with a cosine schedule,
batch 8192, 5 % held out; writes the weights and normalisation to PATH.
infer: the 1001 x 401 C7H16 field (unicorn_amd.fields.config1), the given number of
chained steps at dt = 1e-5 s through the gated surrogate with FP64 fallback and,
separately, through the FP64 step alone (time per step of both); then both in
lockstep for the errors against the FP64 run: the one-step error (the surrogate
step applied to the FP64 run's state) and the chained error, max |dT| and max |dY|
over the major species, with the fallback rate and the sum-of-Y residuals.
One JSON line each.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "surrogate_weights.npz"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cmd_train(a):
    import torch

    from . import surrogate as S
    t0 = time.perf_counter()
    ds = S.generate_pairs(a.pairs, seed=a.seed, log=log)
    t_gen = time.perf_counter() - t0
    stats = S.column_stats(ds, a.scale_quantile)
    net, info = S.train(ds, stats, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    stats["stmf_max"] = ds.stmf_max
    stats["stnd_max"] = ds.stnd_max
    S.save(a.out, net, stats, {"pairs": ds.rows, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
                               "scale_quantile": a.scale_quantile})
    torch.cuda.synchronize()
    print(json.dumps({"train": {"pairs": ds.rows, "generate_s": t_gen, "train_s": info["seconds"],
                                "final_train_mse": info["history"][-1][1], "final_val_mse": info["history"][-1][2],
                                "history": info["history"], "weights": str(a.out)}}))


def _copy_state(dst, src):
    for name in ("rho", "y", "tk", "ht", "stmf", "stnd"):
        getattr(dst, name).copy_(getattr(src, name))


def cmd_infer(a):
    from .fields import config1
    out = evaluate(config1(), a.steps, a.weights, a.tol_el, a.tol_sum, a.tol_dt, a.tol_dy, a.tol_bc)
    out["workload"] = f"1001x401 C7H16 field, {a.steps} chained steps at dt=1e-5 s"
    print(json.dumps({"infer": out}))


def evaluate(f, steps: int, weights=DEFAULT_WEIGHTS, tol_el: float = 1e-6, tol_sum: float = 1e-6,
             tol_dt: float = 5.0, tol_dy: float = 2e-3, tol_bc: float = 0.03) -> dict:
    """Throughput and error of the gated surrogate with FP64 fallback on field f over
    `steps` chained steps from its state, against the FP64 step alone."""
    import torch

    from .chem import ChemStep, Controls, DeviceField
    from .scales import Scales, mechanism
    from .surrogate import MAJORS, SurrogateStep, load
    from types import SimpleNamespace
    a = SimpleNamespace(steps=steps, tol_el=tol_el, tol_sum=tol_sum, tol_dt=tol_dt, tol_dy=tol_dy, tol_bc=tol_bc)
    net, stats = load(weights)
    act = torch.from_numpy(f.active()).cuda()
    nact = int(act.numel())
    ctl = Controls(dt=f.dt)
    names = [s["name"] for s in mechanism()["species"]]
    majors = torch.tensor([names.index(m) for m in MAJORS], device="cuda")
    tstr = Scales().tstr

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = [fn(i) for i in range(a.steps)]
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / a.steps * 1e3, out

    # 1. throughput, FP64 step alone
    ref = DeviceField(f)
    ctx = ChemStep(f.n)
    ms_fp64, _ = timed(lambda i: ctx.step(ref, ctl))
    log(f"fp64: {ms_fp64:.3f} ms/step")
    # 2. throughput, gated surrogate + FP64 fallback
    dev = DeviceField(f)
    sstep = SurrogateStep(net, stats, f, tol_el=a.tol_el, tol_sum=a.tol_sum, tol_dt=a.tol_dt, tol_dy=a.tol_dy,
                          tol_bc=a.tol_bc)
    ms_sur, per = timed(lambda i: sstep.step(dev, ctl))
    rej = [p["rejected"] / nact for p in per]
    log(f"surrogate: {ms_sur:.3f} ms/step, fallback {np.mean(rej):.4f}")
    # 3. accuracy: both paths in lockstep from the same state, the chained error at
    # every step, and the one-step error (the surrogate step applied to the FP64 run's
    # own state)
    ref = DeviceField(f)
    dev = DeviceField(f)
    one = DeviceField(f)
    chained_dT, chained_dY, one_dT, one_dY = [], [], [], []
    for i in range(a.steps):
        _copy_state(one, ref)
        ctx.step(ref, ctl)
        sstep.step(dev, ctl)
        sstep.step(one, ctl)
        for fld, lt, ly in ((dev, chained_dT, chained_dY), (one, one_dT, one_dY)):
            lt.append(float((fld.tk[act] - ref.tk[act]).abs().max()) * tstr)
            ly.append(float((fld.y[majors][:, act] - ref.y[majors][:, act]).abs().max()))
    gr, gs = ref.to_host(), dev.to_host()
    ca = f.active()
    sum_ref = np.abs(gr.y[:, ca].sum(0) + gr.stmf[ca] - 1.0).max()
    sum_sur = np.abs(gs.y[:, ca].sum(0) + gs.stmf[ca] - 1.0).max()
    return {
        "active_cells": nact, "cell_updates_per_s": nact / (ms_sur * 1e-3),
        "fallback_rate_mean": float(np.mean(rej)), "fallback_rate_max": float(np.max(rej)),
        "fallback_rate_first_last": [rej[0], rej[-1]],
        "one_step_max_abs_dT_K": max(one_dT), "one_step_max_abs_major_species_error": max(one_dY),
        "chained_max_abs_dT_K": max(chained_dT), "chained_max_abs_major_species_error": max(chained_dY),
        "chained_final_abs_dT_K": chained_dT[-1], "chained_final_major_species_error": chained_dY[-1],
        "sumY_residual_fp64": float(sum_ref), "sumY_residual_surrogate": float(sum_sur),
        "ms_per_step_fp64": ms_fp64, "ms_per_step_surrogate": ms_sur,
        "second_half": {"ms_step": float(np.mean([p["ms_total"] for p in per[a.steps // 2:]])),
                        "ms_predict": float(np.mean([p["ms_surrogate"] for p in per[a.steps // 2:]])),
                        "fallback_rate": float(np.mean(rej[a.steps // 2:])),
                        "steps_with_fallback": int(sum(p["rejected"] > 0 for p in per[a.steps // 2:]))},
        "speedup": ms_fp64 / ms_sur,
        "gate": {"tol_el": a.tol_el, "tol_sum": a.tol_sum, "tol_dT_K": a.tol_dt, "tol_dY": a.tol_dy, "tol_dBC": a.tol_bc},
        "gate_rejections_last_step": {k: per[-1][k] for k in ("out_of_range", "residual", "sum", "active")},
        "raw_element_residual_max_last_step": per[-1]["raw_element_residual_max"],
        "ms_network_and_gate_last_step": per[-1]["ms_surrogate"],
        "one_step_dT_by_step": one_dT[:: max(1, a.steps // 20)],
        "chained_dT_by_step": chained_dT[:: max(1, a.steps // 20)]}


def main(argv=None):
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=20_000_000)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=7)
    t.add_argument("--out", type=Path, default=DEFAULT_WEIGHTS)
    t.add_argument("--scale-quantile", type=float, default=0.5,
                   help="quantile of |change| that sets each target's asinh compression scale")
    i = sub.add_parser("infer")
    i.add_argument("--weights", type=Path, default=DEFAULT_WEIGHTS)
    i.add_argument("--steps", type=int, default=200)
    i.add_argument("--tol-el", type=float, default=1e-6)
    i.add_argument("--tol-sum", type=float, default=1e-6)
    i.add_argument("--tol-dt", type=float, default=5.0)
    i.add_argument("--tol-dy", type=float, default=2e-3)
    i.add_argument("--tol-bc", type=float, default=0.03)
    a = p.parse_args(argv)
    {"train": cmd_train, "infer": cmd_infer}[a.cmd](a)


if __name__ == "__main__":
    main()
