"""Benchmark: UNICORN's chemistry step on one MI355X.

Workload (one "step" = one chemistry step of the whole field, inputs resident
in HBM): the 1001 x 401 axisymmetric C7H16 field with a 10 % pre-ignition band
at 700-1100 K (unicorn_amd.fields.config1, seed 1), dt = 1e-5 s, the
reference's solver controls (100 sweeps max, relaxation 0.99, tolerance 1e-8,
global stops).  The timed region is K consecutive steps starting from the
seeded state (default K = 20); W warm-up steps run first on the same field and
the state is then restored.

Prints one JSON line: cell-chemistry updates/s (interior cells x steps / s),
the roofline of the dominant kernel k_sweep (FP64, algorithmic operation count
of the sweeps it executed, timed with HIP events on its stream) and the CPU
baseline: the reference Fortran (flang build, oracle/_ref) on a sample of the
same workload -- the same cells' 20 steps with the sweep counts the GPU run
found -- one process per core.  With the trained surrogate weights present, the
line also carries "surrogate": the gated surrogate with FP64 fallback over the same
K steps from the same seeded state (cell-updates/s, speedup over the FP64 step,
fallback rate) and its error against the FP64 run (one-step and chained max |dT|
and max major-species error).  And "coupled": the same field with transport
coefficients (unicorn_amd.fields.synthetic_coupling, seed 3) through
uc_step_coupled, the reference's Gauss-Seidel as a wavefront, with its serial CPU
baseline (--workload coupled prints that measurement as the line).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK = 78.6e12          # MI355X FP64 (vector = matrix), FLOP/s
HBM_PEAK = 8.0e12            # B/s
# algorithmic work of one cell-update (static counts of the reference source):
# 12,760 basic FP64 ops + 650 transcendentals + 5,629 per species sweep
# + 28 per Newton iteration + 18 per enthalpy sweep; 880 B in+out per cell.
F_BASE, F_SWEEP, F_NEWTON, F_HSWEEP = 12760 + 650, 5629, 28, 18
B_CELL = 880
# a coupled cell-sweep adds 4 products and 4 subtractions per relaxed scalar (51 species,
# SST1, SST2; PPP is formed once per step): solvesp.f:424-426
F_SWEEP_CPL = F_SWEEP + 8 * 53
# the coupled workload's coefficients: diffusion number 0.02 per face (dt*D/dx^2 for
# dt = 1e-5 s, D = 2e-5 m^2/s, dx = 1e-4 m: laminar diffusion at this resolution), the
# enthalpy flux 2e-4 of rho*HT, velocities up to 0.3 VREF
CPL_RECIPE = {"strength": 0.02, "flux": 2e-4, "velocity": 0.3}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--cpu-cells", type=int, default=16000, help="CPU baseline sample size (cells)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-surrogate", action="store_true", help="skip the gated-surrogate measurement")
    p.add_argument("--no-coupled", action="store_true", help="skip the coupled-step measurement")
    p.add_argument("--no-pcie", action="store_true", help="skip the PCIe-inclusive (uc_step_host) measurement")
    p.add_argument("--li", type=int, default=1001)
    p.add_argument("--lj", type=int, default=401)
    p.add_argument("--workload", choices=("field", "batch", "coupled"), default="field",
                   help="field: the 1001x401 C7H16 field (the headline line); batch: 256 grids "
                        "of uneven size, fuel and dt in one batched step (autotuned first); coupled: "
                        "the field with transport coefficients (uc_step_coupled)")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB (5 times): the reference "
                        "kernel for scaling FETCH_SIZE / WRITE_SIZE counter readings")
    return p.parse_args()


def copy_probe(mib: int):
    """A plain copy of known size: reads and writes mib MiB per launch."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 5}))


def cpu_baseline(field0, nstars, nhs, ncells, steps):
    """The reference Fortran on `ncells` random interior cells of the same field,
    the same `steps` chained steps, each step forced to the GPU's sweep counts
    (exact: every cell of the reference runs the global count).  P processes in
    parallel, one core each; the 711 x 131 embedding grid's fixed per-step cost
    (an empty run) is subtracted."""
    from oracle import refrun
    if not refrun.available():
        return None
    P, host = host_cores()
    rng = np.random.default_rng(99)
    inter = field0.interior()
    cells = np.sort(rng.choice(inter, size=min(ncells, len(inter)), replace=False))
    chunks = np.array_split(cells, P)

    def run_chunk(ch):
        f = field0.take(ch)
        t = 0.0
        for s in range(steps):
            r = refrun.run(f, nsteps=1, isor1=int(nstars[s]), isor2=int(nhs[s]), force=True)
            t += r.t_chem
            f.rho, f.y, f.tk, f.ht, f.stmf, f.stnd = r.rho, r.y, r.tk, r.ht, r.stmf, r.stnd
        return t

    # the embedding grid's own cost per step (loops over all 711 x 131 cells, ISKIP
    # tests, whole-grid SOLVEH sweeps) with the same sweep counts, from a 1-cell run
    empty = field0.take(inter[:1])

    def run_empty(s):
        return refrun.run(empty, nsteps=1, isor1=int(nstars[s]), isor2=int(nhs[s]), force=True).t_chem

    with ThreadPoolExecutor(P) as ex:
        t_empty = sum(ex.map(run_empty, range(steps)))
        times = list(ex.map(run_chunk, chunks))
    t = max(times) - t_empty
    return {"value": len(cells) * steps / t, "unit": "cell-updates/s", "cores": P, "kind": "reference",
            "host": host, "embedding_overhead_fraction": t_empty / max(times),
            "sample": f"{len(cells)} random interior cells of the same field, {steps} chained steps with the "
                      f"GPU's per-step sweep counts, reference Fortran (flang -O2) in {P} processes of "
                      f"{len(chunks[0])} cells, one per core the process may use; {t_empty:.2f} s of 711x131 "
                      f"embedding-grid overhead ({steps} steps) subtracted from the slowest process's "
                      f"{max(times):.2f} s"}


def host_cores():
    """The cores this process may run on: its CPU affinity, capped by the cgroup's CPU
    quota where one is set; with the host's logical CPU count and CPU model."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    cores = aff if quota is None else max(1, min(aff, int(quota)))
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, {"nproc": os.cpu_count(), "affinity": aff, "cgroup_cpu_quota": quota, "cpu_model": model}


def batch_bench(args):
    """256 grids (unicorn_amd.fields.batch_grids) advanced together by uc_step_batch."""
    import torch

    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import batch_grids

    grids = batch_grids(256, seed=2)
    ncell = sum((f.li - 2) * (f.lj - 2) for f in grids)
    devs = [DeviceField(f) for f in grids]
    ctls = [Controls(dt=f.dt) for f in grids]
    ctx = ChemStep(sum(f.n for f in grids), max_grids=len(grids))
    t0 = time.perf_counter()
    tuning = ctx.autotune(devs, ctls)
    t_tune = time.perf_counter() - t0
    for _ in range(args.warmup):
        ctx.step_batch(devs, ctls)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [ctx.step_batch(devs, ctls) for _ in range(args.steps)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    sweeps = sum(s.cell_sweeps for st in stats for s in st)
    ms_sweep = sum(st[0].ms_sweeps for st in stats)
    launches = sum(st[0].sweep_launches for st in stats)
    out = {"metric": "cell-chem updates/s (batched grids)", "value": ncell * args.steps / el,
           "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic (unicorn_amd.fields.batch_grids, seed 2)",
           "config": {"workload": "256 grids, LI log-uniform 31..301, LJ log-uniform 21..151, fuel per grid "
                                  "CH4/C2H4/C7H16, dt per grid log-uniform 1e-6..1e-5 s, global stops per grid",
                      "grids": len(grids), "cells": ncell},
           "tuning": tuning, "autotune_s": t_tune,
           "k_sweep": {"avg_launch_ms": ms_sweep / launches, "launches": launches,
                       "achieved_tflops": F_SWEEP * sweeps / (ms_sweep * 1e-3) / 1e12},
           "sweeps_per_cell_step": sweeps / (ncell * args.steps),
           "species_sweeps_per_grid_last_step": sorted({s.isor1 for s in stats[-1]})}
    print(json.dumps(out))


def coupled_cpu_baseline(nstar, nh):
    """The reference Fortran's coupled step (flang build, oracle/_ref) on a 201 x 81 field
    of the same recipe with the same coupling recipe, one step forced to the GPU's sweep
    counts, one core (the coupled Gauss-Seidel is serial); scaled per cell-sweep."""
    from oracle import refrun
    from unicorn_amd.fields import flame_field, synthetic_coupling
    if not refrun.available():
        return None
    f = flame_field(201, 81, "C7H16", 1, preignition_band=0.10)
    cp = synthetic_coupling(f, seed=3, **CPL_RECIPE)
    r = refrun.run(f, nsteps=1, isor1=int(nstar), isor2=int(nh), force=True, coupling=cp)
    cells = (f.li - 2) * (f.lj - 2)
    return {"t_s": r.t_chem, "cells": cells, "sweeps": int(nstar),
            "s_per_cell_sweep": r.t_chem / (cells * nstar), "cores": 1}


# SPGSOLV's coefficient assembly (uc_assemble): algorithmic HBM bytes per cell.  k_asm_ekt
# reads TK, VTC, TMU and 52 Y and writes EKT (56 doubles); k_asm_coef reads RHO, U, V, TMU,
# Y (radius), EKT, STDF and 51 VDF and writes 212 coefficients (270 doubles); neighbours'
# values are counted once, as a cell's own.
ASM_BYTES_PER_CELL = 8 * (56 + 270)


def assemble_bench(ctx, dev, ctl, f0, reps=20):
    """Time uc_assemble on the field (synthetic_transport seed 3) with HIP events on the
    stream it runs on; the roofline is HBM (the assembly is per-face memory work)."""
    import torch

    from unicorn_amd.chem import DeviceTransport
    from unicorn_amd.fields import synthetic_transport
    dtr = DeviceTransport(synthetic_transport(f0, seed=3))
    s = torch.cuda.current_stream()
    ctx.assemble(dev, ctl, dtr)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        ctx.assemble(dev, ctl, dtr)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    achieved = ASM_BYTES_PER_CELL * f0.n / (ms * 1e-3)
    return {"kernel": "k_asm_ekt + k_asm_coef (uc_assemble)", "ms_per_call": ms, "bound": "hbm",
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "bytes_per_cell": ASM_BYTES_PER_CELL,
            "note": "SPGSOLV's E/W/P/Q, STE..STS, ZE..ZS and EKT for every cell; equal bit for bit to spgsolv.f"}


# transport properties (uc_transport_props): algorithmic HBM bytes per cell, 53 doubles read
# (Y, T) and 55 written (VMU, VTC, 52 VDF, STDF)
TR_BYTES_PER_CELL = 8 * (53 + 55)


def transport_bench(ctx, dev, f0, reps=20):
    """Time uc_transport_props on the field with HIP events on its stream; with the
    reference Fortran's cost of the same block (main:843-1005) on the host for comparison
    when the oracle's program build is present."""
    import torch
    s = torch.cuda.current_stream()
    ctx.transport_props(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        ctx.transport_props(dev)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    achieved = TR_BYTES_PER_CELL * f0.n / (ms * 1e-3)
    sub = (f0.y[:51] >= 1e-3).sum(axis=0) + 1
    return {"kernel": "k_transport (uc_transport_props)", "ms_per_call": ms, "cells_per_s": f0.n / (ms * 1e-3),
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "bytes_per_cell": TR_BYTES_PER_CELL,
            "submixture_species_mean": float(sub.mean()), "submixture_species_max": int(sub.max()),
            "note": "main:843-1005 for every cell, bit for bit with the reference program's /CB07/; "
                    "work per cell grows with the square of the sub-mixture size (species with Y >= 1e-3)"}


def coupled_bench(args, f0=None, cpu=True):
    """The field with synthetic transport coefficients through uc_step_coupled."""
    import torch

    from unicorn_amd.chem import ChemStep, Controls, DeviceCoupling, DeviceField
    from unicorn_amd.fields import flame_field, synthetic_coupling

    f0 = f0 if f0 is not None else flame_field(args.li, args.lj, "C7H16", 1, preignition_band=0.10)
    cp = synthetic_coupling(f0, seed=3, **CPL_RECIPE)
    ncell = (f0.li - 2) * (f0.lj - 2)
    dev, dcp = DeviceField(f0), DeviceCoupling(cp)
    init = {k: getattr(dev, k).clone() for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    ctl = Controls(dt=f0.dt)
    ctx = ChemStep(f0.n)
    for _ in range(args.warmup):
        ctx.step_coupled(dev, ctl, dcp)
    for k, v in init.items():
        getattr(dev, k).copy_(v)
    ctx.close()
    ctx = ChemStep(f0.n)
    steps = max(1, min(args.steps, 10))   # at ~0.1 s a step, 10 chained steps are plenty
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [ctx.step_coupled(dev, ctl, dcp) for _ in range(steps)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_sweep = sum(s.ms_sweeps for s in stats)
    sweeps = sum(s.cell_sweeps for s in stats)
    ndiag = (f0.li - 2) + (f0.lj - 2) - 1
    wave_steps = sum(s.sweep_launches for s in stats)   # k_sweep_cpl launches = wavefront steps
    achieved = F_SWEEP_CPL * sweeps / (ms_sweep * 1e-3)
    out = {"metric": "cell-chem updates/s, coupled step (transport coefficients, exact Gauss-Seidel order)",
           "value": ncell * steps / el, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
           "warmup": args.warmup, "ms_per_step": el / steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (unicorn_amd.fields.config1 + synthetic_coupling seed 3, "
                   "strength 0.02, flux 2e-4, velocity 0.3)",
           "config": {"workload": f"{f0.li}x{f0.lj} C7H16 field, 10% pre-ignition band, dt=1e-5 s, E/W/P/Q of "
                                  "every species, STE..STS, ZE..ZS, enthalpy flux inputs and velocities; "
                                  f"{steps} chained steps, global stops", "cells": ncell},
           "roofline": {"kernel": "k_sweep_cpl", "bound": "fp64_valu", "achieved": achieved / 1e12,
                        "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK, "traffic": None,
                        "note": "latency bound, not throughput bound: the exact order puts ndiag + 2*(S-1) "
                                "wavefront steps on the critical path, each one cell-sweep of one wave; "
                                f"{F_SWEEP_CPL} FP64 ops per cell-sweep"},
           "wavefront": {"species_ms_per_step": ms_sweep / steps,
                         "us_per_wavefront_step": ms_sweep * 1e3 / max(1, wave_steps),
                         "wavefront_steps_per_step": wave_steps / steps, "diagonals": ndiag},
           "sweeps": {"global_species": [s.isor1 for s in stats], "global_enthalpy": [s.isor2 for s in stats],
                      "first_segment": [s.first_segment for s in stats]}}
    out["assemble"] = assemble_bench(ctx, dev, ctl, f0)
    out["transport"] = transport_bench(ctx, dev, f0)
    if cpu and not args.no_cpu:
        nstar = int(np.median([min(s.isor1, 100) for s in stats]))
        nh = int(np.median([min(s.isor2, 100) for s in stats]))
        c = coupled_cpu_baseline(nstar, nh)
        if c:
            gpu_sweeps_per_cell_step = sweeps / (ncell * steps)
            v = 1.0 / (c["s_per_cell_sweep"] * gpu_sweeps_per_cell_step)
            out["cpu_baseline"] = {"value": v, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                                   "sample": f"reference Fortran (flang -O2), one coupled step of a 201x81 field of "
                                             f"the same recipe, forced to {nstar}/{nh} sweeps ({c['t_s']:.2f} s, "
                                             f"{c['s_per_cell_sweep'] * 1e6:.2f} us per cell-sweep), scaled to "
                                             f"this run's sweeps per cell-step; serial, as the reference's "
                                             f"Gauss-Seidel is"}
            out["gpu_over_cpu"] = out["value"] / v
    return out


def main():
    args = parse()
    if args.workload == "batch":
        batch_bench(args)
        return
    if args.workload == "coupled":
        print(json.dumps(coupled_bench(args)))
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the chemistry step shards by cells)", file=sys.stderr)
        sys.exit(2)
    import torch

    from unicorn_amd.chem import ChemStep, Controls, DeviceField
    from unicorn_amd.fields import flame_field

    f0 = flame_field(args.li, args.lj, "C7H16", 1, preignition_band=0.10)
    ncell = (args.li - 2) * (args.lj - 2)
    dev = DeviceField(f0)
    init = {k: getattr(dev, k).clone() for k in ("rho", "y", "tk", "ht", "stmf", "stnd")}
    ctx = ChemStep(f0.n)
    ctl = Controls(dt=f0.dt)
    for _ in range(args.warmup):
        ctx.step(dev, ctl)
    for k, v in init.items():
        getattr(dev, k).copy_(v)
    # the first-pass cap starts from scratch for the timed region
    ctx.close()
    ctx = ChemStep(f0.n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [ctx.step(dev, ctl) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    el = t1 - t0
    ms_sweep = sum(s.ms_sweeps for s in stats)
    launches = sum(s.sweep_launches for s in stats)
    sweeps = sum(s.cell_sweeps for s in stats)
    # dominant kernel: k_sweep; algorithmic FP64 operations of the sweeps it executed
    achieved = F_SWEEP * sweeps / (ms_sweep * 1e-3)
    nstars = [min(s.isor1, 100) for s in stats]
    nhs = [min(s.isor2, 100) for s in stats]
    traffic = None  # HBM bytes per k_sweep launch from the committed FETCH_SIZE/WRITE_SIZE runs
    tf = ROOT / "profiles" / "k_sweep_hbm_c7h16_1001x401.md"
    if tf.exists():
        m = re.search(r"^traffic_bytes_per_launch: ([0-9.eE+-]+)", tf.read_text(), re.M)
        traffic = float(m.group(1)) if m else None
    out = {
        "metric": "cell-chem updates/s (52 sp, FP64 vs surrogate @ stated err); % FP64/HBM peak",
        "value": ncell * args.steps / el,
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded field, unicorn_amd.fields.config1)",
        "config": {"workload": "1001x401 C7H16 field, 10% pre-ignition band, dt=1e-5 s, "
                               f"{args.steps} chained chemistry steps, global sweep stops; minor species "
                               "log-uniform in [1e-8,1e-4] scaled by 10^(-30*theta), theta=clip((1500 K-T)/1200 K,"
                               "0,1), so cold cells stay where the reference step converges",
                   "cells": ncell, "li": args.li, "lj": args.lj, "fuel": "C7H16", "dt_s": 1e-5,
                   "parallelism": "1 GPU"},
        "roofline": {"kernel": "k_sweep", "bound": "fp64_valu", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK, "traffic": traffic,
                     "avg_launch_ms": ms_sweep / launches, "launches": launches,
                     "note": "FP64 vector (VALU) peak, no MFMA used (MI355X's FP64 vector and matrix peaks are "
                             "both 78.6 TFLOP/s); algorithmic operations of the sweeps executed (5629 per "
                             "cell-sweep); traffic = measured HBM bytes per launch"},
        "hbm_algorithmic_frac": B_CELL * ncell * args.steps / el / HBM_PEAK,
        "ms_kernels": {"k_rates": sum(s.ms_rates for s in stats), "k_sweep": ms_sweep,
                       "k_sources": sum(s.ms_sources for s in stats),
                       "step_total_device": sum(s.ms_total for s in stats)},
        "sweeps": {"global_species": [s.isor1 for s in stats], "global_enthalpy": [s.isor2 for s in stats],
                   "executed_per_cell_step": sweeps / (ncell * args.steps),
                   "redo_cells": [s.redo_cells for s in stats], "first_segment": [s.first_segment for s in stats]},
    }
    if not args.no_surrogate:
        from unicorn_amd.surrogate_run import DEFAULT_WEIGHTS, evaluate
        if DEFAULT_WEIGHTS.exists():
            ev = evaluate(f0, args.steps)
            out["surrogate"] = {k: ev[k] for k in (
                "cell_updates_per_s", "ms_per_step_surrogate", "ms_per_step_fp64", "speedup", "fallback_rate_mean",
                "one_step_max_abs_dT_K", "one_step_max_abs_major_species_error", "chained_max_abs_dT_K",
                "chained_max_abs_major_species_error", "sumY_residual_surrogate", "gate", "second_half")}
            # C, H, O residual of the raw prediction (before the gate's minimum-norm C/H/O
            # correction), max over the last step's cells; the gate tests the corrected one at tol_el
            out["surrogate"]["raw_element_residual_max_CHO"] = ev["raw_element_residual_max_last_step"]
            out["surrogate"]["dtype"] = "bf16 network (fp32 accumulation), fp64 gate and fallback"
        else:
            out["surrogate"] = None
    # PCIe-inclusive: uc_step_host copies the field in and out on every call (never `value`)
    hsteps = 0 if args.no_pcie else 3
    if hsteps:
        hf = f0.copy()
        hctx = ChemStep(f0.n)
        t0 = time.perf_counter()
        for _ in range(hsteps):
            hctx.step_host(hf, ctl)
        el_h = time.perf_counter() - t0
        hctx.close()
        out["pcie_inclusive"] = {"value": ncell * hsteps / el_h, "unit": "cell-updates/s",
                                 "ms_per_step": el_h / hsteps * 1e3, "steps": hsteps,
                                 "device_resident_ms_per_step_same_steps":
                                     sum(st.ms_total for st in stats[:hsteps]) / max(1, min(hsteps, len(stats))),
                                 "note": "uc_step_host from the seeded state: pageable host arrays copied in and "
                                         "out every step (first steps, so more sweeps than the 20-step average)"}
    if not args.no_coupled:
        c = coupled_bench(args, f0)
        out["coupled"] = {k: c[k] for k in ("value", "ms_per_step", "roofline", "wavefront", "sweeps", "assemble",
                                            "transport")}
        if "cpu_baseline" in c:
            out["coupled"]["cpu_baseline"] = c["cpu_baseline"]
            out["coupled"]["gpu_over_cpu"] = c["gpu_over_cpu"]
    if not args.no_cpu:
        cpu = cpu_baseline(f0, nstars, nhs, args.cpu_cells, args.steps)
        if cpu:
            out["cpu_baseline"] = cpu
            out["gpu_over_cpu"] = out["value"] / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
