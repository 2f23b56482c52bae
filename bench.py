#!/usr/bin/env python3
"""Benchmark: RK4 steps of the 8-field compressible MHD system on one MI355X.

Workload (default): one 256 x 512 x 1024 (r x theta x phi) subdomain, isentropic polytrope
(n=1.5, density contrast 20), seeded entropy perturbation (1e-3) and random vector potential
(1e-4), u = 0, dt from the CFL condition at 0.4, stress-free impenetrable radial walls, periodic
phi, SLD diffusion on.  A "step" is one full Step_RK4 (four RHS evaluations + stage updates +
Bound) on the whole subdomain with the state resident in HBM.

`value` is the FP32-bulk path (FP32 stages/RHS/diffusion, FP64 state y0); the FP64 path, the
100-step FP32-vs-FP64 precision study (max/RMS relative error per field every 10 steps),
per-phase kernel times, the roofline of the dominant kernel, the same measurement at
128 x 256 x 512 and a CPU baseline (the FP64 oracle, one step at 128 x 256 x 512) are
reported beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4|c2|c3|c5]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "css-code_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "RK4 Mpt-steps/s + HBM TB/s, FP32 vs FP64 at ≤1e-5 rel err; closure speedup"
HBM_PEAK_GBS = 8000.0
# compulsory stage traffic of one RK4 step, bytes per point (8 fields).  FP64: 17 field
# accesses x 8 B (stage 1 reads y0, writes acc and y1; stages 2-3 read y_s, y0, acc and write
# acc, y_s+1; stage 4 reads y3, y0, acc and writes y0).  Mixed path as stored: FP32 stage states,
# FP32 accumulator and FP32 copy of y0, FP64 y0 read and written once (stage 4, which also
# refreshes the FP32 copy): per field 12 + 20 + 20 + (4 + 4 + 8 + 8 + 4) = 80 B.
STEP_BYTES_F64 = 1088.0
STEP_BYTES_MIXED = 640.0


def _sync():
    import torch
    torch.cuda.synchronize()


def run_path(sol, y0, dt, steps, warmup, precision):
    import torch
    y = y0.clone()
    if warmup:
        w = y0.clone()
        sol.rk4(w, dt, warmup, precision)
        del w
    _sync()
    t0 = time.perf_counter()
    sol.rk4(y, dt, steps, precision)
    _sync()
    t1 = time.perf_counter()
    return y, (t1 - t0) * 1e3 / max(steps, 1)


def profile_phases(sol, y0, dt, steps, precision):
    from cssx import _lib
    y = y0.clone()
    _lib.timing_read()
    _lib.timing_enable(True)
    sol.rk4(y, dt, steps, precision)
    tab = _lib.timing_read()
    _lib.timing_enable(False)
    return tab


def cpu_baseline(sld_coef, shape=(128, 256, 512), seeds=(5, 6)):
    """The FP64 oracle (oracle/*.c, OpenMP) timed on one Step_RK4 of a bounded sample of the
    workload: the same recipe and seeds on a 128 x 256 x 512 grid (1/8 of the 256 x 512 x 1024
    points; about 15-30 s of CPU work on the host's cores)."""
    import oracle
    from cssx.model import Config, Model
    c = Config(nr=shape[0], nth=shape[1], nphi=shape[2], sld_coef=sld_coef, seed_s=seeds[0], seed_a=seeds[1])
    m = Model(c)
    y = m.initial_state()
    t0 = time.perf_counter()
    y = oracle.rk4(m, y, 1e-4, 1)
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(c.npts / el / 1e6, 4), "unit": "Mpt-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (oracle/*.c, OpenMP threads={threads}) one Step_RK4 with SLD of the same recipe "
                      f"on {c.nr}x{c.nth}x{c.nphi} (1/8 of the workload's points) in {el:.1f} s; the rate is per point, "
                      "so it stands for the full workload only as an extrapolation"}


def precision_study(sol, y0, dt, steps, nv, every=10):
    """FP64 and FP32-bulk runs from the same state; per field, the max and RMS of |y32 - y64|
    relative to max|y64| and RMS(y64), every `every` steps."""
    import torch
    y64 = y0.clone()
    y32 = y0.clone()
    rows = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        sol.rk4(y64, dt, k, "f64")
        sol.rk4(y32, dt, k, "f32")
        done += k
        mx, rms = [], []
        for iv in range(nv):
            d = y32[iv] - y64[iv]
            mx.append(d.abs().max().item() / max(y64[iv].abs().max().item(), 1e-300))
            rms.append(d.pow(2).mean().sqrt().item() / max(y64[iv].pow(2).mean().sqrt().item(), 1e-300))
        rows.append({"step": done, "max_rel": [float(f"{x:.3e}") for x in mx],
                     "rms_rel": [float(f"{x:.3e}") for x in rms]})
    torch.cuda.synchronize()
    worst = max(max(r["max_rel"]) for r in rows)
    return {"steps": steps, "fields": ["rho", "w", "v", "u", "s", "A_theta", "A_phi", "A_r"][:nv],
            "max_rel_err_over_run": worst, "meets_1e-5": worst <= 1e-5, "every": rows}


def run_c5(args):
    """Learned closure on the 256x512x1024 recipe (seeds 5/6) box-filtered by 4 to 64x128x256:
    training snapshots from a fine FP32-bulk run, training on the GPU, the a-priori gain on a
    held-out later snapshot, then fine vs coarse (plain SLD) vs coarse + closure runs from the
    same initial state, their statistics against the filtered fine run and their wall times."""
    import numpy as np
    import torch

    from cssx import closure as CL
    from cssx.mhd import Solver
    from cssx.model import Config, Model
    factor = 4
    fc = Config(nr=args.nr, nth=args.nth, nphi=args.nphi, sld_coef=args.sld_coef, seed_s=5, seed_a=6)
    fine = Solver(Model(fc))
    coarse = Solver(Model(CL.coarse_config(fc, factor)))
    y = fine.model.initial_state(device="cuda")
    dt = fine.cfl(y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    snaps = []
    for k in range(args.c5_snapshots):
        if k:
            fine.rk4(y, dt, args.c5_every, "f32")
        snaps.append(CL.snapshot(fine, coarse, y, factor, k * args.c5_every * dt, include_sld=args.c5_sld))
        print(f"c5: snapshot {k}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    del y
    fine.close()
    coarse.close()
    # interior-time snapshots 1 (validation: the per-field gate) and 3 (test) are held out
    val_id, test_id = 1, 3
    train_snaps = [sn for i, sn in enumerate(snaps) if i not in (val_id, test_id)]
    ts = CL.build_training_set(train_snaps, args.c5_samples, seed=7)
    n = ts.X.shape[0]
    net = CL.Network(128, 128, max_rows=262144, seed=7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = CL.train(net, ts, epochs=args.c5_epochs, batch=8192, lr=1e-3, seed=7)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    print(f"c5: trained {args.c5_epochs} epochs in {t_train:.2f} s, loss {hist[0]:.4f} -> {hist[-1]:.4f}",
          file=sys.stderr, flush=True)
    # a field's correction runs only where it removes >= 5% of the validation residual; the
    # rest falls back to the plain coarse RHS (an a-priori gain of a few % does not survive the
    # feedback of a run)
    mask = CL.field_gate(net, snaps[val_id], ts, threshold=0.95)
    before, after = CL.apriori_gain(net, snaps[test_id], ts, mask)
    # the correction's gain, set a posteriori on a calibration case (seeds 8/9) apart from the
    # evaluated one
    gains = [0.0, 0.1, 0.2, 0.3, 0.5, 0.75, 1.0]
    cc = Config(nr=args.nr, nth=args.nth, nphi=args.nphi, sld_coef=args.sld_coef, seed_s=8, seed_a=9)

    def _gain_progress(g, sc, w):
        print(f"c5: calibration gain {g} mean stat err {sc:.5g} worst {w:.5g}", file=sys.stderr, flush=True)

    gain, calib_score, calib_worst = CL.calibrate_gain(cc, factor, net, ts, mask, gains, args.c5_calib_steps,
                                                       args.c5_stat_every, "f32", progress=_gain_progress)
    # the closure tendency of a whole coarse state (the fused inference kernel), timed on the
    # FP32 fluctuation state the coarse run hands it, and on an FP64 absolute state
    from cssx.mhd import Solver
    cm = Model(CL.coarse_config(fc, factor))
    yc = cm.initial_state(device="cuda")
    solc = Solver(cm)
    y32 = solc.to_f32(yc)
    solc.close()

    def _time_tendency(state, fluct):
        ext = CL.tendency(net, state, cm, fluct=fluct)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            CL.tendency(net, state, cm, fluct=fluct, out=ext)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 10

    infer_ms = _time_tendency(y32, True)
    infer_ms_f64 = _time_tendency(yc, False)
    infer_tf = 90112.0 * cm.c.npts / (infer_ms * 1e-3) / 1e12
    del snaps, ts, train_snaps, yc, y32
    torch.cuda.empty_cache()
    def _progress(r):
        print(f"c5: coarse step {r['step']} ke fine/plain/closure "
              f"{r['filtered_fine']['ke']:.4e} {r['coarse']['ke']:.4e} {r.get('coarse_closure', {}).get('ke', 0.0):.4e}",
              file=sys.stderr, flush=True)

    recs, times = CL.compare(fc, factor, net, args.c5_steps, args.c5_stat_every, "f32", progress=_progress)
    ep, kp = CL.stat_errors(recs, "coarse")
    ec, kc = CL.stat_errors(recs, "coarse_closure")
    sp_err = CL.scalar_errors(recs, "coarse")
    sc_err = CL.scalar_errors(recs, "coarse_closure")
    flops_train = 3 * 90112.0 * n * len(hist)
    ach = flops_train / t_train / 1e12
    ratio = times["fine_s"] / max(times["coarse_closure_s"], 1e-12)
    worst_c = float(max(ec.max(), kc, max(sc_err.values())))
    worst_p = float(max(ep.max(), kp, max(sp_err.values())))
    matched = worst_c <= 0.05
    # the closure speedup is defined at matched statistics; an unmatched run reports no value,
    # only its wall-time ratio and by how much the statistics miss
    out = {
        "metric": METRIC + " [closure speedup]",
        "value": round(ratio, 2) if matched else None,
        "unit": "x (wall filtered-fine / wall coarse+closure, statistics within 5%)",
        "matched": matched,
        "target": "coarse+closure statistics within 5% of the filtered fine run, >= 10x less wall time",
        "worst_stat_rel_err": {"coarse_closure": round(worst_c, 4), "coarse_plain": round(worst_p, 4)},
        "wall_time_ratio_unmatched": None if matched else round(ratio, 2),
        "n_gpus": 1, "steps": args.c5_steps, "warmup": 0, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "bf16 network (f32 accumulate), f32-bulk solver", "data": "synthetic",
        "config": {"workload": f"closure_{fc.nr}x{fc.nth}x{fc.nphi}_to_{fc.nr // 4}x{fc.nth // 4}x{fc.nphi // 4}",
                   "factor": factor, "snapshots": args.c5_snapshots, "snapshot_every_fine_steps": args.c5_every,
                   "samples": n, "epochs": len(hist), "batch": 8192, "lr": 1e-3, "network": "216-128-128-8 GELU",
                   "coarse_steps": args.c5_steps, "dt_coarse": times["dt_coarse"], "dt_fine": times["dt_fine"],
                   "targets_include_sld": bool(args.c5_sld)},
        "wall_s": {"fine_run_and_filter": round(times["fine_s"], 3), "coarse_plain": round(times["coarse_plain_s"], 3),
                   "coarse_closure": round(times["coarse_closure_s"], 3), "training_data": round(t_data, 3),
                   "training": round(t_train, 3)},
        "stats_max_rel_err": {"coarse_plain": {"fields": [float(f"{x:.3e}") for x in ep], "ke": float(f"{kp:.3e}"),
                                               "scalar_output": {k: float(f"{v:.3e}") for k, v in sp_err.items()}},
                              "coarse_closure": {"fields": [float(f"{x:.3e}") for x in ec], "ke": float(f"{kc:.3e}"),
                                                 "scalar_output": {k: float(f"{v:.3e}") for k, v in sc_err.items()}},
                              "closure_within_5pct": matched},
        "closure_gain": {"gains": gains, "calibration_mean_stat_rel_err": [float(f"{x:.5g}") for x in calib_score],
                         "calibration_worst_stat_rel_err": [float(f"{x:.5g}") for x in calib_worst],
                         "chosen": gain, "calibration_case": "seeds 8/9", "calibration_coarse_steps": args.c5_calib_steps},
        "apriori_residual_rms": {"fields": list(CL.FIELDS), "gate": [bool(x) for x in mask],
                                 "held_out": {"validation": val_id, "test": test_id},
                                 "plain": [float(f"{x:.3e}") for x in before],
                                 "closure": [float(f"{x:.3e}") for x in after],
                                 "ratio": [float(f"{x:.3f}") for x in after / before]},
        "training_loss": [float(f"{x:.4g}") for x in hist],
        "roofline": {"kernel": "k_closure_infer (features + 3 BF16 MFMA layers, one coarse state)", "bound": "mfma",
                     "achieved": round(infer_tf, 3), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(infer_tf / 2500.0, 5), "traffic": None, "avg_launch_ms": round(infer_ms, 4),
                     "input": "f32 fluctuation state (as the coarse run)", "avg_ms_f64_absolute_state": round(infer_ms_f64, 4),
                     "algorithmic_flop_per_point": 90112},
        "training_throughput": {"kernels": "k_gemm_bf16 + k_mse/k_colsum/k_adam + row gathers",
                                "achieved_TFLOPs": round(ach, 3), "samples_per_s": round(n * len(hist) / t_train)},
        "records": recs,
    }
    net.close()
    print(json.dumps(_clean(out)))


def run_c3(args):
    """Batched uneven derivatives: patches with every axis length uniform in [32, 512] (seed 4),
    FP32 N(0,1) smoothed by [1,2,1]/4 per axis, all three first derivatives of a chunk of patches
    in one persistent launch (variant autotuned on the first chunk), chunks run back to back;
    then the 64-member ensemble of 48x96x96 MHD runs (Ra log-uniform in [1e4, 1e6]) stepped
    eagerly and as one captured HIP graph."""
    import numpy as np
    import torch

    from cssx.patches import Ensemble, PatchChunk, patch_shapes
    shapes = patch_shapes(args.c3_patches, seed=4)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_ms, pts, best, tune_ms = 0.0, 0, None, None
    for c0 in range(0, len(shapes), args.c3_chunk):
        ch = PatchChunk(shapes[c0:c0 + args.c3_chunk], prec="f32")
        ch.fill(4 + c0)
        if best is None:
            best, tune_ms = ch.autotune(1)
        ch.run(best)                       # warm the chunk's tables
        ev0.record()
        ch.run(best)
        ev1.record()
        torch.cuda.synchronize()
        t_ms += ev0.elapsed_time(ev1)
        pts += ch.points
        ch.close()
        del ch
        print(f"c3: chunk {c0 // args.c3_chunk} done", file=sys.stderr, flush=True)
    by = 24.0 * pts
    gbs = by / (t_ms * 1e-3) / 1e9
    ens = Ensemble(members=args.c3_members, shape=(48, 96, 96), seed=4)
    ens.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(args.c3_ens_steps)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / args.c3_ens_steps
    ens.capture()
    ens.replay(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.replay(args.c3_ens_steps)
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / args.c3_ens_steps
    epts = ens.points
    ens.close()
    out = {
        "metric": METRIC + " [batched uneven derivatives]",
        "value": round(pts / (t_ms * 1e-3) / 1e9, 3), "unit": "Gpt/s (3 directional derivatives per point)",
        "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"patches_{len(shapes)}_uniform_32_512_f32", "patches": len(shapes),
                   "chunk": args.c3_chunk, "points": pts, "variant": best, "variant_ms_first_chunk": tune_ms},
        "ms_total": round(t_ms, 3),
        "roofline": {"kernel": "k_patch_derivs", "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_point": 24},
        "ensemble": {"members": args.c3_members, "shape": [48, 96, 96], "ms_per_step_eager": round(t_eager * 1e3, 3),
                     "ms_per_step_graph": round(t_graph * 1e3, 3),
                     "Mpt_steps_per_s_eager": round(epts / t_eager / 1e6, 2),
                     "Mpt_steps_per_s_graph": round(epts / t_graph / 1e6, 2)},
    }
    print(json.dumps(_clean(out)))


def run_deriv(args):
    """The derivative sweeps alone: one seeded field of the --nr/--nth/--nphi shape, the compact
    first and second derivative along each axis (derivatives.f90 Dbydr/Dbydt/Dbydp and
    D2byd*2, boundary code 1 on the walls, phi periodic), each timed over --steps launches with
    HIP events after --warmup, in each precision of --paths.  One JSON line: microseconds per
    sweep and the algorithmic rate (one read and one write of the field)."""
    import torch
    from cssx import operators as O

    nr, nth, nphi = args.nr, args.nth, args.nphi
    ops = O.Operators(nr, nth, nphi, nr - 1.0, nth - 1.0, float(nphi))
    res = {}
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    for prec in args.paths.split(","):
        dt = torch.float32 if prec == "f32" else torch.float64
        f = torch.randn((nr, nphi, nth), dtype=dt, device="cuda", generator=g)
        out = torch.empty_like(f)
        for ax, name in ((0, "r"), (1, "theta"), (2, "phi")):
            for order in (1, 2):
                for _ in range(args.warmup):
                    ops.deriv(f, ax, order, 1, out=out)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                torch.cuda.synchronize()
                e0.record()
                for _ in range(args.steps):
                    ops.deriv(f, ax, order, 1, out=out)
                e1.record()
                torch.cuda.synchronize()
                us = e0.elapsed_time(e1) * 1e3 / args.steps
                tbs = 2.0 * f.numel() * f.element_size() / (us * 1e-6) / 1e12
                res[f"{prec}_{name}_d{order}"] = {"us": round(us, 1), "TB_s": round(tbs, 2)}
                print(f"bench: deriv {nr}x{nth}x{nphi} {prec} {name} order {order}: {us:.1f} us, {tbs:.2f} TB/s",
                      file=sys.stderr, flush=True)
        del f, out
    ops.close()
    print(json.dumps({"metric": "compact FD6 derivative sweep time", "unit": "us", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
                      "data": "synthetic", "config": {"workload": f"deriv_{nr}x{nth}x{nphi}", "nr": nr,
                                                      "nth": nth, "nphi": nphi}, "sweeps": res}))


def measured_traffic(workload):
    """The rocprofv3 FETCH_SIZE/WRITE_SIZE summary committed for this workload (FP32-bulk path):
    the JSON block of profiles/hbm_traffic_<workload>.md, or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    i = text.find("```json")
    if i < 0:
        return None
    j = text.find("```", i + 7)
    try:
        d = json.loads(text[i + 7:j])
    except ValueError:
        return None
    d["source"] = os.path.relpath(path, ROOT)
    return d


def _clean(o):
    """NaN/inf (a precision path that was not run) -> null, so the line stays strict JSON."""
    if isinstance(o, dict):
        return {k: _clean(v) for k, v in o.items()}
    if isinstance(o, list):
        return [_clean(v) for v in o]
    if isinstance(o, float) and not math.isfinite(o):
        return None
    return o


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c5-snapshots", type=int, default=8,
                    help="closure: fine snapshots (1 and 3 are held out for the gate and the test)")
    ap.add_argument("--c5-every", type=int, default=20, help="closure: fine steps between snapshots")
    ap.add_argument("--c5-samples", type=int, default=2_000_000, help="closure: training rows")
    ap.add_argument("--c5-epochs", type=int, default=10)
    ap.add_argument("--c5-steps", type=int, default=1000, help="closure: coarse steps of the comparison runs")
    ap.add_argument("--c5-stat-every", type=int, default=10)
    ap.add_argument("--c5-calib-steps", type=int, default=400, help="closure: coarse steps of the gain calibration run")
    ap.add_argument("--c5-sld", type=int, default=1,
                    help="closure targets include each grid's SLD diffusion (1) or the RHS alone (0)")
    ap.add_argument("--c3-patches", type=int, default=4096)
    ap.add_argument("--c3-chunk", type=int, default=64)
    ap.add_argument("--c3-members", type=int, default=64)
    ap.add_argument("--c3-ens-steps", type=int, default=5)
    ap.add_argument("--config", choices=["c2", "c3", "c4", "c5", "deriv"], default="c4",
                    help="c4: one 256x512x1024 subdomain with the precision study (the headline); c2: 128x256x512; "
                         "deriv: the derivative sweeps alone (--nr/--nth/--nphi, default 256x512x1024)")
    ap.add_argument("--no-c2", action="store_true", help="c4: leave out the 128x256x512 measurement")
    ap.add_argument("--nr", type=int, default=0)
    ap.add_argument("--nth", type=int, default=0)
    ap.add_argument("--nphi", type=int, default=0)
    ap.add_argument("--sld-coef", type=float, default=0.1,
                    help="SLD diffusion coefficient (SLDiffusion is on; 0 keeps its cost with a zero tendency)")
    ap.add_argument("--paths", default="f64,f32", help="precisions to run: f64,f32 | f32 | f64")
    ap.add_argument("--laplacian", type=int, default=0,
                    help="1: the explicit-diffusion physics (Laplacian viscosity, thermal diffusion and "
                         "resistivity) in place of SLD")
    ap.add_argument("--kappa", type=float, default=1e-5, help="kappa_value of the explicit diffusion")
    ap.add_argument("--study-steps", type=int, default=100,
                    help="FP32-vs-FP64 precision study length (untimed; errors every 10 steps; 0 = off)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    shape = {"c2": (128, 256, 512), "c3": (0, 0, 0), "c4": (256, 512, 1024), "c5": (256, 512, 1024),
             "deriv": (256, 512, 1024)}[args.config]
    args.nr, args.nth, args.nphi = (args.nr or shape[0], args.nth or shape[1], args.nphi or shape[2])
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition is a design, not a code path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cssx.mhd import Solver
    from cssx.model import Config, Model

    torch.cuda.set_device(0)
    if args.config == "c5":
        run_c5(args)
        return
    if args.config == "c3":
        run_c3(args)
        return
    if args.config == "deriv":
        run_deriv(args)
        return
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}}))
        return
    seeds = {"c2": (2, 3), "c4": (5, 6)}[args.config]     # entropy / vector-potential seeds
    out = run_mhd(args, (args.nr, args.nth, args.nphi), seeds, args.study_steps, full=True)
    if args.config == "c4" and not args.no_c2:
        # the same measurement at 128 x 256 x 512 (seeds 2/3), timed steps only
        c2 = run_mhd(args, (128, 256, 512), (2, 3), 0, full=False)
        out["c2"] = {k: c2[k] for k in ("value", "unit", "ms_per_step", "fp64", "fp32_speedup_vs_fp64",
                                        "fp32_vs_fp64", "hbm_tbs_algorithmic", "roofline")}
        out["c2"]["workload"] = c2["config"]["workload"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(out["config"]["sld_coef"])
    print(json.dumps(_clean(out)))


def run_mhd(args, shape, seeds, study_steps, full):
    """RK4 steps of one subdomain in FP64 and FP32-bulk: the timed steps, their agreement, the
    precision study and the per-kernel timings (all when `full`)."""
    import torch

    from cssx.mhd import Solver
    from cssx.model import Config, Model
    lap = bool(args.laplacian)
    c = Config(nr=shape[0], nth=shape[1], nphi=shape[2], sld_coef=0.0 if lap else args.sld_coef, seed_s=seeds[0],
               seed_a=seeds[1], Laplacian=lap, kappa_value=args.kappa, SLDiffusion=not lap)
    m = Model(c)
    y0 = m.initial_state(device="cuda")
    sol = Solver(m)
    dt = sol.cfl(y0)
    npts = c.npts
    workload = f"mhd_rk4_{c.nr}x{c.nth}x{c.nphi}_8field" + ("_laplacian" if lap else "")
    print(f"bench: {workload} dt={dt:.3e}", file=sys.stderr, flush=True)

    paths = set(args.paths.split(","))
    y64 = y32 = None
    ms64 = ms32 = float("nan")
    if "f64" in paths:
        y64, ms64 = run_path(sol, y0, dt, args.steps, args.warmup, "f64")
        print(f"bench: {workload} f64 {ms64:.3f} ms/step", file=sys.stderr, flush=True)
    if "f32" in paths:
        y32, ms32 = run_path(sol, y0, dt, args.steps, args.warmup, "f32")
        print(f"bench: {workload} f32 {ms32:.3f} ms/step", file=sys.stderr, flush=True)

    # FP32-bulk vs FP64 after the same K steps from the same state
    cmp = None
    if y64 is not None and y32 is not None:
        rel_state, rel_change, rms_state = [], [], []
        for iv in range(c.nv):
            diff = y32[iv] - y64[iv]
            d = diff.abs().max().item()
            rel_state.append(d / max(y64[iv].abs().max().item(), 1e-300))
            rms_state.append((diff.pow(2).mean().sqrt() / y64[iv].pow(2).mean().sqrt().clamp_min(1e-300)).item())
            ch = (y64[iv] - y0[iv]).abs().max().item()
            rel_change.append(d / max(ch, 1e-300))
        finite = bool(torch.isfinite(y64).all().item() and torch.isfinite(y32).all().item())
        cmp = {"steps": args.steps, "max_rel_err_state": max(rel_state), "max_rel_err_of_change": max(rel_change),
               "rms_rel_err_state": max(rms_state), "per_field_max_rel_err": [float(f"{x:.3e}") for x in rel_state],
               "finite": finite}
    del y64, y32

    study = None
    if full and study_steps > 0 and len(paths) == 2:
        study = precision_study(sol, y0, dt, study_steps, c.nv)
        print(f"bench: {workload} precision study {study['max_rel_err_over_run']:.2e}", file=sys.stderr, flush=True)

    prof_steps = max(2, min(args.steps, 5))
    tab32 = profile_phases(sol, y0, dt, prof_steps, "f32") if "f32" in paths else None
    tab64 = profile_phases(sol, y0, dt, prof_steps, "f64") if ("f64" in paths and full) else None

    families = {"k_deriv": ("deriv_r", "deriv_theta", "deriv_phi"), "k_rhs_final": ("rhs_final",),
                "k_sld": ("sld",), "pointwise_other": ("pointwise",), "rk_stage": ("rk_update",)}

    def family_rows(tab):
        rows = {}
        for name, ph in families.items():
            ms = sum(tab[k][0] for k in ph)
            n = sum(tab[k][1] for k in ph)
            by = sum(tab[k][2] for k in ph)
            if n:
                avg = ms / n
                ach = (by / n) / (avg * 1e-3) / 1e9
                rows[name] = {"ms_per_step": round(ms / prof_steps, 4), "launches": n, "avg_launch_ms": round(avg, 5),
                              "achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}
        return rows

    def roofline(tab):
        """Roofline of the dominant kernel family (largest summed time): achieved = algorithmic
        bytes per launch / average launch time, both from the per-launch hipEvent timer."""
        if tab is None:
            return None
        rows = family_rows(tab)
        name = max(rows, key=lambda k: rows[k]["ms_per_step"])
        ph = families[name]
        ms = sum(tab[k][0] for k in ph)
        n = sum(tab[k][1] for k in ph)
        by = sum(tab[k][2] for k in ph)
        avg_ms = ms / max(n, 1)
        achieved = (by / max(n, 1)) / (avg_ms * 1e-3) / 1e9
        return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(avg_ms, 5), "launches": n,
                "algorithmic_bytes_per_launch": by / max(n, 1)}

    val32 = npts / (ms32 * 1e-3) / 1e6
    val64 = npts / (ms64 * 1e-3) / 1e6
    tr = measured_traffic(workload)
    roof32 = roofline(tab32)
    if roof32 is not None and tr is not None and roof32["kernel"] in tr.get("families", {}):
        roof32["traffic"] = tr["families"][roof32["kernel"]]["bytes_per_launch"]
        roof32["traffic_source"] = tr["source"]
    out = {
        "metric": METRIC,
        "value": round(val32, 2) if "f32" in paths else round(val64, 2),
        "unit": "Mpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32 if "f32" in paths else ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if "f32" in paths else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "nr": c.nr, "nth": c.nth,
                   "nphi": c.nphi, "fields": c.nv, "state_y0": "f64", "dt": dt, "sld_coef": c.sld_coef,
                   "parallelism": "single-gpu",
                   **({"laplacian": True, "kappa_value": c.kappa_value, "sld": False} if lap else {})},
        "fp64": {"value": round(val64, 2), "unit": "Mpt-steps/s", "ms_per_step": round(ms64, 4)},
        "fp32_speedup_vs_fp64": round(ms64 / ms32, 3) if len(paths) == 2 else None,
        "fp32_vs_fp64": cmp,
        "precision_study": study,
        # compulsory stage bytes per step (STEP_BYTES_*) over the step time
        "hbm_tbs_algorithmic": {"f32": round(val32 * 1e6 * STEP_BYTES_MIXED / 1e12, 3),
                                "f64": round(val64 * 1e6 * STEP_BYTES_F64 / 1e12, 3),
                                "bytes_per_point": {"f32": STEP_BYTES_MIXED, "f64": STEP_BYTES_F64}},
        # rocprofv3 FETCH+WRITE bytes of one FP32-bulk step of this same workload (committed
        # counter summary) over this run's FP32 step time
        "hbm_tbs_measured_f32": ({"value": round(tr["total_GB_per_step"] / ms32, 3), "GB_per_step": tr["total_GB_per_step"],
                                  "source": tr["source"]} if tr and "f32" in paths else None),
        "roofline": roof32,
    }
    if full:
        out["roofline_f64"] = roofline(tab64)
        out["phases_ms_per_step"] = {p: ({k: round(v[0] / prof_steps, 4) for k, v in t.items()} if t else None)
                                     for p, t in (("f32", tab32), ("f64", tab64))}
        out["kernel_families"] = {p: (family_rows(t) if t else None) for p, t in (("f32", tab32), ("f64", tab64))}
    sol.close()
    del y0
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
