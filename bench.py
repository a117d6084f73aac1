#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the level step on one MI355X, FP32 bulk against FP64.

Workload (the precision configuration the metric is quoted on): one level of 8192 x 4096 cells
(33.5 M), dx = 12.5 m, cut into boxes of 32 x 32 (the reference's default max_grid_size;
--max-grid-size 0 runs it as one box); regularized rate-and-state fault with aging-law state
evolution, seeded per-face a, L and normal stress and a 25 MPa Gaussian nucleation bump;
Drucker-Prager viscoplasticity (mu 0.6, beta 0.3, relaxation time 0.03 s, cohesion 1 MPa);
PPM normal predictor with primitive limiting; absorbing outer boundaries.  A cell-update is one
valid cell advanced one level step; ghosts are not counted.  Inputs are resident in HBM (set on
the device) before the timed region starts.

`value` is the FP32-bulk rate (bulk update and plastic return in FP32; fault solve, state
exponential, slip and time in FP64); the FP64 rate and the ratio are reported beside it.  The
accuracy leg runs both precisions from the same start for --acc-steps steps (4000 = 7.5 s
simulated, the rupture has spread over about 9 km) and reports, at 64 seeded fault stations
inside the ruptured region, the error of FP32 bulk against FP64 in final slip, peak slip rate
and rupture arrival time (in cell-crossing times dx/cs).  A secondary leg times the plain FP64
drop-in configuration (4096 x 2048, dx = 25 m, no plasticity).  The CPU baseline is the oracle (a
scalar C restatement of the reference) on a bounded sample of the same workload, on this host.

Printed as one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tetemoko_amd import capi, scenarios, serving  # noqa: E402

HBM_PEAK_GBS = 8000.0
G = scenarios.GHOST
METRIC = "cell-updates/s/GPU, FP32-bulk vs FP64 at ≤1e-4 slip err; scenario p50/p99 latency"


def stored_bytes(plastic, tsize):
    """Bytes one cell-update moves in the stored layout the update kernel touches: per side it
    reads vx vy sxx syy sxy szz (+ gam with plasticity) and writes them (+ lam and gam), both
    sides; lam/gam in the bulk type."""
    per_side = 12 + (3 if plastic else 0)
    return per_side * 2 * tsize


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=4096)
    ap.add_argument("--dx", type=float, default=12.5)
    ap.add_argument("--shape-fp32", type=str, default="",
                    help="update shape of the FP32-bulk levels, march<rows> or tile<bx>x<by> (default: autotuned)")
    ap.add_argument("--shape-fp64", type=str, default="",
                    help="update shape of the FP64 levels, march<rows> or tile<bx>x<by> (default: autotuned)")
    ap.add_argument("--max-grid-size", type=int, default=32,
                    help="boxes of at most this many cells per side (the reference's default max_grid_size "
                         "is 32); 0 = the level as one box")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default update shape instead of autotuning it per level")
    ap.add_argument("--acc-steps", type=int, default=4000, help="steps of the accuracy leg (0: skip it)")
    ap.add_argument("--no-c1", action="store_true", help="skip the secondary 4096x2048 FP64 drop-in leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-dropin", action="store_true", help="skip the per-patch drop-in leg")
    ap.add_argument("--serving-n", type=int, default=120,
                    help="scenarios of the serving leg (the C4 stream at 8/s; 0: skip it)")
    return ap.parse_args()


def measured_traffic(workload, precision, shape):
    """HBM bytes per launch of the interior update kernel from the committed counter runs
    (profiles/traffic_*.json: FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    copy-kernel calibration), if one was taken on this workload, precision and update shape
    ("march32", "tile32x16", ...)."""
    d = os.path.join(ROOT, "profiles")
    for name in sorted(os.listdir(d)) if os.path.isdir(d) else []:
        if not (name.startswith("traffic_") and name.endswith(".json")):
            continue
        with open(os.path.join(d, name)) as f:
            t = json.load(f)
        if t.get("workload") == workload and t.get("precision") == precision and t.get("shape") == shape:
            return t.get("bytes_per_launch"), "profiles/" + name
    return None, None


def level_boxes(nx, ny, max_grid_size):
    if max_grid_size <= 0:
        return [(0, 0, nx - 1, ny - 1)]
    return scenarios.split_boxes(nx, ny, max_grid_size)


def parse_shape(text):
    """"march32" -> (32, 0, 0); "tile64x16" -> (0, 64, 16); "" -> None"""
    if not text:
        return None
    if text.startswith("march"):
        return int(text[5:]), 0, 0
    if text.startswith("tile"):
        bx, by = text[4:].split("x")
        return 0, int(bx), int(by)
    raise SystemExit("bench.py: update shape %r is neither march<rows> nor tile<bx>x<by>" % text)


SHAPES = {}  # precision -> fixed update shape (march_rows, bx, by); absent: autotuned


def make_level(cfg, precision, boxes):
    """A device-resident level holding the configuration's initial state (uniform background,
    seeded fault data), set on the device box by box (no host copy of the state)."""
    ctx = capi.Context(cfg, precision)
    lev = capi.Level(ctx, boxes)
    if precision in SHAPES:
        lev.set_update_shape(*SHAPES[precision])
    B = scenarios.fault_arrays(cfg)
    bg = scenarios.background_vector(cfg)
    for k, b in enumerate(boxes):
        lev.fill(k, bg, scenarios.box_bview(B, b) if b[1] == 0 else None)
    return ctx, lev


TUNE = True  # autotune each level's update shape before it is timed


def timed_leg(cfg, precision, boxes, warmup, steps):
    """The level's update shape autotuned from its initial state (tmk_level_autotune: every
    strip height and tile shape timed, then the fastest with the other workgroup order, the level
    returned to that state), then W untimed
    graph-replayed steps, then exactly K graph-replayed steps bracketed by HIP events on the
    level's stream (after a device synchronisation), then per-launch durations of the update
    kernels inside a replayed graph (stamps on each kernel's stream)."""
    ctx, lev = make_level(cfg, precision, boxes)
    dt = scenarios.dt_of(cfg)
    tuning = lev.autotune(0.0, dt, nsteps=10) if TUNE and precision not in SHAPES else None
    shape = lev.update_shape()
    t = 0.0
    lev.run(warmup, t, dt)
    t += warmup * dt
    lev.sync()
    ms = lev.run_timed(steps, t, dt)
    lev.sync()
    t += steps * dt
    prof = lev.profile_kernels(max(10, min(steps, 50)), t, dt)
    lev.sync()
    lev.close()
    ctx.close()
    ncell = cfg["nx"] * cfg["ny"]
    prof["march_rows"], prof["bx"], prof["by"] = shape
    return {"ms_per_step": ms / steps, "value": ncell * steps / (ms * 1e-3), "profile": prof, "shape": shape,
            "tuning": tuning}


def shape_record(r):
    """The update shape a timed leg ran, and what the tuner measured for each candidate."""
    rows, bx, by = r["shape"]
    out = {"march_rows": rows, "tile": [bx, by], "autotuned": r["tuning"] is not None}
    if r["tuning"] is not None:
        out["workgroup_order"] = "xcd bands" if r["tuning"]["xcd"] else "row-major"
        out["tried_ms_per_step"] = {(("march%d" % t[0]) if t[0] else ("tile%dx%d" % (t[1], t[2]))) +
                                    ("" if t[3] else "-rowmajor"): round(t[4], 4) for t in r["tuning"]["tried"]}
    return out


def roofline(prof, plastic, tsize, workload, precision):
    bpc = stored_bytes(plastic, tsize)
    achieved = prof["cells_interior"] * bpc / (prof["ms_interior"] * 1e-3) / 1e9
    rows, bx, by = prof["march_rows"], prof["bx"], prof["by"]
    shape = ("march%d" % rows) if rows else ("tile%dx%d" % (bx, by))
    traffic, src = measured_traffic(workload, precision, shape)
    kernel = ("k_godunov_march (interior strips of 58 x %d)" % rows) if rows else \
        ("k_godunov_interior (interior tiles %dx%d)" % (bx, by))
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": src,
            "kernel": kernel, "kernel_ms": prof["ms_interior"],
            "kernel_ms_timed_in_graph": bool(prof["in_graph"]),
            "cells_per_launch": prof["cells_interior"], "bytes_per_cell_update": bpc,
            "bytes_per_cell_update_22comp": 22 * 2 * tsize,
            "frac_22comp": prof["cells_interior"] * 22 * 2 * tsize / (prof["ms_interior"] * 1e-3) / 1e9
                           / HBM_PEAK_GBS,
            "boundary_kernel_ms": prof["ms_boundary"], "boundary_cells": prof["cells_boundary"],
            "step_ms_in_profile": prof["ms_step"]}


def station_faces(cfg, n=64, half_width_m=4500.0, seed=1234):
    """n distinct seeded fault faces within +-half_width_m of the nucleation centre; each
    station sits on its face centre, x = (i + 1/2) dx, where the reference's station
    interpolation (weight in metres, AMRLevelLinElast.cpp:262-270) returns the face value."""
    dx = cfg["dx"]
    xc = cfg["rs"]["nuc_x"]
    lo, hi = int(np.ceil((xc - half_width_m) / dx)), int(np.floor((xc + half_width_m) / dx))
    faces = np.sort(np.random.default_rng(seed).choice(np.arange(lo, hi), size=n, replace=False))
    return faces, (faces + 0.5) * dx


def accuracy_leg(cfg, precision, boxes, nsteps, xs, shape, chunk=500):
    """nsteps steps from t = 0, in the update shape the timed leg of this precision ran, with
    fault stations recorded on the device after every step: final slip, peak slip rate (over
    every step) and rupture time at each station."""
    ctx, lev = make_level(cfg, precision, boxes)
    lev.set_update_shape(*shape)
    st = capi.Stations(lev, xs, capacity=chunk)
    dt = scenarios.dt_of(cfg)
    peak = np.zeros(len(xs))
    last = None
    t, done = 0.0, 0
    t0 = time.perf_counter()
    while done < nsteps:
        n = min(chunk, nsteps - done)
        lev.run(n, t, dt)
        for _ in range(n):
            t = t + dt
        done += n
        rec = st.read()
        assert rec.shape[0] == n
        peak = np.maximum(peak, np.max(rec[:, :, 1 + 2], axis=0))
        last = rec[-1]
    lev.sync()
    wall = time.perf_counter() - t0
    st.close()
    lev.close()
    ctx.close()
    return {"slip": last[:, 1 + 8], "peak": peak, "rupture": last[:, 1 + 9], "wall_s": wall, "time": last[0, 0]}


def host_cpu():
    """Core count and model of the host this runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def drop_in_patch_rate(precision, seconds_hint=3.0):
    """The minimal drop-in: LinElastPatchGodunov::updateState replaced by one
    tmk_patch_update_state call per box on host FArrayBox data (host -> device -> host each
    call), over a 512x256 level cut into 32x32 boxes as the reference's default max_grid_size.
    PCIe-inclusive; reported beside the resident level path, never as its value."""
    cfg = scenarios.config_c1(seed=0, nx=512, ny=256, dx=25.0)
    U, B = scenarios.level_arrays(cfg)
    boxes = scenarios.split_boxes(cfg["nx"], cfg["ny"], 32)
    views = [(b, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None) for b in boxes]
    ctx = capi.Context(cfg, precision)
    dt = scenarios.dt_of(cfg)

    def one_step():
        for b, ub, bb in views:
            ubox = (b[0] - G, b[1] - G, b[2] + G, b[3] + G)
            bbox = (b[0] - G, 0, b[2] + G, 0) if bb is not None else None
            ctx.update_state(b, ub, ubox, bb, bbox, time=0.0, dt=dt)

    one_step()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds_hint:
        one_step()
        n += 1
    el = time.perf_counter() - t0
    ctx.close()
    return {"value": cfg["nx"] * cfg["ny"] * n / el, "unit": "cell-updates/s",
            "sample": "512x256 level, 128 boxes of 32x32, one tmk_patch_update_state call per box on host "
                      "arrays, %d level steps in %.1f s" % (n, el)}


def cpu_baseline(cfg_full, nthreads, seconds_hint):
    """Oracle (scalar C restatement) on a bounded sample: 512 x 256 cells of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tmo
    cfg = scenarios.config_c3(seed=0, nx=512, ny=256, dx=cfg_full["dx"])
    U, B = scenarios.level_arrays(cfg)
    op = tmo.OracleParams(cfg)
    dt = scenarios.dt_of(cfg)
    # calibrate the number of steps so the sample takes roughly seconds_hint
    t0 = time.perf_counter()
    assert tmo.level_steps(op, U, B, 0.0, dt, 1, nthreads=nthreads, maxbox=32) == 0
    one = time.perf_counter() - t0
    nsteps = max(2, min(200, int(seconds_hint / max(one, 1e-3))))
    t0 = time.perf_counter()
    assert tmo.level_steps(op, U, B, dt, dt, nsteps, nthreads=nthreads, maxbox=32) == 0
    el = time.perf_counter() - t0
    ncpu, model = host_cpu()
    return {"value": cfg["nx"] * cfg["ny"] * nsteps / el, "unit": "cell-updates/s", "cores": nthreads,
            "kind": "port",
            "sample": "512x256 cells of the same rate-and-state + Drucker-Prager workload (dx %.1f m), %d level "
                      "steps, boxes of 32x32, FP64, oracle C restatement (gcc -O3, no FMA contraction), %d "
                      "thread(s), %.1f s; host: %d logical CPUs, %s" % (cfg["dx"], nsteps, nthreads, el, ncpu,
                                                                          model)}


def serving_leg(n, rate=8.0):
    """The first n scenarios of the serving stream (grids 256x128..2048x1024, slip weakening or
    rate-and-state, 0-2 refined levels, 1000-4000 coarse steps, FP32 bulk), arriving as a Poisson
    process at `rate`, served with the loop's tuned update shapes: time to solution p50 / p99
    (completion - arrival) and throughput."""
    print("bench: serving leg, %d scenarios at %g/s" % (n, rate), file=sys.stderr, flush=True)
    sl = serving.draw_scenarios(n, 7, max_levels=2)
    shapes = serving.tune_shapes(sl, capi.PREC_FP32_BULK)
    _, st = serving.serve(sl, rate, precision=capi.PREC_FP32_BULK, shapes=shapes)
    return {"scenarios": st["scenarios"], "failed": st["failed"], "arrival_rate_hz": rate,
            "p50_time_to_solution_s": st["p50_latency_s"], "p99_time_to_solution_s": st["p99_latency_s"],
            "scenarios_per_s": st["scenarios_per_s"], "cell_updates_per_s": st["cell_updates_per_s"],
            "wall_s": st["wall_s"], "refined_levels": "0-2 per scenario",
            "update_shapes": {"coarsest_rows": shapes["single"], "refined_rows": shapes["refined"]},
            "stream": "python -m tetemoko_amd.serving --n %d --rate %g --steps-scale 1.0 (its first scenarios)"
                      % (n, rate)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only; every box here has a single MI355X. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if not capi.lib().tmk_device_ok():
        print("bench.py: no gfx950 device visible", file=sys.stderr)
        sys.exit(3)
    global TUNE
    TUNE = not a.no_tune
    for prec, text in ((capi.PREC_FP32_BULK, a.shape_fp32), (capi.PREC_FP64, a.shape_fp64)):
        if parse_shape(text):
            SHAPES[prec] = parse_shape(text)
    cfg = scenarios.config_c3(seed=0, nx=a.nx, ny=a.ny, dx=a.dx)
    boxes = level_boxes(cfg["nx"], cfg["ny"], a.max_grid_size)
    wl = "rs-aging+drucker-prager %dx%d dx=%g" % (cfg["nx"], cfg["ny"], cfg["dx"])
    print("bench: %s, %d boxes, FP32 bulk, warmup %d, steps %d" % (wl, len(boxes), a.warmup, a.steps),
          file=sys.stderr, flush=True)
    r32 = timed_leg(cfg, capi.PREC_FP32_BULK, boxes, a.warmup, a.steps)
    print("bench: FP64, warmup %d, steps %d" % (a.warmup, a.steps), file=sys.stderr, flush=True)
    r64 = timed_leg(cfg, capi.PREC_FP64, boxes, a.warmup, a.steps)
    out = {
        "metric": METRIC,
        "value": r32["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk (f64 fault solve, state exponential, slip, time)",
        "data": "synthetic (seeded per-face friction parameters and normal stress, background prestress)",
        "config": {"workload": "rate-and-state (aging law) + Drucker-Prager viscoplasticity, single level "
                               "%dx%d (%.1f M cells) in %d boxes, dx=%g m, PPM, absorbing boundaries, FP32 bulk "
                               "with FP64 fault; FP64 beside it" % (cfg["nx"], cfg["ny"], cfg["nx"] * cfg["ny"] / 1e6,
                                                                     len(boxes), cfg["dx"]),
                   "nx": cfg["nx"], "ny": cfg["ny"], "boxes": len(boxes), "max_grid_size": a.max_grid_size,
                   "dx_m": cfg["dx"], "dt_s": scenarios.dt_of(cfg), "normal_predictor": "PPM",
                   "plasticity": {"mu": cfg["p_mu"], "beta": cfg["p_beta"], "eta_MPa_s": cfg["p_eta"],
                                  "cohesion_MPa": cfg["p_b"]},
                   "update_shape_fp32_bulk": shape_record(r32), "update_shape_fp64": shape_record(r64)},
        "roofline": roofline(r32["profile"], True, 4, wl, "fp32_bulk"),
        "fp64": {"value": r64["value"], "ms_per_step": r64["ms_per_step"],
                 "roofline": roofline(r64["profile"], True, 8, wl, "fp64")},
        "speedup_fp32_bulk_vs_fp64": r64["ms_per_step"] / r32["ms_per_step"],
    }
    if a.acc_steps > 0:
        faces, xs = station_faces(cfg)
        print("bench: accuracy leg, %d steps per precision, 64 stations" % a.acc_steps, file=sys.stderr, flush=True)
        s64 = accuracy_leg(cfg, capi.PREC_FP64, boxes, a.acc_steps, xs, r64["shape"])
        s32 = accuracy_leg(cfg, capi.PREC_FP32_BULK, boxes, a.acc_steps, xs, r32["shape"])
        r64t, r32t = s64["rupture"], s32["rupture"]
        both = (r64t < 1e9) & (r32t < 1e9)
        cells = np.abs(r64t[both] - r32t[both]) * cfg["cs"] / cfg["dx"]
        slip_err = float(np.max(np.abs(s32["slip"] - s64["slip"])) / max(np.max(np.abs(s64["slip"])), 1e-300))
        arrival = float(np.max(cells)) if cells.size else None
        out["accuracy"] = {
            "steps": a.acc_steps, "simulated_s": s64["time"], "stations": len(xs),
            "station_faces": [int(faces[0]), int(faces[-1])],
            "slip_rel_err_max": slip_err,
            "peak_slip_rate_rel_err_max": float(np.max(np.abs(s32["peak"] - s64["peak"])) /
                                                max(np.max(np.abs(s64["peak"])), 1e-300)),
            "rupture_arrival_err_cells_max": arrival,
            "ruptured_stations_fp64": int(np.sum(r64t < 1e9)), "ruptured_stations_fp32": int(np.sum(r32t < 1e9)),
            "max_slip_fp64_m": float(np.max(s64["slip"])), "max_peak_slip_rate_fp64": float(np.max(s64["peak"])),
            "slip_err_le_1e-4": slip_err <= 1e-4,
            "arrival_err_le_1_cell": arrival is not None and arrival <= 1.0 and
                                     int(np.sum(r64t < 1e9)) == int(np.sum(r32t < 1e9)),
            "wall_s_fp64": s64["wall_s"], "wall_s_fp32": s32["wall_s"],
        }
    if not a.no_c1:
        c1 = scenarios.config_c1(seed=0, nx=4096, ny=2048, dx=25.0)
        print("bench: secondary leg, 4096x2048 FP64 drop-in configuration", file=sys.stderr, flush=True)
        b1 = level_boxes(c1["nx"], c1["ny"], a.max_grid_size)
        wl1 = "rs-aging 4096x2048 dx=25"
        q64 = timed_leg(c1, capi.PREC_FP64, b1, a.warmup, a.steps)
        q32 = timed_leg(c1, capi.PREC_FP32_BULK, b1, a.warmup, a.steps)
        out["secondary_4096x2048_rs"] = {
            "workload": "rate-and-state (aging law), single level 4096x2048 in %d boxes, dx=25 m, PPM, no "
                        "plasticity" % len(b1),
            "fp64": {"value": q64["value"], "ms_per_step": q64["ms_per_step"], "update_shape": shape_record(q64),
                     "roofline": roofline(q64["profile"], False, 8, wl1, "fp64")},
            "fp32_bulk": {"value": q32["value"], "ms_per_step": q32["ms_per_step"],
                          "update_shape": shape_record(q32),
                          "roofline": roofline(q32["profile"], False, 4, wl1, "fp32_bulk")},
        }
    if not a.no_dropin:
        print("bench: per-patch drop-in leg", file=sys.stderr, flush=True)
        out["drop_in_patch_fp64"] = drop_in_patch_rate(capi.PREC_FP64)
    if a.serving_n > 0:
        out["serving"] = serving_leg(a.serving_n)
    if not a.no_cpu:
        print("bench: CPU baseline leg", file=sys.stderr, flush=True)
        cb = cpu_baseline(cfg, 1, 12.0)
        nthr = min(16, os.cpu_count() or 1)
        cb_all = cpu_baseline(cfg, nthr, 6.0)
        cb["all_threads"] = {"value": cb_all["value"], "cores": nthr, "sample": cb_all["sample"]}
        cb["gpu_fp32_bulk_over_cpu_1_thread"] = out["value"] / cb["value"]
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
