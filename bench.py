#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the level step on one MI355X.

Workload (the plain FP64 drop-in configuration): one level of 4096 x 2048 cells cut
into 8192 boxes of 32 x 32 (the reference's default max_grid_size; --max-grid-size 0
runs it as one box), dx = 25 m, regularized rate-and-state fault with aging-law state
evolution and
seeded per-face a, L and normal stress, PPM normal predictor with primitive
limiting, absorbing outer boundaries.  A cell-update is one valid cell advanced
one level step; ghosts are not counted.  Inputs are resident in HBM before the
timed region starts.

Printed as one JSON line.  The same workload is also run with FP32 bulk
arithmetic (fault solve, state, slip and time in FP64) and its slip error
against the FP64 run is reported beside it.  The CPU baseline is the oracle
(a scalar C restatement of the reference) on a bounded sample of the same
workload, run on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tetemoko_amd import capi, scenarios  # noqa: E402

HBM_PEAK_GBS = 8000.0
# bytes one cell-update must move with the state layout the kernel touches:
# mode-II stencil planes (vx vy sxx syy sxy) + szz, read and written, both sides
LIVE_COMPONENTS = 12
ALL_COMPONENTS = 22


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--tile", type=str, default="")
    ap.add_argument("--max-grid-size", type=int, default=32,
                    help="boxes of at most this many cells per side (the reference's default max_grid_size "
                         "is 32); 0 = the level as one box")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-bulk leg")
    return ap.parse_args()


def measured_traffic(nx, ny, tile):
    """HBM bytes per launch of the FP64 update kernel from the committed counter runs
    (profiles/traffic_k_godunov_fp64.json: FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    the copy-kernel calibration), if they were taken on this workload and tile shape."""
    path = os.path.join(ROOT, "profiles", "traffic_k_godunov_fp64.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("nx") != nx or t.get("ny") != ny or t.get("tile") != (list(tile) if tile else "default"):
        return None
    return t.get("bytes_per_launch")


def level_boxes(cfg, max_grid_size):
    if max_grid_size <= 0:
        return [(0, 0, cfg["nx"] - 1, cfg["ny"] - 1)]
    return scenarios.split_boxes(cfg["nx"], cfg["ny"], max_grid_size)


def run_level(cfg, U, B, precision, warmup, steps, tile, boxes):
    ctx = capi.Context(cfg, precision)
    if tile:
        ctx.set_tile(*tile)
    lev = capi.Level(ctx, boxes)
    for k, b in enumerate(boxes):
        lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
    dt = scenarios.dt_of(cfg)
    t = 0.0
    lev.run(warmup, t, dt)
    t += warmup * dt
    lev.sync()
    ms = lev.run_timed(steps, t, dt)
    lev.sync()
    t += steps * dt
    Uo, Bo = np.full_like(U, np.nan), np.full_like(B, np.nan)
    g = scenarios.GHOST
    for k, b in enumerate(boxes):
        x0, y0, x1, y1 = b
        ub = np.empty((U.shape[0], y1 - y0 + 1 + 2 * g, x1 - x0 + 1 + 2 * g))
        bb = np.empty((B.shape[0], x1 - x0 + 1 + 2 * g)) if y0 == 0 else None
        lev.download(k, ub, bb)
        Uo[:, y0 + g:y1 + g + 1, x0 + g:x1 + g + 1] = ub[:, g:-g, g:-g]
        if bb is not None:
            Bo[:, x0 + g:x1 + g + 1] = bb[:, g:-g]
    # per-launch durations of the update kernels (eager steps; interior tiles on the level's
    # stream, boundary tiles concurrently on its side stream), each from HIP events on the
    # stream it runs on; continues the same simulation
    kms = lev.profile_kernels(max(10, steps // 4), t, dt)
    lev.sync()
    lev.close()
    ctx.close()
    return ms, kms, Uo, Bo


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
    g = scenarios.GHOST
    views = [(b, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None) for b in boxes]
    ctx = capi.Context(cfg, precision)
    dt = scenarios.dt_of(cfg)

    def one_step():
        for b, ub, bb in views:
            ubox = (b[0] - g, b[1] - g, b[2] + g, b[3] + g)
            bbox = (b[0] - g, 0, b[2] + g, 0) if bb is not None else None
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
    cfg = scenarios.config_c1(seed=0, nx=512, ny=256, dx=25.0)
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
            "sample": "512x256 cells of the same rate-and-state workload, %d level steps, boxes of 32x32, "
                      "FP64, oracle C restatement (gcc -O3, no FMA contraction), %d thread(s), %.1f s; "
                      "host: %d logical CPUs, %s" % (nsteps, nthreads, el, ncpu, model)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only; every box here has a single MI355X. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if not capi.lib().tmk_device_ok():
        print("bench.py: no gfx950 device visible", file=sys.stderr)
        sys.exit(3)
    tile = tuple(int(x) for x in a.tile.split("x")) if a.tile else None
    cfg = scenarios.config_c1(seed=0, nx=a.nx, ny=a.ny, dx=25.0)
    U0, B0 = scenarios.level_arrays(cfg)
    ncell = cfg["nx"] * cfg["ny"]
    print("bench: %dx%d FP64, warmup %d, steps %d" % (cfg["nx"], cfg["ny"], a.warmup, a.steps),
          file=sys.stderr, flush=True)
    boxes = level_boxes(cfg, a.max_grid_size)
    ms64, kms64, U64, B64 = run_level(cfg, U0, B0, capi.PREC_FP64, a.warmup, a.steps, tile, boxes)
    value = ncell * a.steps / (ms64 * 1e-3)
    # dominant kernel: the interior-tile update; its algorithmic bytes are the live state planes
    # of the cells it updates, read once and written once
    bytes_live = kms64["cells_interior"] * LIVE_COMPONENTS * 2 * 8
    achieved = bytes_live / (kms64["ms_interior"] * 1e-3) / 1e9
    out = {
        "metric": "cell-updates/s/GPU (FP64 plain level step; FP32-bulk reported beside it)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms64 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded per-face friction parameters, background prestress)",
        "config": {"workload": "rate-and-state, aging law, single level %dx%d in %d boxes, dx=25 m, PPM, "
                               "absorbing boundaries, FP64" % (cfg["nx"], cfg["ny"], len(boxes)),
                   "nx": cfg["nx"], "ny": cfg["ny"], "boxes": len(boxes), "max_grid_size": a.max_grid_size,
                   "dx_m": 25.0, "dt_s": scenarios.dt_of(cfg),
                   "normal_predictor": "PPM", "tile": list(tile) if tile else "default"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(cfg["nx"], cfg["ny"], tile),
                     "kernel": "k_godunov_interior (interior tiles)",
                     "kernel_ms": kms64["ms_interior"],
                     "cells_per_launch": kms64["cells_interior"],
                     "bytes_per_cell_update": LIVE_COMPONENTS * 2 * 8,
                     "bytes_per_cell_update_22comp": ALL_COMPONENTS * 2 * 8,
                     "frac_22comp": kms64["cells_interior"] * ALL_COMPONENTS * 2 * 8 /
                                    (kms64["ms_interior"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "boundary_kernel_ms": kms64["ms_boundary"], "boundary_cells": kms64["cells_boundary"],
                     "step_ms_eager": kms64["ms_step"]},
    }
    if not a.no_fp32:
        print("bench: FP32-bulk leg", file=sys.stderr, flush=True)
        ms32, kms32, U32, B32 = run_level(cfg, U0, B0, capi.PREC_FP32_BULK, a.warmup, a.steps, tile, boxes)
        g = scenarios.GHOST
        slip64, slip32 = B64[8, g:-g], B32[8, g:-g]
        V64, V32 = B64[2, g:-g], B32[2, g:-g]
        out["fp32_bulk"] = {
            "value": ncell * a.steps / (ms32 * 1e-3),
            "ms_per_step": ms32 / a.steps,
            "kernel_ms": kms32["ms_interior"],
            "boundary_kernel_ms": kms32["ms_boundary"],
            "speedup_vs_fp64": ms64 / ms32,
            "roofline_frac": kms32["cells_interior"] * LIVE_COMPONENTS * 2 * 4 / (kms32["ms_interior"] * 1e-3) / 1e9
                             / HBM_PEAK_GBS,
            "slip_rel_err": float(np.max(np.abs(slip32 - slip64)) / max(np.max(np.abs(slip64)), 1e-300)),
            "slip_rate_rel_err": float(np.max(np.abs(V32 - V64)) / max(np.max(np.abs(V64)), 1e-300)),
        }
    print("bench: per-patch drop-in leg", file=sys.stderr, flush=True)
    out["drop_in_patch_fp64"] = drop_in_patch_rate(capi.PREC_FP64)
    if not a.no_cpu:
        print("bench: CPU baseline leg", file=sys.stderr, flush=True)
        cb = cpu_baseline(cfg, 1, 12.0)
        nthr = min(16, os.cpu_count() or 1)
        cb_all = cpu_baseline(cfg, nthr, 6.0)
        cb["all_threads"] = {"value": cb_all["value"], "cores": nthr, "sample": cb_all["sample"]}
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
