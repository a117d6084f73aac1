"""FP32-bulk versus FP64: error of final slip, peak slip rate and rupture time.

Runs the rate-and-state + Drucker-Prager workload (8192 x 4096 cells, dx = 12.5 m
at full size) twice from the same seeded initial state: once with FP64
everywhere, once with FP32 bulk arithmetic and FP64 fault/state/slip/time.
The errors are read at 64 fault stations spaced evenly over +-span/2 around the nucleation
centre (the part of the fault the rupture reaches in the simulated time):
  * final slip (relative to the FP64 slip at the station);
  * peak slip rate (the largest slip rate seen at the station, sampled every
    `sample` steps);
  * rupture-front arrival time (in cells: |dt_arrival| * cs / dx).
"""
import argparse
import json
import time

import numpy as np

from . import capi, scenarios

G = scenarios.GHOST


def run(cfg, precision, nsteps, sample, stations, tile=None):
    ctx = capi.Context(cfg, precision)
    if tile:
        ctx.set_tile(*tile)
    nx, ny = cfg["nx"], cfg["ny"]
    lev = capi.Level(ctx, [(0, 0, nx - 1, ny - 1)])
    B = scenarios.fault_arrays(cfg)
    lev.fill(0, scenarios.background_vector(cfg), B)
    dt = scenarios.dt_of(cfg)
    peak = np.zeros(len(stations))
    Bb = np.empty_like(B)
    t = 0.0
    done = 0
    t0 = time.perf_counter()
    while done < nsteps:
        n = min(sample, nsteps - done)
        lev.run(n, t, dt)
        for _ in range(n):
            t = t + dt
        done += n
        lev.sync()
        lev.download_fault(0, Bb)
        peak = np.maximum(peak, Bb[2, G + stations])
    wall = time.perf_counter() - t0
    lev.close()
    ctx.close()
    return Bb, peak, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=4096)
    ap.add_argument("--dx", type=float, default=12.5)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--span-km", type=float, default=30.0)
    ap.add_argument("--sample", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    cfg = scenarios.config_c3(seed=a.seed, nx=a.nx, ny=a.ny, dx=a.dx)
    xc = cfg["rs"]["nuc_x"] / a.dx
    half = 0.5 * a.span_km * 1000.0 / a.dx
    stations = np.clip(np.round(np.linspace(xc - half, xc + half, 64)).astype(int), 0, a.nx - 1)
    B64, pk64, w64 = run(cfg, capi.PREC_FP64, a.steps, a.sample, stations)
    B32, pk32, w32 = run(cfg, capi.PREC_FP32_BULK, a.steps, a.sample, stations)
    s64, s32 = B64[8, G + stations], B32[8, G + stations]
    rt64, rt32 = B64[9, G + stations], B32[9, G + stations]
    both = (rt64 < 1e9) & (rt32 < 1e9)
    cells = np.abs(rt64[both] - rt32[both]) * cfg["cs"] / cfg["dx"]
    out = {
        "workload": "rate-and-state + Drucker-Prager %dx%d dx=%.1f m, %d steps (%.2f s)" % (
            a.nx, a.ny, a.dx, a.steps, a.steps * scenarios.dt_of(cfg)),
        "stations_span_km": a.span_km,
        "stations": 64,
        "slip_rel_err_max": float(np.max(np.abs(s32 - s64)) / max(np.max(np.abs(s64)), 1e-300)),
        "peak_slip_rate_rel_err_max": float(np.max(np.abs(pk32 - pk64)) / max(np.max(np.abs(pk64)), 1e-300)),
        "ruptured_stations_fp64": int(np.sum(rt64 < 1e9)), "ruptured_stations_fp32": int(np.sum(rt32 < 1e9)),
        "ruptured_faces_fp64": int(np.sum(B64[9, G:-G] < 1e9)), "ruptured_faces_fp32": int(np.sum(B32[9, G:-G] < 1e9)),
        "max_slip_fp64_m": float(np.max(B64[8, G:-G])), "max_slip_rate_fp64": float(np.max(pk64)),
        "rupture_time_err_cells_max": float(np.max(cells)) if cells.size else None,
        "fp64_wall_s": w64, "fp32_wall_s": w32,
        "cell_updates": a.nx * a.ny * a.steps,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
