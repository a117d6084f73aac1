"""Continuous-batching serving loop for rupture scenarios on one MI355X.

Rupture scenarios arrive as a Poisson process. Each scenario has its own grid size,
friction law and parameters, nucleation position and step count. A scenario is
admitted when the free device memory covers its footprint. Its level data then
stays resident until it finishes. Every admitted scenario's boxes go into the same
fused launch as the others (tmk_batch). A scenario that arrives while others are
running joins at the next chunk boundary. A finished one leaves, its fault data is
read back, and its memory is freed for the queue.

Reported: scenarios/s, cell-updates/s and p50/p99 time-to-solution
(completion - arrival, wall clock).

The reference runs one scenario per process with no serving mode
(amrGodunov.cpp:81-166). Each result is bit-identical alone and in the mix, because
nothing couples the members of a batch.
"""
import argparse
import json
import math
import time

import numpy as np

from . import capi, scenarios

GHOST = scenarios.GHOST


def draw_scenarios(n, seed, nx_range=(256, 2048), steps_range=(1000, 4000), steps_scale=1.0):
    """Seeded scenario list.

    Grids run from 256x128 to 2048x1024, with the aspect ratio 2:1. Friction is
    slip weakening (config_c0 distributions) or rate-and-state (config_c1
    distributions), drawn per scenario. The nucleation position is random.
    """
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        lg = rng.uniform(math.log2(nx_range[0]), math.log2(nx_range[1]))
        nx = int(2 ** lg) // 32 * 32
        nx = max(nx_range[0], min(nx_range[1], nx))
        ny = nx // 2
        kind = "sw" if rng.uniform() < 0.5 else "rs"
        # rate-and-state keeps the config_c1 spacing: with L ~ 0.02 m its explicit state
        # update needs the cohesive zone resolved (coarser grids drive psi to overflow)
        dx = float(rng.choice([25.0, 50.0, 100.0])) if kind == "sw" else 25.0
        if kind == "sw":
            cfg = scenarios.config_c0(seed=seed * 100003 + k, nx=nx, ny=ny, dx=dx)
            # move the nucleation patch: shift the per-face shear by a random offset
            shift = int(rng.integers(-nx // 4, nx // 4))
            cfg["faces"]["tau"] = np.roll(cfg["faces"]["tau"], shift)
        else:
            cfg = scenarios.config_c1(seed=seed * 100003 + k, nx=nx, ny=ny, dx=dx)
            cfg["rs"]["nuc_x"] = float(rng.uniform(0.25, 0.75) * nx * dx)
            cfg["rs"]["nuc_R"] = 1500.0 * dx / 25.0
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1) * steps_scale)
        out.append({"id": k, "cfg": cfg, "steps": max(steps, 2)})
    return out


class Running:
    __slots__ = ("sc", "ctx", "lev", "done_steps", "arrival", "admitted", "dt", "peak_V")

    def __init__(self, sc, ctx, lev, arrival, admitted, dt):
        self.sc, self.ctx, self.lev = sc, ctx, lev
        self.arrival, self.admitted, self.dt = arrival, admitted, dt
        self.done_steps = 0
        self.peak_V = None


def _start(sc, precision, tile):
    cfg = sc["cfg"]
    ctx = capi.Context(cfg, precision)
    if tile:
        ctx.set_tile(*tile)
    box = (0, 0, cfg["nx"] - 1, cfg["ny"] - 1)
    lev = capi.Level(ctx, [box])
    B = scenarios.fault_arrays(cfg)
    lev.fill(0, scenarios.background_vector(cfg), B)
    return ctx, lev


def _result(r, B):
    cfg = r.sc["cfg"]
    nb_slip = 8 if cfg["fault"] == "rs" else 7
    nb_rt = 9 if cfg["fault"] == "rs" else 8
    nb_V = 2 if cfg["fault"] == "rs" else 4
    slip = B[nb_slip, GHOST:-GHOST]
    rt = B[nb_rt, GHOST:-GHOST]
    V = B[nb_V, GHOST:-GHOST]
    ruptured = rt < 1e9
    return {"id": r.sc["id"], "fault": cfg["fault"], "nx": cfg["nx"], "ny": cfg["ny"], "steps": r.sc["steps"],
            "max_slip": float(np.max(slip)), "mean_slip": float(np.mean(slip)),
            "final_max_slip_rate": float(np.max(V)),
            "ruptured_fraction": float(np.mean(ruptured)),
            "first_rupture_time": float(np.min(rt[ruptured])) if ruptured.any() else None}


def serve(scen_list, rate_hz, precision=capi.PREC_FP32_BULK, chunk=50, mem_reserve=8 << 30, tile=None,
          log=None, keep_fault=False):
    """Run the arrival stream; returns (results, stats).

    The loop keeps the device busy while the host works: chunk k+1 of the shared
    launches is enqueued before the results of chunk k are handled, and new
    scenarios are set up (context, level allocation, initial upload) while a chunk
    runs.  Membership changes take effect at chunk boundaries.
    """
    rng = np.random.default_rng(12345)
    gaps = rng.exponential(1.0 / rate_hz, size=len(scen_list)) if rate_hz > 0 else np.zeros(len(scen_list))
    arrivals = np.cumsum(gaps) - gaps[0]
    pending = list(zip(arrivals, range(len(scen_list))))
    queue = []
    batch = capi.Batch()
    running = []      # members of the batch
    ready = []        # set up, joining the batch at the next chunk boundary
    results, latencies = [], []
    n_failed = 0
    cell_updates = 0
    footprint = {}    # (fault, nx, ny) -> device bytes of one scenario's level
    t0 = time.perf_counter()
    ip = 0
    phase = {"admit": 0.0, "launch": 0.0, "sync": 0.0, "retire": 0.0, "idle": 0.0}
    inflight = 0      # steps of the chunk in flight (0: none)

    def now():
        return time.perf_counter() - t0

    class _clock:  # host time spent per phase of the loop
        def __init__(self, name):
            self.name = name

        def __enter__(self):
            self.t = time.perf_counter()

        def __exit__(self, *exc):
            phase[self.name] += time.perf_counter() - self.t

    def admit():
        # admission by free device memory, in arrival order
        while queue:
            arr, k = queue[0]
            sc = scen_list[k]
            cfg = sc["cfg"]
            key = (cfg["fault"], cfg["nx"], cfg["ny"])
            ctx = None
            if key not in footprint:
                ctx = capi.Context(cfg, precision)
                footprint[key] = capi.level_footprint(ctx, [(0, 0, cfg["nx"] - 1, cfg["ny"] - 1)])
            free, _ = capi.mem_info()
            if footprint[key] + mem_reserve > free:
                if ctx is not None:
                    ctx.close()
                break
            queue.pop(0)
            if ctx is not None:
                ctx.close()
            ctx, lev = _start(sc, precision, tile)
            ready.append(Running(sc, ctx, lev, arr, now(), scenarios.dt_of(cfg)))

    def retire(r, B, failed_code=None):
        r.lev.close()
        r.ctx.close()
        running.remove(r)
        if failed_code is not None:
            return
        tdone = now()
        res = _result(r, B)
        res["arrival_s"] = float(r.arrival)
        res["latency_s"] = float(tdone - r.arrival)
        res["queue_s"] = float(r.admitted - r.arrival)
        if keep_fault:
            res["fault_data"] = B
        results.append(res)
        latencies.append(tdone - r.arrival)
        if log:
            cfg = r.sc["cfg"]
            log("done %d (%dx%d %s, %d steps) latency %.2fs running %d queued %d" %
                (res["id"], cfg["nx"], cfg["ny"], cfg["fault"], r.sc["steps"], res["latency_s"],
                 len(running), len(queue)))

    while ip < len(pending) or queue or running or ready or inflight:
        t = now()
        while ip < len(pending) and pending[ip][0] <= t:
            queue.append(pending[ip])
            ip += 1
        with _clock("admit"):
            admit()
        if inflight:
            failed = set()
            with _clock("sync"):
                try:
                    batch.sync()
                except capi.TmkError:
                    # a scenario's own error (e.g. a NaN state variable) retires only that scenario
                    failed = {id(r) for r in running if batch.member_error(r.lev) != 0}
            n = inflight
            inflight = 0
            for r in running:
                r.done_steps += n
                cell_updates += r.sc["cfg"]["nx"] * r.sc["cfg"]["ny"] * n
            with _clock("retire"):
                for r in [r for r in running if id(r) in failed]:
                    code = batch.member_error(r.lev)
                    batch.remove(r.lev)
                    res = {"id": r.sc["id"], "failed": capi.strerror(code), "arrival_s": float(r.arrival),
                           "latency_s": float(now() - r.arrival)}
                    results.append(res)
                    n_failed += 1
                    retire(r, None, code)
                    if log:
                        log("failed %d: %s" % (res["id"], res["failed"]))
                for r in [r for r in running if r.done_steps >= r.sc["steps"]]:
                    batch.remove(r.lev)
                    cfg = r.sc["cfg"]
                    B = np.empty((scenarios.NB[cfg["fault"]], cfg["nx"] + 2 * GHOST))
                    r.lev.download_fault(0, B)
                    retire(r, B)
        with _clock("launch"):
            for r in ready:
                batch.add(r.lev, 0.0, r.dt)
                running.append(r)
            ready.clear()
            if running:
                n = min(chunk, min(r.sc["steps"] - r.done_steps for r in running))
                batch.run(n)
                inflight = n
        if not running and not ready and not queue and ip < len(pending):
            with _clock("idle"):
                time.sleep(max(0.0, pending[ip][0] - now()))
    wall = now()
    batch.close()
    lat = np.array(latencies)
    stats = {"scenarios": len(results) - n_failed, "failed": n_failed, "wall_s": wall,
             "scenarios_per_s": (len(results) - n_failed) / wall,
             "cell_updates_per_s": cell_updates / wall,
             "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
             "mean_latency_s": float(np.mean(lat)), "arrival_rate_hz": rate_hz, "chunk_steps": chunk,
             "host_seconds": {k: round(v, 3) for k, v in phase.items()}}
    return results, stats


def main():
    ap = argparse.ArgumentParser(description="rupture-scenario serving loop (continuous batching)")
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--rate", type=float, default=4.0, help="Poisson arrival rate, scenarios/s")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--steps-scale", type=float, default=0.1,
                    help="scale of the 1000-4000 step range (1.0 = full length)")
    ap.add_argument("--nx-max", type=int, default=2048)
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--fp64", action="store_true")
    a = ap.parse_args()
    sl = draw_scenarios(a.n, a.seed, nx_range=(256, a.nx_max), steps_scale=a.steps_scale)
    prec = capi.PREC_FP64 if a.fp64 else capi.PREC_FP32_BULK
    res, stats = serve(sl, a.rate, precision=prec, chunk=a.chunk,
                       log=lambda m: print(m, flush=True))
    stats.update({"n": a.n, "steps_scale": a.steps_scale, "precision": "fp64" if a.fp64 else "fp32-bulk",
                  "grid_range": "256x128..%dx%d" % (a.nx_max, a.nx_max // 2)})
    print(json.dumps(stats), flush=True)


if __name__ == "__main__":
    main()
