"""Continuous-batching serving loop for rupture scenarios on one MI355X.

Rupture scenarios arrive as a Poisson process. Each scenario has its own grid size,
friction law and parameters, nucleation position and step count. A scenario is
admitted when the free device memory covers its footprint. Its level data then
stays resident until it finishes. Every admitted scenario's boxes go into the same
fused launch as the others (tmk_batch). A scenario that arrives while others are
running joins at the next chunk boundary. A finished one leaves, its fault data is
read back, and its memory is freed for the queue.

Scenarios with refined levels (0-2, ratio 2, around the nucleation) join a hierarchy
batch of their level count (tmk_amr_add): level l of every such scenario is updated in
the same launches, with the coarse-fine fills, registers and averages of all of them in
shared launches too, one graph per two coarse steps on the batch's own stream, beside the
single-level batch.

Reported: scenarios/s, cell-updates/s (every level's cells times its steps) and p50/p99
time-to-solution (completion - arrival, wall clock).

The reference runs one scenario per process with no serving mode
(amrGodunov.cpp:81-166). Each result is bit-identical alone and in the mix, because
nothing couples the members of a batch.
"""
import argparse
import json
import math
import time

import numpy as np

from . import amr, capi, scenarios

GHOST = scenarios.GHOST


AMR_RATIO = 2


def lens_boxes(nx, ny, xn, half_frac, height_frac, scale, align=8, maxbox=64):
    """Boxes (in the cells of a level refined by scale) of a region on the fault around coarse
    cell xn: x within +-half_frac * nx, y below height_frac * ny (coarse cells), edges on
    multiples of align fine cells, boxes of at most maxbox per side."""
    fx0 = max(0, int((xn - half_frac * nx) * scale) // align * align)
    fx1 = min(nx * scale, -(-int((xn + half_frac * nx) * scale) // align) * align)
    fy1 = min(ny * scale, -(-int(height_frac * ny * scale) // align) * align)
    boxes = []
    for y0 in range(0, fy1, maxbox):
        for x0 in range(fx0, fx1, maxbox):
            boxes.append((x0, y0, min(x0 + maxbox, fx1) - 1, min(y0 + maxbox, fy1) - 1))
    return boxes


def amr_layout(sc):
    """Box lists of a scenario with sc["levels"] refined levels (ratio AMR_RATIO): the base as one
    box, level 1 over +-1/8 of the grid width and 1/8 of its height around the nucleation, level 2
    over half of that."""
    cfg = sc["cfg"]
    nx, ny = cfg["nx"], cfg["ny"]
    out = [[(0, 0, nx - 1, ny - 1)]]
    if sc["levels"] >= 1:
        out.append(lens_boxes(nx, ny, sc["xn"], 1.0 / 8, 1.0 / 8, AMR_RATIO))
    if sc["levels"] >= 2:
        out.append(lens_boxes(nx, ny, sc["xn"], 1.0 / 16, 1.0 / 16, AMR_RATIO * AMR_RATIO))
    return out


def draw_scenarios(n, seed, nx_range=(256, 2048), steps_range=(1000, 4000), steps_scale=1.0, max_levels=2):
    """Seeded scenario list.

    Grids run from 256x128 to 2048x1024, with the aspect ratio 2:1. Friction is
    slip weakening (config_c0 distributions) or rate-and-state (config_c1
    distributions), drawn per scenario. The nucleation position is random.  Each
    scenario has 0 to max_levels refined levels (ratio 2) around its nucleation,
    drawn from a stream of its own (the other draws do not depend on max_levels).
    """
    rng = np.random.default_rng(seed)
    lrng = np.random.default_rng(seed + 7777)
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
            xn = nx // 2 + shift
        else:
            cfg = scenarios.config_c1(seed=seed * 100003 + k, nx=nx, ny=ny, dx=dx)
            cfg["rs"]["nuc_x"] = float(rng.uniform(0.25, 0.75) * nx * dx)
            cfg["rs"]["nuc_R"] = 1500.0 * dx / 25.0
            xn = int(cfg["rs"]["nuc_x"] / dx)
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1) * steps_scale)
        levels = int(lrng.integers(0, max_levels + 1)) if max_levels > 0 else 0
        out.append({"id": k, "cfg": cfg, "steps": max(steps, 2), "levels": levels, "xn": xn})
    return out


class Running:
    __slots__ = ("sc", "ctxs", "levels", "done_steps", "arrival", "admitted", "dt", "peak_V")

    def __init__(self, sc, ctxs, levels, arrival, admitted, dt):
        self.sc, self.ctxs, self.levels = sc, ctxs, levels
        self.arrival, self.admitted, self.dt = arrival, admitted, dt
        self.done_steps = 0
        self.peak_V = None

    @property
    def lev(self):
        return self.levels[0]

    @property
    def nlev(self):
        return len(self.levels)

    def close(self):
        for lev in self.levels:
            lev.close()
        for ctx in self.ctxs:
            ctx.close()


def _start(sc, precision, tile, shapes=None):
    """Contexts and levels of a scenario: one level, or a hierarchy around its nucleation.
    shapes: {"single": rows, "refined": rows} strip heights of the row-marching update for the
    coarsest level and for the refined ones (every member of a launch must agree)."""
    ctxs, levels = amr.make_levels(sc["cfg"], amr_layout(sc), AMR_RATIO, precision, tile)
    if shapes:
        for l, lev in enumerate(levels):
            lev.set_update_shape(shapes["single"] if l == 0 else shapes["refined"])
    return ctxs, levels


TUNE_ROWS = (8, 16, 32)


def _time_batch(scs, precision, rows0, rowsf, steps):
    """ms per (coarse) step of one batch of scenarios scs with strip heights rows0 (coarsest
    level) and rowsf (refined levels), best of two timed runs after a warm-up."""
    shapes = {"single": rows0, "refined": rowsf}
    started = [_start(sc, precision, None, shapes) for sc in scs]
    nlev = len(started[0][1])
    B = capi.Batch() if nlev == 1 else capi.AmrBatch(nlev, AMR_RATIO)
    for sc, (_, levels) in zip(scs, started):
        if nlev == 1:
            B.add(levels[0], 0.0, scenarios.dt_of(sc["cfg"]))
        else:
            B.add(levels, 0.0, scenarios.dt_of(sc["cfg"]))
    go = B.run if nlev == 1 else B.advance
    go(2)
    B.sync()
    best = float("inf")
    for _ in range(2):
        t = time.perf_counter()
        go(steps)
        B.sync()
        best = min(best, (time.perf_counter() - t) / steps * 1e3)
    for _, levels in started:
        B.remove(levels[0])
    B.close()
    for ctxs, levels in started:
        for lev in levels:
            lev.close()
        for c in ctxs:
            c.close()
    return best


def tune_shapes(scen_list, precision=capi.PREC_FP32_BULK, members=12, steps=40, log=None):
    """Update shapes for the loop's launches, timed on batches of the stream's own scenarios.

    Every member of a launch must run the same update shape, so the loop takes one row-marching
    strip height for the coarsest level of every scenario and one for the refined levels.  Each
    candidate (TUNE_ROWS) runs the first `members` single-level scenarios of the stream as one
    batch; the fastest height becomes the coarsest levels'.  With it, the first scenarios with
    the most refined levels run as one hierarchy batch for each candidate height of their refined
    levels, and the fastest of those becomes the refined levels'.  A level tuned alone
    (tmk_level_autotune) favours the shortest strips, its latency; in a batch the launch is
    full and the choice differs.  Returns {"single", "refined", "timings"}."""
    single = [sc for sc in scen_list if sc.get("levels", 0) == 0][:members]
    most = max(sc.get("levels", 0) for sc in scen_list)
    refined = [sc for sc in scen_list if sc.get("levels", 0) == most][:members] if most > 0 else []
    timings = {"single": {}, "refined": {}}
    rows0 = 32
    if single:
        for r in TUNE_ROWS:
            timings["single"][r] = _time_batch(single, precision, r, r, steps)
        rows0 = min(timings["single"], key=timings["single"].get)
    rowsf = rows0
    if refined:
        for r in TUNE_ROWS:
            timings["refined"][r] = _time_batch(refined, precision, rows0, r, max(steps // 4, 4))
        rowsf = min(timings["refined"], key=timings["refined"].get)
    if log:
        log("update shapes: coarsest levels %d-row strips, refined levels %d-row strips (ms per step %s)" %
            (rows0, rowsf, json.dumps(timings)))
    return {"single": rows0, "refined": rowsf, "timings": timings}


def _footprint(sc, precision):
    """Device bytes of a scenario: its levels' arrays plus, with refined levels, the flux arrays
    and face masks of each level's islands (bounded by the boxes' bounding box)."""
    layout = amr_layout(sc)
    cfg = sc["cfg"]
    tot = 0
    for l, boxes in enumerate(layout):
        tot += capi.level_footprint_params(cfg, boxes, precision)
        if len(layout) > 1:
            w = max(b[2] for b in boxes) - min(b[0] for b in boxes) + 1
            h = max(b[3] for b in boxes) - min(b[1] for b in boxes) + 1
            tot += 2 * scenarios.NCOMP * 8 * (w + 1) * (h + 1) + (w + 2) * (h + 2)
        cfg = scenarios.refine_cfg(cfg, AMR_RATIO)
    return tot


def _cell_updates_per_step(sc):
    return sum(sum((b[2] - b[0] + 1) * (b[3] - b[1] + 1) for b in boxes) * AMR_RATIO ** l
               for l, boxes in enumerate(amr_layout(sc)))


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
            "levels": r.sc.get("levels", 0),
            "max_slip": float(np.max(slip)), "mean_slip": float(np.mean(slip)),
            "final_max_slip_rate": float(np.max(V)),
            "ruptured_fraction": float(np.mean(ruptured)),
            "first_rupture_time": float(np.min(rt[ruptured])) if ruptured.any() else None}


class _Lane:
    """One batch of the loop -- the single-level batch, or the hierarchy batch of one level count
    -- with its members and its own chunk cadence."""

    def __init__(self, nlev):
        self.nlev = nlev
        self.b = capi.Batch() if nlev == 1 else capi.AmrBatch(nlev, AMR_RATIO)
        self.running = []   # members
        self.ready = []     # set up, joining at this lane's next chunk boundary
        self.inflight = 0   # steps of the chunk in flight (0: none)

    def add(self, r):
        if self.nlev == 1:
            self.b.add(r.lev, 0.0, r.dt)
        else:
            self.b.add(r.levels, 0.0, r.dt)

    def launch(self, n):
        (self.b.run if self.nlev == 1 else self.b.advance)(n)
        self.inflight = n


def serve(scen_list, rate_hz, precision=capi.PREC_FP32_BULK, chunk=50, mem_reserve=8 << 30, tile=None,
          log=None, keep_fault=False, shapes=None):
    """Run the arrival stream; returns (results, stats).

    Each batch (the single-level one, and one hierarchy batch per level count) is a lane with its
    own chunk cadence: the loop polls the lanes without waiting (tmk_batch_done, tmk_amr_done),
    and a lane whose chunk has completed retires its finished members, takes in the scenarios
    that joined it and starts its next chunk while the other lanes' chunks run.  New scenarios
    are set up (context, levels, initial state) while chunks run.  Membership changes take effect
    at the lane's chunk boundaries.
    """
    rng = np.random.default_rng(12345)
    gaps = rng.exponential(1.0 / rate_hz, size=len(scen_list)) if rate_hz > 0 else np.zeros(len(scen_list))
    arrivals = np.cumsum(gaps) - gaps[0]
    pending = list(zip(arrivals, range(len(scen_list))))
    queue = []
    lanes = {}        # level count -> lane
    results, latencies = [], []
    n_failed = 0
    cell_updates = 0
    footprint = {}    # (fault, nx, ny) -> device bytes of one scenario's level
    t0 = time.perf_counter()
    ip = 0
    phase = {"admit": 0.0, "join": 0.0, "launch": 0.0, "poll": 0.0, "sync": 0.0, "retire": 0.0, "idle": 0.0}

    def now():
        return time.perf_counter() - t0

    class _clock:  # host time spent per phase of the loop
        def __init__(self, name):
            self.name = name

        def __enter__(self):
            self.t = time.perf_counter()

        def __exit__(self, *exc):
            phase[self.name] += time.perf_counter() - self.t

    def lane_of(nlev):
        if nlev not in lanes:
            lanes[nlev] = _Lane(nlev)
        return lanes[nlev]

    def admit():
        # admission by free device memory, in arrival order
        while queue:
            arr, k = queue[0]
            sc = scen_list[k]
            cfg = sc["cfg"]
            key = (cfg["fault"], cfg["nx"], cfg["ny"], sc.get("levels", 0), sc.get("xn", 0))
            if key not in footprint:
                footprint[key] = _footprint(sc, precision)
            free, _ = capi.mem_info()
            if footprint[key] + mem_reserve > free:
                break
            queue.pop(0)
            ctxs, levels = _start(sc, precision, tile, shapes)
            r = Running(sc, ctxs, levels, arr, now(), scenarios.dt_of(cfg))
            lane_of(r.nlev).ready.append(r)

    def retire(lane, r, B, failed_code=None):
        lane.running.remove(r)
        r.close()
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
            log("done %d (%dx%d %s, %d refined levels, %d steps) latency %.2fs running %d queued %d" %
                (res["id"], cfg["nx"], cfg["ny"], cfg["fault"], r.sc.get("levels", 0), r.sc["steps"],
                 res["latency_s"], sum(len(l.running) for l in lanes.values()), len(queue)))

    def tend(lane):
        """A lane whose chunk has completed: account it, retire, join, start the next chunk."""
        nonlocal n_failed, cell_updates
        if lane.inflight:
            failed = {}
            with _clock("sync"):
                # a scenario's own error (e.g. a NaN state variable) retires only that scenario
                try:
                    lane.b.sync()
                except capi.TmkError:
                    for r in lane.running:
                        code = lane.b.member_error(r.lev)
                        if code != 0:
                            failed[id(r)] = code
            n = lane.inflight
            lane.inflight = 0
            for r in lane.running:
                r.done_steps += n
                cell_updates += _cell_updates_per_step(r.sc) * n
            with _clock("retire"):
                for r in [r for r in lane.running if id(r) in failed]:
                    code = failed[id(r)]
                    lane.b.remove(r.lev)
                    res = {"id": r.sc["id"], "failed": capi.strerror(code), "arrival_s": float(r.arrival),
                           "latency_s": float(now() - r.arrival)}
                    results.append(res)
                    n_failed += 1
                    retire(lane, r, None, code)
                    if log:
                        log("failed %d: %s" % (res["id"], res["failed"]))
                for r in [r for r in lane.running if r.done_steps >= r.sc["steps"]]:
                    lane.b.remove(r.lev)
                    cfg = r.sc["cfg"]
                    B = np.empty((scenarios.NB[cfg["fault"]], cfg["nx"] + 2 * GHOST))
                    r.lev.download_fault(0, B)
                    retire(lane, r, B)
        with _clock("join"):
            for r in lane.ready:
                lane.add(r)
                lane.running.append(r)
            lane.ready.clear()
        if lane.running:
            n = min(chunk, min(r.sc["steps"] - r.done_steps for r in lane.running))
            with _clock("launch"):
                lane.launch(n)

    def busy():
        return any(l.running or l.ready or l.inflight for l in lanes.values())

    while ip < len(pending) or queue or busy():
        t = now()
        while ip < len(pending) and pending[ip][0] <= t:
            queue.append(pending[ip])
            ip += 1
        with _clock("admit"):
            admit()
        tended = False
        # the single-level lane first: its chunks are the shortest
        for nlev in sorted(lanes):
            lane = lanes[nlev]
            if lane.inflight:
                with _clock("poll"):
                    if not lane.b.done():
                        continue
            if lane.inflight or lane.ready:
                tend(lane)
                tended = True
        if not tended:
            if not busy() and not queue and ip < len(pending):
                with _clock("idle"):
                    time.sleep(max(0.0, pending[ip][0] - now()))
            else:
                with _clock("poll"):
                    time.sleep(2e-4)
    wall = now()
    for lane in lanes.values():
        lane.b.close()
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
    ap.add_argument("--max-levels", type=int, default=2, help="refined levels per scenario drawn from 0..this")
    ap.add_argument("--no-tune", action="store_true", help="the contexts' default update shape, untuned")
    a = ap.parse_args()
    sl = draw_scenarios(a.n, a.seed, nx_range=(256, a.nx_max), steps_scale=a.steps_scale, max_levels=a.max_levels)
    prec = capi.PREC_FP64 if a.fp64 else capi.PREC_FP32_BULK
    log = lambda m: print(m, flush=True)
    shapes = None if a.no_tune else tune_shapes(sl, prec, log=log)
    res, stats = serve(sl, a.rate, precision=prec, chunk=a.chunk, log=log, shapes=shapes)
    if shapes:
        stats["update_shapes"] = {"coarsest_rows": shapes["single"], "refined_rows": shapes["refined"]}
    stats.update({"n": a.n, "steps_scale": a.steps_scale, "precision": "fp64" if a.fp64 else "fp32-bulk",
                  "scenarios_by_levels": {str(k): sum(1 for sc in sl if sc.get("levels", 0) == k)
                                          for k in range(a.max_levels + 1)},
                  "grid_range": "256x128..%dx%d" % (a.nx_max, a.nx_max // 2)})
    print(json.dumps(stats), flush=True)


if __name__ == "__main__":
    main()
