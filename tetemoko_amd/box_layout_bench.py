"""Batching across many boxes: one level cut into boxes of different sizes.

The reference updates a level box by box (LinElastLevelGodunov.cpp:284-374, max_grid_size 32
by default, 8-64 cells per side on refined AMR levels).  Here every box of a level goes into the
same two update launches, with a ghost exchange between boxes each step.  This measures the
cost of a box layout: a 1024 x 512 rate-and-state level as one box and as boxes of 64, 32, 16
and 8 cells per side, FP64 and FP32 bulk.  Times come from HIP events on the level's stream
(graph-replayed steps); the per-kernel split from tmk_level_profile_kernels.
"""
import argparse
import json

from . import capi, scenarios


def run(nx=1024, ny=512, sizes=(1024, 64, 32, 16, 8), steps=40, log=print):
    cfg = scenarios.config_c1(seed=0, nx=nx, ny=ny, dx=25.0)
    U, B = scenarios.level_arrays(cfg)
    dt = scenarios.dt_of(cfg)
    out = {}
    for prec, name in ((capi.PREC_FP64, "fp64"), (capi.PREC_FP32_BULK, "fp32")):
        for mb in sizes:
            boxes = scenarios.split_boxes(nx, ny, mb)
            ctx = capi.Context(cfg, prec)
            lev = capi.Level(ctx, boxes)
            for k, b in enumerate(boxes):
                lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
            lev.run(4, 0.0, dt)
            lev.sync()
            ms = lev.run_timed(steps, 4 * dt, dt)
            lev.sync()
            p = lev.profile_kernels(5, (4 + steps) * dt, dt)
            lev.sync()
            lev.close()
            ctx.close()
            key = "%s_box%d" % (name, mb)
            out[key] = {"boxes": len(boxes), "ms_per_step": ms / steps,
                        "cell_updates_per_s": nx * ny * steps / (ms * 1e-3),
                        "tiles_interior": p["tiles_interior"], "tiles_boundary": p["tiles_boundary"],
                        "ms_interior_kernel": p["ms_interior"], "ms_boundary_kernel": p["ms_boundary"]}
            if log:
                log("%s %s" % (key, json.dumps(out[key])))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--steps", type=int, default=40)
    a = ap.parse_args()
    run(a.nx, a.ny, steps=a.steps, log=lambda m: print(m, flush=True))


if __name__ == "__main__":
    main()
