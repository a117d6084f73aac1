"""Tile-shape autotuner for the fused update kernel.

For a workload (a scenario config and a box layout), time every tile shape the
library is built for. Each shape is timed with tmk_level_profile (HIP events on
the level's stream, one fused launch per step), and the fastest is returned.
The same inputs always produce the same state whatever the tile shape: tiling
does not change any computed value. So the choice only affects speed.
"""
import numpy as np

from . import capi, scenarios

SHAPES = ((16, 16), (32, 8), (32, 12), (32, 16), (32, 32), (64, 4), (64, 8), (64, 16))


def tune(cfg, boxes=None, precision=capi.PREC_FP64, steps=20, shapes=SHAPES, log=None):
    """Returns (best_shape, {shape: ms_per_step})."""
    boxes = boxes or [(0, 0, cfg["nx"] - 1, cfg["ny"] - 1)]
    U, B = scenarios.level_arrays(cfg)
    dt = scenarios.dt_of(cfg)
    times = {}
    for shape in shapes:
        ctx = capi.Context(cfg, precision)
        ctx.set_tile(*shape)
        lev = capi.Level(ctx, boxes)
        for k, b in enumerate(boxes):
            lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
        lev.profile(3, 0.0, dt)          # warm-up (code objects, caches)
        ms = lev.profile(steps, 3 * dt, dt)
        lev.sync()
        lev.close()
        ctx.close()
        times[shape] = ms
        if log:
            log("tile %dx%d: %.4f ms/step" % (shape[0], shape[1], ms))
    best = min(times, key=times.get)
    return best, times


if __name__ == "__main__":
    import json
    import sys
    nx = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    cfg = scenarios.config_c1(seed=0, nx=nx, ny=nx // 2, dx=25.0)
    out = {}
    for prec, name in ((capi.PREC_FP64, "fp64"), (capi.PREC_FP32_BULK, "fp32_bulk")):
        best, times = tune(cfg, precision=prec, log=lambda m: print(m, flush=True))
        out[name] = {"best": "%dx%d" % best, "ms_per_step": {"%dx%d" % k: v for k, v in times.items()}}
    print(json.dumps(out))
