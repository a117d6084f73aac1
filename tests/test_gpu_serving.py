"""The serving loop on the device: continuous batching of single-level scenarios beside
hierarchies with refined levels, each result the same as the scenario run alone and as the
oracle's run of it.

Nothing couples the scenarios: single-level ones share update launches, hierarchies run their
own graphs on their own streams, so every scenario's fault data at completion equals the same
scenario stepped alone (bit for bit in FP64; FP32 bulk reorders nothing either)."""
import numpy as np
import pytest

import amr as oamr
import tmo
from faultcheck import SENTINEL, fault_errors, rupture_comp
from tetemoko_amd import amr, capi, scenarios, serving

pytestmark = pytest.mark.gpu

G = scenarios.GHOST


def _alone(sc, precision):
    cfg = sc["cfg"]
    if sc["levels"] > 0:
        H = amr.Hierarchy(cfg, serving.amr_layout(sc), serving.AMR_RATIO, precision)
        H.run(sc["steps"])
        H.sync()
        B = np.empty((scenarios.NB[cfg["fault"]], cfg["nx"] + 2 * G))
        H.levels[0].download_fault(0, B)
        H.close()
        return B
    ctx = capi.Context(cfg, precision)
    lev = capi.Level(ctx, [(0, 0, cfg["nx"] - 1, cfg["ny"] - 1)])
    lev.fill(0, scenarios.background_vector(cfg), scenarios.fault_arrays(cfg))
    lev.run(sc["steps"], 0.0, scenarios.dt_of(cfg))
    lev.sync()
    B = np.empty((scenarios.NB[cfg["fault"]], cfg["nx"] + 2 * G))
    lev.download_fault(0, B)
    lev.close()
    ctx.close()
    return B


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_mixed_levels_each_result_equals_alone(precision):
    sl = serving.draw_scenarios(8, 11, nx_range=(256, 512), steps_range=(40, 120), steps_scale=1.0)
    assert {sc["levels"] for sc in sl} >= {0, 1, 2}
    results, stats = serving.serve(sl, 0.0, precision=precision, chunk=16, keep_fault=True)
    assert stats["failed"] == 0 and stats["scenarios"] == len(sl)
    by_id = {r["id"]: r for r in results}
    for sc in sl:
        B = _alone(sc, precision)
        assert np.array_equal(by_id[sc["id"]]["fault_data"], B), sc["id"]


def _oracle(sc):
    """Fault data (with ghosts) of a scenario run by the oracle: tmo level steps for one level,
    the multi-level oracle for a hierarchy; the same initial state as serving's (background,
    initial fault data) and the same clock."""
    cfg = sc["cfg"]
    dt = scenarios.dt_of(cfg)
    if sc["levels"] > 0:
        boxes = serving.amr_layout(sc)
        cfgs = [cfg]
        for _ in range(1, len(boxes)):
            cfgs.append(scenarios.refine_cfg(cfgs[-1], serving.AMR_RATIO))
        h = oamr.Hierarchy(cfg, boxes, serving.AMR_RATIO)
        h.init_state(lambda l: scenarios.level_arrays(cfgs[l])[0], lambda l: scenarios.level_arrays(cfgs[l])[1])
        h.set_dt(dt)
        assert h.advance(sc["steps"]) == 0
        return h.levels[0].B
    U, B = scenarios.level_arrays(cfg)
    assert tmo.level_steps(tmo.OracleParams(cfg), U, B, 0.0, dt, sc["steps"], nthreads=16, maxbox=64) == 0
    return B


def test_served_results_match_oracle():
    """Six scenarios (slip weakening and rate-and-state, 0, 1 and 2 refined levels, 256 to 288
    cells wide, 10 to 20 steps) served together in FP64: each one's fault data at completion
    against the oracle's run of that scenario alone, on every face of the base level.  Slip
    weakening bit for bit; rate-and-state within 1e-12 relative per component (exp/log/asinh
    differ in the last bit between the device and glibc), rupture times per face; every
    scenario has ruptured faces."""
    sl = serving.draw_scenarios(6, 30, nx_range=(256, 320), steps_range=(40, 120), steps_scale=0.2)
    assert {sc["levels"] for sc in sl} == {0, 1, 2} and {sc["cfg"]["fault"] for sc in sl} == {"sw", "rs"}
    results, stats = serving.serve(sl, 0.0, precision=capi.PREC_FP64, chunk=4, keep_fault=True)
    assert stats["failed"] == 0 and stats["scenarios"] == len(sl)
    by_id = {r["id"]: r for r in results}
    for sc in sl:
        Br = _oracle(sc)[:, G:-G]
        Bg = by_id[sc["id"]]["fault_data"][:, G:-G]
        if sc["cfg"]["fault"] == "sw":
            assert np.array_equal(Bg, Br), sc["id"]
        else:
            e = fault_errors(Bg, Br)
            assert np.all(e <= 1e-12), (sc["id"], e)
        assert np.count_nonzero(Br[rupture_comp(Br)] != SENTINEL) > 0, sc["id"]


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_tuned_shapes_serve_the_same_results(precision):
    """The loop's shape tuner (batches of the stream's own scenarios, each strip height timed)
    picks heights it offers, and the scenarios served with them give the same fault data as
    with the contexts' default shape: the strip height changes which lanes compute a cell,
    never its arithmetic."""
    sl = serving.draw_scenarios(8, 11, nx_range=(256, 512), steps_range=(40, 120), steps_scale=1.0)
    shapes = serving.tune_shapes(sl, precision, members=3, steps=4)
    assert shapes["single"] in serving.TUNE_ROWS and shapes["refined"] in serving.TUNE_ROWS
    assert set(shapes["timings"]["single"]) == set(serving.TUNE_ROWS)
    assert set(shapes["timings"]["refined"]) == set(serving.TUNE_ROWS)
    # force heights other than the default so that the comparison means something
    shapes = {"single": 8, "refined": 16}
    results, stats = serving.serve(sl, 0.0, precision=precision, chunk=16, keep_fault=True, shapes=shapes)
    assert stats["failed"] == 0 and stats["scenarios"] == len(sl)
    by_id = {r["id"]: r for r in results}
    for sc in sl:
        assert np.array_equal(by_id[sc["id"]]["fault_data"], _alone(sc, precision)), sc["id"]
