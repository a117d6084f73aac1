"""The serving loop on the device: continuous batching of single-level scenarios beside
hierarchies with refined levels, each result the same as the scenario run alone.

Nothing couples the scenarios: single-level ones share update launches, hierarchies run their
own graphs on their own streams, so every scenario's fault data at completion equals the same
scenario stepped alone (bit for bit in FP64; FP32 bulk reorders nothing either)."""
import numpy as np
import pytest

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
