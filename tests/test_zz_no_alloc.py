"""tmk_patch_update_state_noalloc (no device allocation, fixed per-device scratch) against
tmk_patch_update_state through the C ABI: the same kernels on the same tiles, so every output --
state, fault data, fluxes -- must be bit-identical, in FP64 and in FP32 bulk.  Two calls in a row
on the same context, the second on a smaller box, check that the shared scratch carries nothing
from one call into the next; a box too large for the scratch is refused before any device work."""
import numpy as np
import pytest

from tetemoko_amd import capi, scenarios

pytestmark = pytest.mark.gpu

G = scenarios.GHOST


def _inputs(cfg, valid, seed):
    U, B = scenarios.level_arrays(cfg)
    scenarios.perturb(U, seed=seed, amp_v=1.0, amp_s=5.0)
    rng = np.random.default_rng(seed + 1)
    B[10] += rng.uniform(-0.02, 0.02, size=B.shape[1])
    B[8] = rng.uniform(0, 1, size=B.shape[1])
    B[9] = np.where(rng.uniform(size=B.shape[1]) < 0.5, 1e10, 0.01)
    ubox = (valid[0] - G, valid[1] - G, valid[2] + G, valid[3] + G)
    Ub = np.ascontiguousarray(U[:, ubox[1] + G:ubox[3] + G + 1, ubox[0] + G:ubox[2] + G + 1])
    bbox = (valid[0] - G, 0, valid[2] + G, 0)
    Bb = np.ascontiguousarray(B[:, bbox[0] + G:bbox[2] + G + 1]) if valid[1] == 0 else None
    return Ub, ubox, Bb, bbox


def _call(ctx, valid, Ub, ubox, Bb, bbox, noalloc, dt):
    U = Ub.copy()
    B = Bb.copy() if Bb is not None else None
    nfx = (valid[2] - valid[0] + 2, valid[3] - valid[1] + 1)
    nfy = (valid[2] - valid[0] + 1, valid[3] - valid[1] + 2)
    F0 = np.full((22, nfx[1], nfx[0]), np.nan)
    F1 = np.full((22, nfy[1], nfy[0]), np.nan)
    mws = ctx.update_state(valid, U, ubox, B, bbox if B is not None else None, F0, F1, time=0.41, dt=dt,
                           noalloc=noalloc)
    return mws, U, B, F0, F1


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_noalloc_matches_update_state_two_calls(precision):
    cfg = scenarios.config_c1(seed=5, nx=160, ny=96, dx=25.0)
    cfg["rs"]["state_law"] = 1
    dt = scenarios.dt_of(cfg)
    ctx = capi.Context(cfg, precision)
    try:
        # a box on the fault with interior tiles and boundary pieces, then a smaller box off the fault
        for valid, seed in (((8, 0, 135, 69), 21), ((40, 20, 71, 43), 22)):
            Ub, ubox, Bb, bbox = _inputs(cfg, valid, seed)
            ref = _call(ctx, valid, Ub, ubox, Bb, bbox, False, dt)
            got = _call(ctx, valid, Ub, ubox, Bb, bbox, True, dt)
            assert got[0] == ref[0] == max(cfg["cs"], cfg["cp"])
            for name, a, b in zip(("U", "B", "F0", "F1"), got[1:], ref[1:]):
                if b is None:
                    assert a is None
                    continue
                assert not np.isnan(b).any(), name
                assert np.array_equal(a, b), (valid, name, np.max(np.abs(a - b)))
            assert not np.array_equal(got[1], Ub)  # the call updated the state
    finally:
        ctx.close()


def test_noalloc_refuses_a_box_beyond_its_scratch():
    cfg = scenarios.config_c1(seed=6, nx=400, ny=240, dx=25.0)
    dt = scenarios.dt_of(cfg)
    ctx = capi.Context(cfg, capi.PREC_FP64)
    try:
        # 400 x 200 cells: two FP64 state copies alone exceed 16 MiB
        valid = (0, 0, 399, 199)
        Ub, ubox, Bb, bbox = _inputs(cfg, valid, 31)
        U = Ub.copy()
        with pytest.raises(capi.TmkError) as ei:
            ctx.update_state(valid, U, ubox, Bb.copy(), bbox, time=0.0, dt=dt, noalloc=True)
        assert ei.value.code == capi.E_SCRATCH
        assert np.array_equal(U, Ub)  # refused before any device work
        # the context and the allocation-free path still work afterwards
        valid = (0, 0, 63, 31)
        Ub, ubox, Bb, bbox = _inputs(cfg, valid, 32)
        ref = _call(ctx, valid, Ub, ubox, Bb, bbox, False, dt)
        got = _call(ctx, valid, Ub, ubox, Bb, bbox, True, dt)
        for a, b in zip(got[1:], ref[1:]):
            assert np.array_equal(a, b)
    finally:
        ctx.close()
