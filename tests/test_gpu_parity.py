"""GPU parity: the HIP path through the C ABI against the CPU oracle on seeded inputs.

FP64 tolerance: 1e-12 relative on state, slip and rupture time.  State and fault data are
measured per component as max|gpu - oracle| / max|oracle|; the rupture time per face
(faultcheck.fault_errors): unruptured faces must hold the 1e10 sentinel exactly and ruptured
faces their oracle time to 1e-12 of that time.  Patches that do not touch the fault involve
no transcendental function and are expected to be bit-identical (both sides evaluate the
same expressions without FMA contraction).
"""
import ctypes
import itertools

import numpy as np
import pytest

import tmo
from faultcheck import fault_errors, new_ruptures
from tetemoko_amd import capi, scenarios

pytestmark = pytest.mark.gpu

G = scenarios.GHOST
TOL = 1e-12


def _cfg(fault, pred, nx=48, ny=40, seed=7, prim=1, char=0, fourth=1, bc=(0, 0, 0, 0), antiplane=False):
    if fault == "rs":
        cfg = scenarios.config_c1(seed=seed, nx=nx, ny=ny, dx=25.0)
        cfg["rs"]["state_law"] = 0 if seed % 2 else 1
    else:
        cfg = scenarios.config_c0(seed=seed, nx=nx, ny=ny, dx=100.0)
    cfg["numerics"] = {"normal_predictor": pred, "fourth_order": fourth, "prim_limiting": prim,
                       "char_limiting": char}
    cfg["bc"] = bc
    if antiplane:
        bg = list(cfg["background"])
        bg[4] = 3.0   # sxz0
        bg[5] = -2.0  # syz0
        cfg["background"] = tuple(bg)
    return cfg


def _state(cfg, seed, antiplane=False):
    U, B = scenarios.level_arrays(cfg)
    scenarios.perturb(U, seed=seed, amp_v=1.0, amp_s=5.0, antiplane=antiplane)
    rng = np.random.default_rng(seed + 1)
    if cfg["fault"] == "rs":
        B[10] += rng.uniform(-0.02, 0.02, size=B.shape[1])
        B[8] = rng.uniform(0, 1, size=B.shape[1])
        B[9] = np.where(rng.uniform(size=B.shape[1]) < 0.5, 1e10, 0.01)
    else:
        B[7] = rng.uniform(0, 0.6, size=B.shape[1])
        B[4] = rng.uniform(0, 0.01, size=B.shape[1])
    return U, B


def _oracle_patch(cfg, valid, U, ubox, B, bbox, time, dt, with_flux=False):
    op = tmo.OracleParams(cfg)
    L = tmo.lib()
    Uf = tmo.FabArray(ubox[:2], ubox[2:], 22, data=U)
    Bf = tmo.FabArray((bbox[0], 0), (bbox[2], 0), B.shape[0], data=B.reshape(B.shape[0], 1, -1)) if B is not None else None
    F0 = F1 = None
    if with_flux:
        F0 = tmo.FabArray((valid[0], valid[1]), (valid[2] + 1, valid[3]), 22)
        F1 = tmo.FabArray((valid[0], valid[1]), (valid[2], valid[3] + 1), 22)
    e = L.tmo_update_state(Uf.fab, Bf.fab if Bf else None, F0.fab if F0 else None, F1.fab if F1 else None,
                           op.ref, tmo.Box.make(*valid), time, dt)
    return e, (F0.a if F0 else None), (F1.a if F1 else None)


def _relerr(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    scale = np.maximum(np.max(np.abs(b), axis=1), 1e-300)
    return np.max(np.abs(a - b), axis=1) / scale


def _patch_case(cfg, valid, seed, precision=capi.PREC_FP64, antiplane=False, with_flux=False, tile=None):
    nx, ny = cfg["nx"], cfg["ny"]
    U, B = _state(cfg, seed, antiplane)
    ubox = (valid[0] - G, valid[1] - G, valid[2] + G, valid[3] + G)
    Ub = np.ascontiguousarray(U[:, ubox[1] + G:ubox[3] + G + 1, ubox[0] + G:ubox[2] + G + 1])
    touches = valid[1] == 0
    bbox = (valid[0] - G, 0, valid[2] + G, 0)
    Bb = np.ascontiguousarray(B[:, bbox[0] + G:bbox[2] + G + 1]) if touches else None
    dt = scenarios.dt_of(cfg)
    time = 0.37
    Ur, Br = Ub.copy(), (Bb.copy() if Bb is not None else None)
    e, F0r, F1r = _oracle_patch(cfg, valid, Ur, ubox, Br, bbox, time, dt, with_flux)
    assert e == 0
    ctx = capi.Context(cfg, precision)
    if tile:
        ctx.set_tile(*tile)
    Ug, Bg = Ub.copy(), (Bb.copy() if Bb is not None else None)
    F0g = F1g = None
    if with_flux:
        F0g = np.zeros_like(F0r)
        F1g = np.zeros_like(F1r)
    mws = ctx.update_state(valid, Ug, ubox, Bg, bbox if touches else None, F0g, F1g, time=time, dt=dt)
    ctx.close()
    assert mws == max(cfg["cs"], cfg["cp"])
    if touches and cfg["fault"] == "rs":
        # the step stamps rupture times (the seeded slip rates exceed the threshold)
        assert new_ruptures(Bb, Br) > 0
    return (Ug, Ur), (Bg, Br), (F0g, F0r), (F1g, F1r), touches


BOXES = {
    "interior": (12, 10, 29, 25),
    "fault": (8, 0, 27, 15),
    "xlo_fault_corner": (0, 0, 19, 11),
    "xhi": (30, 12, 47, 30),
    "yhi": (10, 24, 33, 39),
    "xhi_yhi_corner": (34, 28, 47, 39),
    "whole": (0, 0, 47, 39),
}


@pytest.mark.parametrize("pred,prim,char,fourth", [
    ("CTU", 1, 0, 1), ("PLM", 1, 0, 1), ("PLM", 1, 0, 0), ("PLM", 0, 1, 1), ("PLM", 0, 0, 0),
    ("PPM", 1, 0, 1), ("PPM", 0, 1, 1), ("PPM", 0, 0, 1),
])
@pytest.mark.parametrize("fault", ["rs", "sw"])
@pytest.mark.parametrize("where", list(BOXES))
def test_patch_update_fp64(pred, prim, char, fourth, fault, where):
    print("case", pred, prim, char, fourth, fault, where, flush=True)
    cfg = _cfg(fault, pred, prim=prim, char=char, fourth=fourth, bc=(0, 1, 0, 1) if where.endswith("corner") else (0, 0, 0, 0))
    (Ug, Ur), (Bg, Br), _, _, touches = _patch_case(cfg, BOXES[where], seed=11)
    eu = _relerr(Ug, Ur)
    assert np.all(eu <= TOL), "state rel err per comp: %s" % eu
    if not touches:
        assert np.array_equal(Ug, Ur), "non-fault patch not bit-identical: %s" % eu
    if touches:
        eb = fault_errors(Bg, Br)
        assert np.all(eb <= TOL), "fault rel err per comp: %s" % eb


@pytest.mark.parametrize("where", ["interior", "fault", "whole"])
def test_patch_antiplane(where):
    cfg = _cfg("rs", "PPM", antiplane=True)
    (Ug, Ur), (Bg, Br), _, _, touches = _patch_case(cfg, BOXES[where], seed=3, antiplane=True)
    eu = _relerr(Ug, Ur)
    assert np.all(eu <= TOL), eu
    if touches:
        assert np.all(fault_errors(Bg, Br) <= TOL)


@pytest.mark.parametrize("tile", [(16, 16), (32, 8), (32, 12), (32, 16), (64, 4), (64, 8), (32, 32), (64, 16)])
def test_patch_tile_shapes(tile):
    cfg = _cfg("rs", "PPM")
    (Ug, Ur), (Bg, Br), _, _, _ = _patch_case(cfg, BOXES["whole"], seed=5, tile=tile)
    assert np.all(_relerr(Ug, Ur) <= TOL)
    assert np.all(fault_errors(Bg, Br) <= TOL)


def test_patch_fluxes():
    cfg = _cfg("rs", "PPM")
    _, _, (F0g, F0r), (F1g, F1r), _ = _patch_case(cfg, BOXES["fault"], seed=9, with_flux=True)
    assert np.all(_relerr(F0g, F0r) <= TOL)
    assert np.all(_relerr(F1g, F1r) <= TOL)


def _level_case(cfg, boxes, nsteps, seed, precision=capi.PREC_FP64, use_run=False, tile=None, march=None):
    U, B = _state(cfg, seed)
    Ur, Br = U.copy(), B.copy()
    dt = scenarios.dt_of(cfg)
    assert tmo.level_steps(tmo.OracleParams(cfg), Ur, Br, 0.0, dt, nsteps) == 0
    if cfg["fault"] == "rs":
        assert new_ruptures(B[:, G:-G], Br[:, G:-G]) > 0  # rupture times are stamped in the window
    ctx = capi.Context(cfg, precision)
    if tile:
        ctx.set_tile(*tile)
    if march is not None:
        ctx.set_march(march)
    elif tile:
        ctx.set_march(0)  # a tile shape under test: the interior runs the tile kernels
    lev = capi.Level(ctx, boxes)
    for k, b in enumerate(boxes):
        lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
    if use_run:
        lev.run(nsteps, 0.0, dt)
    else:
        t = 0.0
        for _ in range(nsteps):
            lev.step(t, dt)
            t += dt
    lev.sync()
    Ug, Bg = np.full_like(U, np.nan), np.full_like(B, np.nan)
    for k, b in enumerate(boxes):
        ub = scenarios.box_view(U, b)
        bb = scenarios.box_bview(B, b) if b[1] == 0 else None
        lev.download(k, ub, bb)
        x0, y0, x1, y1 = b
        Ug[:, y0 + G:y1 + G + 1, x0 + G:x1 + G + 1] = ub[:, G:-G, G:-G]
        if bb is not None:
            Bg[:, x0 + G:x1 + G + 1] = bb[:, G:-G]
    lev.close()
    ctx.close()
    return (Ug[:, G:-G, G:-G], Ur[:, G:-G, G:-G]), (Bg[:, G:-G], Br[:, G:-G])


@pytest.mark.parametrize("fault", ["rs", "sw"])
@pytest.mark.parametrize("maxbox", [None, 16, 20])
def test_level_steps(fault, maxbox):
    cfg = _cfg(fault, "PPM", nx=64, ny=40)
    boxes = [(0, 0, 63, 39)] if maxbox is None else scenarios.split_boxes(64, 40, maxbox)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, boxes, 6, seed=21)
    assert np.all(_relerr(Ug, Ur) <= TOL), _relerr(Ug, Ur)
    assert np.all(fault_errors(Bg, Br) <= TOL), fault_errors(Bg, Br)


@pytest.mark.parametrize("pred,prim,char,fourth", [("PPM", 1, 0, 1), ("PPM", 0, 1, 1), ("PLM", 1, 0, 1)])
@pytest.mark.parametrize("tile", [None, (32, 16), (32, 12), (64, 4), (64, 8)])
def test_level_interior_tiles(pred, prim, char, fourth, tile):
    """A grid with many interior tiles (the specialised interior kernel: one thread per cell and
    side on the small FP64 tiles, one thread per cell on 32x16 and 64x8) and boundary tiles (the
    general kernel, on the side stream).  Away from the fault's reach (3 cells a step) the state
    is bit-identical to the oracle; everywhere within 1e-12."""
    cfg = _cfg("rs", pred, nx=200, ny=120, prim=prim, char=char, fourth=fourth)
    nsteps = 4
    (Ug, Ur), (Bg, Br) = _level_case(cfg, [(0, 0, 199, 119)], nsteps, seed=31, tile=tile)
    assert np.all(_relerr(Ug, Ur) <= TOL), _relerr(Ug, Ur)
    assert np.all(fault_errors(Bg, Br) <= TOL), fault_errors(Bg, Br)
    far = 3 * nsteps + 3
    assert np.array_equal(Ug[:, far:, :], Ur[:, far:, :])


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
@pytest.mark.parametrize("tile", [None, (32, 16)])
def test_level_many_boxes_partial_tiles(precision, tile):
    """Boxes of 24x24: tiles cut short at box edges run the interior kernel when their stencil lies
    inside the domain, reading neighbour data from the exchanged ghosts."""
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    boxes = scenarios.split_boxes(200, 120, 24)
    nsteps = 4
    (Ug, Ur), (Bg, Br) = _level_case(cfg, boxes, nsteps, seed=33, precision=precision, tile=tile)
    if precision == capi.PREC_FP64:
        assert np.all(_relerr(Ug, Ur) <= TOL), _relerr(Ug, Ur)
        assert np.all(fault_errors(Bg, Br) <= TOL), fault_errors(Bg, Br)
        far = 3 * nsteps + 3
        assert np.array_equal(Ug[:, far:, :], Ur[:, far:, :])
    else:
        assert np.all(_relerr(Ug, Ur) <= 1e-5), _relerr(Ug, Ur)


def test_level_interior_tiles_fp32():
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, [(0, 0, 199, 119)], 4, seed=32, precision=capi.PREC_FP32_BULK)
    assert np.all(_relerr(Ug, Ur) <= 1e-5), _relerr(Ug, Ur)
    assert np.all(_relerr(Bg[[6, 8, 10]], Br[[6, 8, 10]]) <= 1e-4), fault_errors(Bg, Br)


def test_profile_kernels():
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    U, B = _state(cfg, 5)
    ctx = capi.Context(cfg, capi.PREC_FP64)
    ctx.set_tile(32, 16)
    ctx.set_march(0)
    lev = capi.Level(ctx, [(0, 0, 199, 119)])
    lev.upload(0, U, B)
    p = lev.profile_kernels(3, 0.0, scenarios.dt_of(cfg))
    lev.sync()
    lev.close()
    ctx.close()
    assert p["cells_interior"] + p["cells_boundary"] == 200 * 120
    # 7 x 9 tiles of at most 30 x 14; each tile within 3 cells of the domain's edges gives its core
    # to the interior kernel and its bands (3 cells thick) to the boundary pieces
    assert p["tiles_interior"] == 7 * 9
    assert p["cells_interior"] == (200 - 6) * (120 - 6)
    # one band piece per edge a tile touches: 7 tiles along each row band, 9 along each column band
    assert p["tiles_boundary"] == 2 * 7 + 2 * 9
    assert p["ms_interior"] > 0 and p["ms_boundary"] > 0 and p["ms_step"] > 0


BIG = (0, 0, 199, 119)  # a 200x120 patch: interior tiles in every thread layout, plus boundary pieces


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
@pytest.mark.parametrize("tile", [None, (32, 16), (64, 4)])
def test_big_patch_antiplane(precision, tile):
    """Live anti-plane fields (three slots per sweep) on interior and boundary tiles."""
    cfg = _cfg("rs", "PPM", nx=200, ny=120, antiplane=True)
    (Ug, Ur), (Bg, Br), _, _, _ = _patch_case(cfg, BIG, seed=13, precision=precision, antiplane=True, tile=tile)
    tol = TOL if precision == capi.PREC_FP64 else 1e-5
    assert np.all(_relerr(Ug, Ur) <= tol), _relerr(Ug, Ur)
    assert np.all(_relerr(Bg[[6, 8, 10]], Br[[6, 8, 10]]) <= (TOL if tol == TOL else 1e-4))
    if precision == capi.PREC_FP64:
        assert np.array_equal(Ug[:, G + 4:, :], Ur[:, G + 4:, :])


@pytest.mark.parametrize("pred,prim,char,fourth", [("CTU", 1, 0, 1), ("PLM", 0, 0, 0), ("PLM", 1, 0, 0),
                                                   ("PLM", 0, 1, 1), ("PPM", 0, 0, 1)])
@pytest.mark.parametrize("tile", [None, (32, 16)])
def test_big_patch_predictors(pred, prim, char, fourth, tile):
    """Every predictor/limiter variant through the interior kernels (runtime predictor kind):
    bit-identical to the oracle beyond the fault's reach."""
    cfg = _cfg("sw", pred, nx=200, ny=120, prim=prim, char=char, fourth=fourth)
    (Ug, Ur), (Bg, Br), _, _, _ = _patch_case(cfg, BIG, seed=14, tile=tile)
    assert np.all(_relerr(Ug, Ur) <= TOL), _relerr(Ug, Ur)
    assert np.array_equal(Ug[:, G + 4:, :], Ur[:, G + 4:, :])


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
@pytest.mark.parametrize("tile", [None, (32, 16)])
def test_big_patch_fluxes(precision, tile):
    """Face fluxes for flux registers written by the interior kernels (both layouts)."""
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    cfg["rs"]["state_law"] = 1  # the slip law overflows psi on this seeded perturbation, in the oracle too
    _, _, (F0g, F0r), (F1g, F1r), _ = _patch_case(cfg, BIG, seed=15, precision=precision, with_flux=True,
                                                  tile=tile)
    tol = TOL if precision == capi.PREC_FP64 else 1e-5
    assert np.all(_relerr(F0g, F0r) <= tol), _relerr(F0g, F0r)
    assert np.all(_relerr(F1g, F1r) <= tol), _relerr(F1g, F1r)


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
@pytest.mark.parametrize("tile", [None, (32, 16)])
def test_big_level_plasticity(precision, tile):
    """The fused Drucker-Prager return on interior tiles, one thread per cell and per side."""
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    cfg["plasticity"] = 1
    cfg["p_mu"] = 0.6
    cfg["p_beta"] = 0.3
    cfg["p_eta"] = 0.03 * scenarios.G_MPA
    cfg["p_b"] = 1.0
    cfg["rs"]["state_law"] = 1
    cfg["background"] = (-120.0, -120.0, -120.0, 70.0, 0.0, 0.0)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, [(0, 0, 199, 119)], 4, seed=16, precision=precision, tile=tile)
    tol = TOL if precision == capi.PREC_FP64 else 1e-5
    assert np.all(_relerr(Ug, Ur) <= tol), _relerr(Ug, Ur)
    assert np.max(Ur[9]) > 0.0, "plastic flow expected somewhere"


def test_level_graph_run_matches_steps():
    cfg = _cfg("rs", "PPM", nx=64, ny=40)
    a = _level_case(cfg, [(0, 0, 63, 39)], 7, seed=4, use_run=True)
    (Ug, Ur), (Bg, Br) = a
    assert np.all(_relerr(Ug, Ur) <= TOL)
    assert np.all(fault_errors(Bg, Br) <= TOL)


def test_level_plasticity():
    cfg = _cfg("rs", "PPM", nx=64, ny=40)
    cfg["plasticity"] = 1
    cfg["p_mu"] = 0.6
    cfg["p_beta"] = 0.3
    cfg["p_eta"] = 0.03 * scenarios.G_MPA
    cfg["p_b"] = 1.0
    cfg["rs"]["state_law"] = 1
    cfg["background"] = (-120.0, -120.0, -120.0, 70.0, 0.0, 0.0)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, scenarios.split_boxes(64, 40, 20), 5, seed=8)
    assert np.all(_relerr(Ug, Ur) <= TOL), _relerr(Ug, Ur)
    assert np.all(fault_errors(Bg, Br) <= TOL)
    assert np.max(Ur[9]) > 0.0, "plastic flow expected somewhere"


def test_fp32_bulk_level():
    cfg = _cfg("rs", "PPM", nx=64, ny=40)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, [(0, 0, 63, 39)], 6, seed=21, precision=capi.PREC_FP32_BULK)
    assert np.all(_relerr(Ug, Ur) <= 1e-5), _relerr(Ug, Ur)
    assert np.all(_relerr(Bg[[6, 8, 10]], Br[[6, 8, 10]]) <= 1e-4), fault_errors(Bg, Br)


def test_psi_nan_error():
    cfg = _cfg("rs", "PPM")
    cfg["rs"]["state_law"] = 0
    U, B = _state(cfg, 2)
    B[10, 20] = np.nan
    valid = (8, 0, 27, 15)
    ubox = (valid[0] - G, valid[1] - G, valid[2] + G, valid[3] + G)
    Ub = np.ascontiguousarray(U[:, ubox[1] + G:ubox[3] + G + 1, ubox[0] + G:ubox[2] + G + 1])
    bbox = (valid[0] - G, 0, valid[2] + G, 0)
    Bb = np.ascontiguousarray(B[:, bbox[0] + G:bbox[2] + G + 1])
    ctx = capi.Context(cfg)
    with pytest.raises(capi.TmkError) as ei:
        ctx.update_state(valid, Ub, ubox, Bb, bbox, time=0.0, dt=scenarios.dt_of(cfg))
    assert ei.value.code == capi.E_PSI_NAN
    ctx.close()


def test_argument_errors():
    cfg = _cfg("rs", "PPM")
    U, B = _state(cfg, 2)
    ctx = capi.Context(cfg)
    valid = (8, 0, 27, 15)
    small = (valid[0] - 2, valid[1] - 2, valid[2] + 2, valid[3] + 2)
    Us = np.zeros((22, small[3] - small[1] + 1, small[2] - small[0] + 1))
    with pytest.raises(capi.TmkError) as ei:
        ctx.update_state(valid, Us, small, None, None, time=0.0, dt=0.001)
    assert ei.value.code == capi.E_BOX
    ubox = (valid[0] - G, valid[1] - G, valid[2] + G, valid[3] + G)
    Ub = np.zeros((22, ubox[3] - ubox[1] + 1, ubox[2] - ubox[0] + 1))
    with pytest.raises(capi.TmkError) as ei:
        ctx.update_state(valid, Ub, ubox, None, None, time=0.0, dt=0.001)
    assert ei.value.code == capi.E_NO_BDRY
    ctx.close()
    bad = dict(cfg)
    bad["bc"] = (0, 5, 0, 0)
    with pytest.raises(capi.TmkError) as ei:
        capi.Context(bad)
    assert ei.value.code == capi.E_BAD_BC


def _fresh_level(cfg, seed, precision=capi.PREC_FP32_BULK, boxes=None):
    U, B = scenarios.level_arrays(cfg)
    scenarios.perturb(U, seed=seed, amp_v=0.2, amp_s=1.0)
    ctx = capi.Context(cfg, precision)
    boxes = boxes or [(0, 0, cfg["nx"] - 1, cfg["ny"] - 1)]
    lev = capi.Level(ctx, boxes)
    for k, b in enumerate(boxes):
        lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
    return ctx, lev, U, B, boxes


def _download(lev, U, B, boxes):
    out_u, out_b = np.full_like(U, np.nan), np.full_like(B, np.nan)
    for k, b in enumerate(boxes):
        ub = scenarios.box_view(U, b)
        bb = scenarios.box_bview(B, b) if b[1] == 0 else None
        lev.download(k, ub, bb)
        x0, y0, x1, y1 = b
        out_u[:, y0 + G:y1 + G + 1, x0 + G:x1 + G + 1] = ub[:, G:-G, G:-G]
        if bb is not None:
            out_b[:, x0 + G:x1 + G + 1] = bb[:, G:-G]
    return out_u[:, G:-G, G:-G], out_b[:, G:-G]


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_batch_member_equals_alone(precision):
    """Continuous batching: a scenario's result is bit-identical alone and in the mix,
    with members joining and leaving at different steps."""
    cfgs = [scenarios.config_c1(seed=1, nx=96, ny=48, dx=25.0),
            scenarios.config_c1(seed=2, nx=64, ny=64, dx=40.0),
            scenarios.config_c0(seed=3, nx=80, ny=40)]
    cfgs[2]["fault"] = "sw"
    layouts = [None, scenarios.split_boxes(64, 64, 24)]
    # alone, with the box layout each scenario has in the mix
    alone = []
    for k, cfg in enumerate(cfgs[:2]):
        ctx, lev, U, B, boxes = _fresh_level(cfg, 10 + k, precision, boxes=layouts[k])
        lev.run(9, 0.0, scenarios.dt_of(cfg))
        lev.sync()
        alone.append(_download(lev, U, B, boxes))
        lev.close()
        ctx.close()
    # mixed: scenario 1 joins after 3 steps of scenario 0; scenario 0 leaves after its 9th step
    ctx0, lev0, U0, B0, bx0 = _fresh_level(cfgs[0], 10, precision)
    ctx1, lev1, U1, B1, bx1 = _fresh_level(cfgs[1], 11, precision, boxes=layouts[1])
    batch = capi.Batch()
    batch.add(lev0, 0.0, scenarios.dt_of(cfgs[0]))
    batch.run(3)
    batch.add(lev1, 0.0, scenarios.dt_of(cfgs[1]))
    batch.run(6)
    batch.sync()
    batch.remove(lev0)
    batch.run(3)
    batch.remove(lev1)
    lev1.run(0, 0.0, scenarios.dt_of(cfgs[1]))
    got0 = _download(lev0, U0, B0, bx0)
    got1 = _download(lev1, U1, B1, bx1)
    assert np.array_equal(got0[0], alone[0][0]) and np.array_equal(got0[1], alone[0][1])
    assert np.array_equal(got1[0], alone[1][0]) and np.array_equal(got1[1], alone[1][1])
    batch.close()
    for o in (lev0, lev1, ctx0, ctx1):
        o.close()


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_batch_mixed_plasticity(precision):
    """A batch of a scenario with plasticity and one without runs the row march's variant with
    plasticity for both (each member's own flag decides); alone, the scenario without plasticity
    runs the variant that never reads gam or writes lam and gam.  Each member matches its run
    alone: FP64 bit for bit, FP32 bulk within 1e-5 (its two variants may contract differently)."""
    cp = _cfg("rs", "PPM", nx=200, ny=120)
    cp["plasticity"] = 1
    cp["p_mu"], cp["p_beta"], cp["p_eta"], cp["p_b"] = 0.6, 0.3, 0.03 * scenarios.G_MPA, 1.0
    cp["rs"]["state_law"] = 1
    cp["background"] = (-120.0, -120.0, -120.0, 70.0, 0.0, 0.0)
    cn = _cfg("rs", "PPM", nx=160, ny=96, seed=8)
    cases = [(cp, 16), (cn, 17)]

    def make(cfg, seed):
        U, B = _state(cfg, seed)
        ctx = capi.Context(cfg, precision)
        ctx.set_march(8)
        boxes = scenarios.split_boxes(cfg["nx"], cfg["ny"], 32)
        lev = capi.Level(ctx, boxes)
        for k, b in enumerate(boxes):
            lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
        return ctx, lev, U, B, boxes

    alone = []
    for cfg, seed in cases:
        ctx, lev, U, B, boxes = make(cfg, seed)
        lev.run(6, 0.0, scenarios.dt_of(cfg))
        lev.sync()
        alone.append(_download(lev, U, B, boxes))
        lev.close()
        ctx.close()
    made = [make(cfg, seed) for cfg, seed in cases]
    batch = capi.Batch()
    for (cfg, _), (ctx, lev, U, B, boxes) in zip(cases, made):
        batch.add(lev, 0.0, scenarios.dt_of(cfg))
    batch.run(6)
    batch.sync()
    got = []
    for (cfg, _), (ctx, lev, U, B, boxes) in zip(cases, made):
        batch.remove(lev)
        lev.run(0, 0.0, scenarios.dt_of(cfg))
        got.append(_download(lev, U, B, boxes))
    batch.close()
    for ctx, lev, *_ in made:
        lev.close()
        ctx.close()
    assert np.max(alone[0][0][9]) > 0.0, "plastic flow expected somewhere"
    for (gu, gb), (au, ab) in zip(got, alone):
        if precision == capi.PREC_FP64:
            assert np.array_equal(gu, au) and np.array_equal(gb, ab)
        else:
            assert np.all(_relerr(gu, au) <= 1e-5), _relerr(gu, au)
            assert np.all(_relerr(gb[[6, 8, 10]], ab[[6, 8, 10]]) <= 1e-5)


@pytest.mark.parametrize("precision,tol", [(capi.PREC_FP64, 0.0), (capi.PREC_FP32_BULK, 1e-4)])
def test_box_layout_invariance(precision, tol):
    """The same scenario on one box and on 9 boxes.  FP64: bit-identical (every kernel rounds like
    the reference).  FP32 bulk: the interior and boundary kernels are built with FMA contraction,
    which may round a cell differently depending on which kernel its tile runs in; the two
    layouts agree within FP32-bulk's stated 1e-4 (measured: 2e-5 on velocities, which ride on
    120 MPa stresses, 3e-7 on stresses after 9 steps)."""
    cfg = scenarios.config_c1(seed=2, nx=64, ny=64, dx=40.0)
    dt = scenarios.dt_of(cfg)
    out = []
    for boxes in (None, scenarios.split_boxes(64, 64, 24)):
        ctx, lev, U, B, bx = _fresh_level(cfg, 11, precision, boxes=boxes)
        lev.run(9, 0.0, dt)
        lev.sync()
        out.append(_download(lev, U, B, bx))
        lev.close()
        ctx.close()
    if tol == 0.0:
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    else:
        assert np.all(_relerr(out[1][0], out[0][0]) <= tol), _relerr(out[1][0], out[0][0])
        assert np.all(_relerr(out[1][1][[6, 8, 10]], out[0][1][[6, 8, 10]]) <= tol)


def test_batch_member_error_isolated():
    """A member that fails (NaN psi on one fault face) reports its own error; the other member's
    result is bit-identical to running alone, and the batch carries on once it is removed."""
    good = scenarios.config_c1(seed=5, nx=96, ny=48, dx=25.0)
    bad = scenarios.config_c1(seed=6, nx=96, ny=48, dx=25.0)
    ctxa, leva, Ua, Ba, bxa = _fresh_level(good, 21, capi.PREC_FP64)
    leva.run(5, 0.0, scenarios.dt_of(good))
    leva.sync()
    alone = _download(leva, Ua, Ba, bxa)
    leva.close()
    ctxa.close()
    ctx0, lev0, U0, B0, bx0 = _fresh_level(good, 21, capi.PREC_FP64)
    U1, B1 = scenarios.level_arrays(bad)
    B1[10, G + 40] = np.nan
    ctx1 = capi.Context(bad, capi.PREC_FP64)
    lev1 = capi.Level(ctx1, [(0, 0, 95, 47)])
    lev1.upload(0, U1, B1)
    batch = capi.Batch()
    batch.add(lev0, 0.0, scenarios.dt_of(good))
    batch.add(lev1, 0.0, scenarios.dt_of(bad))
    batch.run(3)
    with pytest.raises(capi.TmkError) as ei:
        batch.sync()
    assert ei.value.code == capi.E_PSI_NAN
    assert batch.member_error(lev1) == capi.E_PSI_NAN
    assert batch.member_error(lev0) == 0
    batch.remove(lev1)
    batch.run(2)
    batch.sync()
    batch.remove(lev0)
    got = _download(lev0, U0, B0, bx0)
    assert np.array_equal(got[0], alone[0]) and np.array_equal(got[1], alone[1])
    batch.close()
    for o in (lev0, lev1, ctx0, ctx1):
        o.close()


def test_fault_stations_match_oracle():
    """Station records written on the device after each level step (graph-captured and eager)
    equal the reference expression B[i] (1 - d) + B[i+1] d on the oracle's fault data."""
    from tetemoko_amd import stations as stmod
    cfg = _cfg("rs", "PPM", nx=96, ny=48)
    U, B = _state(cfg, 9)
    dt = scenarios.dt_of(cfg)
    xs = [12.5 + 25.0 * 10, 25.0 * 40 + 3.0, 25.0 * 70 + 20.0, 25.0 * 95 - 12.5]
    ctx = capi.Context(cfg, capi.PREC_FP64)
    lev = capi.Level(ctx, scenarios.split_boxes(96, 48, 32))
    for k, b in enumerate(lev.boxes):
        lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
    st = capi.Stations(lev, xs, capacity=8)
    t = 0.0
    lev.run(3, t, dt)
    for _ in range(3):
        t = t + dt
    lev.step(t, dt)
    rec = st.read()
    assert rec.shape == (4, len(xs), 12)
    op = tmo.OracleParams(cfg)
    Ur, Br = U.copy(), B.copy()
    tr = 0.0
    for k in range(4):
        assert tmo.level_steps(op, Ur, Br, tr, dt, 1) == 0
        tr = tr + dt
        for s_, x in enumerate(xs):
            i, d = stmod.placement(x, cfg["dx"])
            want = Br[:, G + i] * (1.0 - d) + Br[:, G + i + 1] * d
            assert rec[k, s_, 0] == tr
            scale = np.maximum(np.max(np.abs(Br[:, G:-G]), axis=1), 1e-300)
            assert np.all(np.abs(rec[k, s_, 1:] - want) <= TOL * scale), (k, s_)
    assert st.read().shape[0] == 0
    st.close()
    lev.close()
    ctx.close()


def test_fault_stations_overflow():
    cfg = _cfg("sw", "PPM", nx=64, ny=40)
    U, B = _state(cfg, 3)
    ctx = capi.Context(cfg, capi.PREC_FP64)
    lev = capi.Level(ctx, [(0, 0, 63, 39)])
    lev.upload(0, U, B)
    st = capi.Stations(lev, [30.0 * 100], capacity=2)
    lev.run(3, 0.0, scenarios.dt_of(cfg))
    with pytest.raises(capi.TmkError) as ei:
        st.read()
    assert ei.value.code == capi.E_OVERFLOW
    lev.run(1, 0.0, scenarios.dt_of(cfg))
    got = st.read()
    assert got.shape == (1, 1, 10)
    with pytest.raises(capi.TmkError):
        capi.Stations(lev, [30.0 * 10], capacity=2)  # a second station set on the same level
    st.close()
    with pytest.raises(capi.TmkError):
        capi.Stations(lev, [1e9], capacity=2)  # a station off the fault
    lev.close()
    ctx.close()


def test_mem_info_and_footprint():
    free, total = capi.mem_info()
    assert 0 < free <= total
    cfg = scenarios.config_c1(seed=1, nx=96, ny=48, dx=25.0)
    ctx = capi.Context(cfg)
    n = capi.level_footprint(ctx, [(0, 0, 95, 47)])

    def cls(b):  # bytes the level-array cache allocates for a request of b
        g = max(256, (1 << (b.bit_length() - 1)) >> 3)
        return max(256, -(-b // g) * g)

    # three state arrays over the ghosted box, its cell-owner map, two fault-data rows and their
    # face-owner map
    g = (96 + 8) * (48 + 8)
    assert n == 3 * cls(22 * g * 8) + cls(g * 4) + 2 * cls((96 + 8) * 11 * 8) + cls((96 + 8) * 4)
    # and that is what a level of that layout holds
    live0 = capi.cache_info()[1]
    lev = capi.Level(ctx, [(0, 0, 95, 47)])
    assert capi.cache_info()[1] - live0 == n
    lev.close()
    assert capi.cache_info()[1] == live0
    # the context-free form agrees on every layout (one island, split islands, boxes off the fault)
    for boxes in ([(0, 0, 95, 47)], [(0, 0, 47, 47), (48, 0, 95, 47)], [(0, 0, 15, 15), (64, 24, 95, 47)]):
        assert capi.level_footprint_params(cfg, boxes) == capi.level_footprint(ctx, boxes), boxes
    ctx.close()


@pytest.mark.parametrize("fault", ["rs", "sw"])
@pytest.mark.parametrize("maxbox", [None, 16])
def test_tags_bit_exact(fault, maxbox):
    cfg = _cfg(fault, "PPM", nx=64, ny=40)
    if fault == "sw":
        cfg["sw"]["patches"] = [(2000.0, 800.0, 5.0)]
    cfg["rs"] = cfg.get("rs", {})
    boxes = [(0, 0, 63, 39)] if maxbox is None else scenarios.split_boxes(64, 40, maxbox)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, boxes, 3, seed=12)
    tps = [dict(refine_thresh=0.4, plastic_thresh=1e-5, tag_buffer=2),
           dict(refine_thresh=0.3, plastic_thresh=1e-5, tag_buffer=1, x_coarsen=(300.0, 900.0),
                y_coarsen=(-1.0, 400.0), slope_coarsen=2.0, width_coarsen=0.005 * 1000, level=1,
                coarsen2_max_level=1, x_coarsen2=(0.0, 1200.0), y_coarsen2=(0.0, 600.0)),
           dict(refine_thresh=5.0, plastic_thresh=1.0, tag_buffer=0, stress_thresh=4.0, sliprate_thresh=0.5)]
    # rebuild the level from the oracle's evolved state and tag on the device
    U = np.full((22, 40 + 2 * G, 64 + 2 * G), 0.0)
    U[:, G:-G, G:-G] = Ur
    B = np.zeros((Br.shape[0], 64 + 2 * G))
    B[:, G:-G] = Br
    ctx = capi.Context(cfg)
    lev = capi.Level(ctx, boxes)
    for k, b in enumerate(boxes):
        lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
    for tpf in tps:
        full = dict(refine_thresh=0.01, plastic_thresh=1e-5, tag_buffer=2, x_coarsen=(-1e40, 1e40),
                    y_coarsen=(-1e40, 1e40), slope_coarsen=0.0, width_coarsen=1e40, x_coarsen2=(-1e40, 1e40),
                    y_coarsen2=(-1e40, 1e40), level=0, coarsen2_max_level=3, stress_thresh=0.0, sliprate_thresh=0.0)
        full.update(tpf)
        want = tmo.tag_level(tmo.OracleParams(cfg), U, B, full)
        got = lev.tag(capi.TagParams.make(**full), 64, 40)
        assert np.array_equal(got, want), (np.sum(got), np.sum(want))
        assert 0 < np.sum(want) < 64 * 40
    lev.close()
    ctx.close()


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_patch_plastic_update_matches_oracle(precision):
    """tmk_patch_plastic_update, the drop-in for LinElastPatchGodunov::plasticUpdate ->
    PLASTICCORRECTION (LinElastPatchGodunov.cpp:161-170, PlasticPhysicsF.ChF:30-69), on a box
    whose shear stresses make about half of its cells yield (Drucker-Prager with cohesion,
    the multi-level workload's parameters), against the oracle's plastic correction.  FP64
    evaluates the same expressions in the same order: bit-identical.  FP32 bulk: within 2e-6
    of each component's largest magnitude.  Ghost cells are not updated."""
    cfg = scenarios.config_c3(seed=3, nx=48, ny=40, dx=25.0)
    valid = (6, 4, 37, 29)
    ubox = (valid[0] - G, valid[1] - G, valid[2] + G, valid[3] + G)
    rng = np.random.default_rng(17)
    U, _ = scenarios.level_arrays(cfg)
    Ub = np.ascontiguousarray(U[:, ubox[1] + G:ubox[3] + G + 1, ubox[0] + G:ubox[2] + G + 1])
    for o in (0, 11):
        for c in (3, 4, 5):
            Ub[o + c] += rng.uniform(-40.0, 40.0, size=Ub.shape[1:])
        Ub[o + 6] += rng.uniform(-150.0, 150.0, size=Ub.shape[1:])
        Ub[o + 10] = rng.uniform(0.0, 1e-3, size=Ub.shape[1:])
    dt = scenarios.dt_of(cfg)
    Ur = Ub.copy()
    op = tmo.OracleParams(cfg)
    Uf = tmo.FabArray(ubox[:2], ubox[2:], 22, data=Ur)
    tmo.lib().tmo_plastic_correction(Uf.fab, op.ref, dt, tmo.Box.make(*valid))
    inside = np.zeros(Ub.shape[1:], dtype=bool)
    inside[G:-G, G:-G] = True
    yielded = (Ur[9] > 0) & inside
    assert 0.3 < yielded.sum() / inside.sum() < 0.9  # yielding and elastic cells
    ctx = capi.Context(cfg, precision)
    Ug = Ub.copy()
    ctx.plastic_update(valid, Ug, ubox, dt)
    ctx.close()
    # ghost cells come back as they went in; the FP32-bulk drop-in holds the patch as its
    # difference from the background in FP32
    bg = scenarios.background_vector(cfg)[:, None, None]
    Ubr = Ub if precision == capi.PREC_FP64 else (Ub - bg).astype(np.float32).astype(np.float64) + bg
    assert np.array_equal(Ug[:, ~inside], Ubr[:, ~inside])
    if precision == capi.PREC_FP64:
        assert np.array_equal(Ug, Ur), np.max(np.abs(Ug - Ur))
    else:
        assert np.array_equal((Ug[9] > 0) & inside, yielded)
        e = _relerr(Ug[:, inside], Ur[:, inside])
        assert np.all(e <= 2e-6), e


# ---------------------------------------------------------------------------
# the row-marching interior update (k_godunov_march, tmk_set_march)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rows", [1, 5, 16, 64])
@pytest.mark.parametrize("maxbox", [None, 24, 40])
def test_march_level_matches_oracle(rows, maxbox):
    """Strips of 58 x rows over the cells whose stencil lies in the domain (one box, or boxes of
    24 / 40 cells merged into island rectangles, with partial strips at their edges), the band
    along the domain boundary in the general kernel: bit-identical to the oracle away from the
    fault's reach, within 1e-12 on it."""
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    boxes = [(0, 0, 199, 119)] if maxbox is None else scenarios.split_boxes(200, 120, maxbox)
    nsteps = 4
    (Ug, Ur), (Bg, Br) = _level_case(cfg, boxes, nsteps, seed=41, march=rows, use_run=rows == 16)
    assert np.all(_relerr(Ug, Ur) <= TOL), _relerr(Ug, Ur)
    assert np.all(fault_errors(Bg, Br) <= TOL), fault_errors(Bg, Br)
    far = 3 * nsteps + 3
    assert np.array_equal(Ug[:, far:, :], Ur[:, far:, :])


@pytest.mark.parametrize("rows", [1, 2, 3, 24])
@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_march_level_plasticity(precision, rows):
    """The Drucker-Prager return fused into the marching update, with strips of 1-3 rows (the LDS
    ring's waits at a strip's start, where its first two rows store nothing) and of 24."""
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    cfg["plasticity"] = 1
    cfg["p_mu"], cfg["p_beta"], cfg["p_eta"], cfg["p_b"] = 0.6, 0.3, 0.03 * scenarios.G_MPA, 1.0
    cfg["rs"]["state_law"] = 1
    cfg["background"] = (-120.0, -120.0, -120.0, 70.0, 0.0, 0.0)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, scenarios.split_boxes(200, 120, 32), 4, seed=16, precision=precision,
                                     march=rows)
    if precision == capi.PREC_FP64:
        assert np.all(_relerr(Ug, Ur) <= TOL), _relerr(Ug, Ur)
        assert np.array_equal(Ug[:, 15:, :], Ur[:, 15:, :])
    else:
        assert np.all(_relerr(Ug, Ur) <= 1e-5), _relerr(Ug, Ur)
    assert np.max(Ur[9]) > 0.0, "plastic flow expected somewhere"


def test_march_fp32_bulk_level():
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, [(0, 0, 199, 119)], 6, seed=42, precision=capi.PREC_FP32_BULK, march=32)
    assert np.all(_relerr(Ug, Ur) <= 1e-5), _relerr(Ug, Ur)
    assert np.all(_relerr(Bg[[6, 8, 10]], Br[[6, 8, 10]]) <= 1e-4), fault_errors(Bg, Br)


@pytest.mark.parametrize("rows", [1, 2, 3])
def test_march_fp32_short_strips(rows):
    """FP32 strips of 1-3 rows over boxes of 40 cells: the LDS ring's waits at the start of a
    strip, where the first two rows of the march store nothing, and strips that end while rows are
    still in flight.  Within the FP32 tolerance of the oracle, and bit-identical to 8-row strips (a
    cell's arithmetic does not depend on the strip height)."""
    cfg = _cfg("rs", "PPM", nx=200, ny=120)
    boxes = scenarios.split_boxes(200, 120, 40)
    kw = dict(precision=capi.PREC_FP32_BULK)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, boxes, 6, seed=45, march=rows, **kw)
    assert np.all(_relerr(Ug, Ur) <= 1e-5), _relerr(Ug, Ur)
    assert np.all(_relerr(Bg[[6, 8, 10]], Br[[6, 8, 10]]) <= 1e-4), fault_errors(Bg, Br)
    (U8, _), (B8, _) = _level_case(cfg, boxes, 6, seed=45, march=8, **kw)
    assert np.array_equal(Ug, U8, equal_nan=True) and np.array_equal(Bg, B8, equal_nan=True)


def test_march_equals_tile_kernels_fp64():
    """The marching update and the tile kernels give the same FP64 state, bit for bit, on a
    slip-weakening level (no transcendental function anywhere)."""
    cfg = _cfg("sw", "PPM", nx=160, ny=96)
    boxes = scenarios.split_boxes(160, 96, 32)
    a = _level_case(cfg, boxes, 5, seed=43)
    b = _level_case(cfg, boxes, 5, seed=43, march=8)
    assert np.array_equal(a[0][0], b[0][0]) and np.array_equal(a[1][0], b[1][0])


def test_march_falls_back_for_other_numerics():
    """PLM (and characteristic limiting, anti-plane fields) keep the tile kernels under
    tmk_set_march: the results match the oracle as before."""
    cfg = _cfg("rs", "PLM", nx=120, ny=80, prim=1, char=0, fourth=1)
    (Ug, Ur), (Bg, Br) = _level_case(cfg, [(0, 0, 119, 79)], 3, seed=44, march=16)
    assert np.all(_relerr(Ug, Ur) <= TOL)
    assert np.all(fault_errors(Bg, Br) <= TOL)


# ---------------------------------------------------------------------------
# update shapes of a level and their autotuning (tmk_level_set_update_shape, tmk_level_autotune)
# ---------------------------------------------------------------------------
def _uploaded_level(cfg, boxes, precision, seed):
    U, B = _state(cfg, seed)
    ctx = capi.Context(cfg, precision)
    lev = capi.Level(ctx, boxes)
    for k, b in enumerate(boxes):
        lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
    return ctx, lev, U, B


def _downloaded(lev, boxes, U, B):
    Ug, Bg = np.full_like(U, np.nan), np.full_like(B, np.nan)
    for k, b in enumerate(boxes):
        ub = scenarios.box_view(U, b)
        bb = scenarios.box_bview(B, b) if b[1] == 0 else None
        lev.download(k, ub, bb)
        x0, y0, x1, y1 = b
        Ug[:, y0 + G:y1 + G + 1, x0 + G:x1 + G + 1] = ub[:, G:-G, G:-G]
        if bb is not None:
            Bg[:, x0 + G:x1 + G + 1] = bb[:, G:-G]
    return Ug[:, G:-G, G:-G], Bg[:, G:-G]


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
@pytest.mark.parametrize("shape", [(8, 0, 0), (0, 32, 16)])
def test_workgroup_order_does_not_change_results(precision, shape):
    """Interior workgroups in XCD bands or in row-major order (tmk_level_set_tile_order) run the
    same arithmetic on the same cells: bit-identical states and fault data after several steps.
    The level is large enough (hundreds of interior strips or tiles) for the orders to differ."""
    cfg = _cfg("sw", "PPM", nx=512, ny=256)
    boxes = [(0, 0, 511, 255)]
    dt = scenarios.dt_of(cfg)
    out = []
    for xcd in (1, 0):
        ctx, lev, U, B = _uploaded_level(cfg, boxes, precision, seed=52)
        lev.set_update_shape(*shape)
        lev.set_tile_order(xcd)
        lev.run(5, 0.0, dt)
        lev.sync()
        out.append(_downloaded(lev, boxes, U, B))
        lev.close()
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_FP32_BULK])
def test_autotune_restores_state_and_sets_fastest(precision):
    """Every strip height and tile shape is timed from the level's state; the level then holds
    that state again, bit for bit, and runs the fastest shape.  FP64: the tuned level steps to
    the same state as an untuned one, bit for bit (every shape evaluates the same expressions;
    slip weakening has no transcendental function whose library code could differ between
    kernels); FP32 bulk: within 1e-5 of each component's largest magnitude."""
    cfg = _cfg("sw", "PPM", nx=200, ny=120)
    boxes = scenarios.split_boxes(200, 120, 40)
    dt = scenarios.dt_of(cfg)
    ctx, lev, U, B = _uploaded_level(cfg, boxes, precision, seed=51)
    before = _downloaded(lev, boxes, U, B)
    r = lev.autotune(0.0, dt, nsteps=3)
    after = _downloaded(lev, boxes, U, B)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    tried = r["tried"]
    assert len(tried) == 21  # 12 strip heights, 8 tile shapes, the fastest in the other order
    assert [t[3] for t in tried] == [1] * 20 + [0]
    assert tried[20][:3] == min(tried[:20], key=lambda t: t[4])[:3]
    assert r["ms_per_step"] == min(t[4] for t in tried) > 0.0
    assert (r["march_rows"], r["bx"], r["by"]) == lev.update_shape()
    assert (r["march_rows"], r["bx"], r["by"], r["xcd"], r["ms_per_step"]) in [tuple(t) for t in tried]
    lev.run(4, 0.0, dt)
    lev.sync()
    tuned = _downloaded(lev, boxes, U, B)
    lev.close()
    ctx.close()
    ctx, lev, U, B = _uploaded_level(cfg, boxes, precision, seed=51)
    lev.set_update_shape(0, 32, 8)
    lev.run(4, 0.0, dt)
    lev.sync()
    plain = _downloaded(lev, boxes, U, B)
    lev.close()
    ctx.close()
    if precision == capi.PREC_FP64:
        assert np.array_equal(tuned[0], plain[0]) and np.array_equal(tuned[1], plain[1])
    else:
        assert np.all(_relerr(tuned[0], plain[0]) <= 1e-5), _relerr(tuned[0], plain[0])


def test_update_shape_arguments():
    """Shapes outside the kernels' set are refused; numerics without the marching update
    report the tile kernels."""
    cfg = _cfg("rs", "PLM", nx=96, ny=64)
    ctx = capi.Context(cfg, capi.PREC_FP64)
    lev = capi.Level(ctx, [(0, 0, 95, 63)])
    with pytest.raises(capi.TmkError):
        lev.set_update_shape(16, 24, 24)
    with pytest.raises(capi.TmkError):
        lev.set_update_shape(-1, 32, 8)
    lev.set_update_shape(16, 64, 8)
    assert lev.update_shape() == (0, 64, 8)  # PLM keeps the tile kernels
    lev.close()
    ctx.close()
