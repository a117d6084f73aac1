"""Whole configurations on the GPU against the oracle (seeded inputs, many steps).

C0 (slip weakening, 400x200, dx = 100 m, 1500 steps at CFL 0.9, boxes of 32x32 as the
reference's default max_grid_size): slip weakening involves no transcendental function, so the
FP64 level path reproduces the oracle bit for bit over the whole run -- state, slip and rupture
time on every face.  The oracle runs box-parallel on the host's cores (boxes of 32x32; results
do not depend on the box layout).

C1 reduced to 256x128 (rate-and-state, aging law, dx = 25 m, 400 steps): exp/log/asinh of the
device math library and of glibc differ in the last bit on fault faces; the run stays within
1e-12 relative of the oracle on state, slip and psi, and of each ruptured face's own rupture
time (unruptured faces hold the 1e10 sentinel exactly).  The same run in FP32 bulk is
bit-identical between strip heights and within 1e-3 of the FP64 oracle.
"""
import os

import numpy as np
import pytest

import tmo
from faultcheck import fault_errors, new_ruptures
from tetemoko_amd import capi, scenarios

pytestmark = pytest.mark.gpu

G = scenarios.GHOST
NTHREADS = min(16, os.cpu_count() or 1)


def _run_level(cfg, boxes, nsteps, U, B, precision=capi.PREC_FP64, march=None):
    ctx = capi.Context(cfg, precision)
    if march is not None:
        ctx.set_march(march)
    lev = capi.Level(ctx, boxes)
    for k, b in enumerate(boxes):
        lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(B, b) if b[1] == 0 else None)
    lev.run(nsteps, 0.0, scenarios.dt_of(cfg))
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
    return Ug[:, G:-G, G:-G], Bg[:, G:-G]


def test_c0_full_run_bit_identical():
    cfg = scenarios.config_c0(seed=0)
    nsteps = cfg["steps"]
    U, B = scenarios.level_arrays(cfg)
    Ur, Br = U.copy(), B.copy()
    assert tmo.level_steps(tmo.OracleParams(cfg), Ur, Br, 0.0, scenarios.dt_of(cfg), nsteps,
                           nthreads=NTHREADS, maxbox=32) == 0
    Ug, Bg = _run_level(cfg, scenarios.split_boxes(cfg["nx"], cfg["ny"], 32), nsteps, U, B)
    Ur, Br = Ur[:, G:-G, G:-G], Br[:, G:-G]
    # with the configuration's background shear (0.55-0.60 of static strength, below the dynamic
    # strength) the rupture stays inside the 3 km nucleation patch: 30 faces at dx = 100 m
    ruptured = Br[8] < 1e9
    assert ruptured.sum() >= 25
    assert np.max(Br[7]) > 0.1
    assert np.array_equal(Ug, Ur)
    assert np.array_equal(Bg, Br)


def test_c1_reduced_long_run_within_tolerance():
    cfg = scenarios.config_c1(seed=0, nx=256, ny=128, dx=25.0)
    nsteps = 400
    U, B = scenarios.level_arrays(cfg)
    Ur, Br, B0 = U.copy(), B.copy(), B.copy()
    assert tmo.level_steps(tmo.OracleParams(cfg), Ur, Br, 0.0, scenarios.dt_of(cfg), nsteps,
                           nthreads=NTHREADS, maxbox=32) == 0
    Ug, Bg = _run_level(cfg, [(0, 0, 255, 127)], nsteps, U, B)
    Ur, Br = Ur[:, G:-G, G:-G], Br[:, G:-G]
    for c in range(22):
        scale = max(np.max(np.abs(Ur[c])), 1e-300)
        assert np.max(np.abs(Ug[c] - Ur[c])) <= 1e-12 * scale, c
    e = fault_errors(Bg, Br)
    assert np.all(e[[8, 9, 10]] <= 1e-12), e  # slip, rupture time (per face), psi
    assert np.max(Br[2]) > 1e-9  # the nucleation has started slipping
    assert new_ruptures(B0[:, G:-G], Br) >= 20  # and ruptured faces inside the window


def test_c1_reduced_long_run_fp32_bulk():
    """The same 400 steps in FP32 bulk (the row march without plasticity, its LDS ring of 4
    slots): with strips of 8 and of 16 rows the results are bit-identical (a row read from a
    slot before its loads landed would differ between the two), and within 1e-3 of the FP64
    oracle on state, slip and psi."""
    cfg = scenarios.config_c1(seed=0, nx=256, ny=128, dx=25.0)
    nsteps = 400
    U, B = scenarios.level_arrays(cfg)
    Ur, Br = U.copy(), B.copy()
    assert tmo.level_steps(tmo.OracleParams(cfg), Ur, Br, 0.0, scenarios.dt_of(cfg), nsteps,
                           nthreads=NTHREADS, maxbox=32) == 0
    Ur, Br = Ur[:, G:-G, G:-G], Br[:, G:-G]
    boxes = scenarios.split_boxes(256, 128, 32)
    runs = [_run_level(cfg, boxes, nsteps, U.copy(), B.copy(), capi.PREC_FP32_BULK, rows) for rows in (8, 16)]
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    Ug, Bg = runs[0]
    for c in range(22):
        scale = max(np.max(np.abs(Ur[c])), 1e-300)
        assert np.max(np.abs(Ug[c] - Ur[c])) <= 1e-3 * scale, (c, np.max(np.abs(Ug[c] - Ur[c])) / scale)
    for c in (8, 10):  # slip, psi
        scale = max(np.max(np.abs(Br[c])), 1e-300)
        assert np.max(np.abs(Bg[c] - Br[c])) <= 1e-3 * scale, (c, np.max(np.abs(Bg[c] - Br[c])) / scale)
