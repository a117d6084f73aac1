"""Multi-level hierarchies on the GPU (tmk_amr_*) against the multi-level oracle.

The coarse-fine fills, registers, reflux and average-down are integer-indexed gathers and
fixed-order sums, and the FP64 level update reproduces the oracle bit for bit off the fault,
so slip-weakening hierarchies (no transcendental functions on the fault) match the oracle
bit for bit on every level: state on every valid cell and fault data on every valid face.
Rate-and-state hierarchies stay within 1e-12 relative (per component, against its largest
magnitude; rupture times per face against their own value, unruptured faces holding the 1e10
sentinel exactly) -- the last-bit differences of exp/log/asinh on fault faces, as on one level.
Runs of 3 and 4 coarse steps cover both the replayed two-step graph and a direct step.
"""
import numpy as np
import pytest

import amr as oamr
import tmo
from faultcheck import SENTINEL, fault_errors
from tetemoko_amd import amr, capi, scenarios

pytestmark = pytest.mark.gpu

G = scenarios.GHOST


def _bump_arrays(cfgs, xc, yc, radius):
    def arrays(l):
        cfg = cfgs[l]
        nx, ny, dx = cfg["nx"], cfg["ny"], cfg["dx"]
        x = (np.arange(-G, nx + G) + 0.5) * dx
        y = (np.arange(-G, ny + G) + 0.5) * dx
        X, Y = np.meshgrid(x, y)
        R = np.hypot(X - xc, Y - yc)
        w = np.where(R < radius, np.cos(0.5 * np.pi * R / radius) ** 2, 0.0)
        U = np.empty((scenarios.NCOMP, ny + 2 * G, nx + 2 * G))
        U[:] = scenarios.background_vector(cfg)[:, None, None]
        for s in (0, 11):
            U[s + 0] += w
            U[s + 1] -= 0.5 * w
            U[s + 3] += 3.0 * w
            U[s + 4] -= 2.0 * w
            U[s + 6] += 1.5 * w
        return U, scenarios.fault_arrays(cfg)
    return arrays


def _background_arrays(cfgs):
    def arrays(l):
        return scenarios.level_arrays(cfgs[l])
    return arrays


def _cfgs(cfg0, nlev, r):
    out = [cfg0]
    for _ in range(1, nlev):
        out.append(scenarios.refine_cfg(out[-1], r))
    return out


def _compare(cfg0, boxes, r, nsteps, arrays, exact, precision=capi.PREC_FP64):
    h = oamr.Hierarchy(cfg0, boxes, r)
    h.init_state(lambda l: arrays(l)[0], lambda l: arrays(l)[1])
    h.set_dt(scenarios.dt_of(cfg0))
    assert h.advance(nsteps) == 0
    H = amr.Hierarchy(cfg0, boxes, r, precision, arrays=arrays)
    H.run(nsteps)
    H.sync()
    for l, lev in enumerate(h.levels):
        Ug, Bg = H.download(l)
        mask = lev.valid
        Ur = lev.U[:, G:-G, G:-G][:, mask]
        Ugv = Ug[:, G:-G, G:-G][:, mask]
        fmask = lev.fvalid
        Br = lev.B[:, G:-G][:, fmask]
        Bgv = Bg[:, G:-G][:, fmask]
        if exact:
            assert np.array_equal(Ugv, Ur), (l, np.nanmax(np.abs(Ugv - Ur)))
            assert np.array_equal(Bgv, Br), l
        else:
            for c in range(scenarios.NCOMP):
                scale = max(np.max(np.abs(Ur[c])), 1e-300)
                assert np.max(np.abs(Ugv[c] - Ur[c])) <= 1e-12 * scale, (l, c)
            e = fault_errors(Bgv, Br)
            assert np.all(e <= 1e-12), (l, e)
    H.close()
    return h


# level 0: 48 x 32 in two boxes; level 1 (ratio 2) two boxes on the fault; level 2 one box
BOXES_3 = [[(0, 0, 23, 31), (24, 0, 47, 31)],
           [(24, 0, 55, 23), (56, 0, 71, 23)],
           [(64, 0, 111, 15)]]


@pytest.mark.parametrize("nsteps", [3, 4])
def test_slip_weakening_rupture_three_levels_bit_identical(nsteps):
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    h = _compare(cfg, BOXES_3, 2, nsteps, _background_arrays(_cfgs(cfg, 3, 2)), exact=True)
    # the nucleation patch slipped on every level
    for lev in h.levels:
        assert np.nanmax(lev.B[7, G:-G][lev.fvalid]) > 0.0


def test_disturbance_across_coarse_fine_interface_bit_identical():
    cfg = scenarios.config_c0(seed=3, nx=48, ny=32, dx=100.0)
    cfg["faces"]["tau"] = np.full(48, 20.0)
    boxes = [[(0, 0, 47, 31)], [(48, 20, 71, 43)]]
    cfgs = _cfgs(cfg, 2, 2)
    h = _compare(cfg, boxes, 2, 3, _bump_arrays(cfgs, 2400.0, 1600.0, 300.0), exact=True)
    assert np.abs(h.regs[0].reg).max() > 0.0


def test_ratio_four_with_plasticity_bit_identical():
    cfg = scenarios.config_c0(seed=4, nx=48, ny=32, dx=100.0)
    cfg["plasticity"] = 1
    cfg["p_mu"], cfg["p_beta"], cfg["p_eta"], cfg["p_b"] = 0.6, 0.3, 0.03 * scenarios.G_MPA, 1.0
    boxes = [[(0, 0, 47, 31)], [(64, 0, 127, 31)]]
    _compare(cfg, boxes, 4, 3, _background_arrays(_cfgs(cfg, 2, 4)), exact=True)


@pytest.mark.parametrize("nsteps", [3, 4])
def test_rate_and_state_three_levels_within_tolerance(nsteps):
    cfg = scenarios.config_c1(seed=5, nx=48, ny=32, dx=25.0)
    cfg["rs"]["nuc_R"] = 400.0
    cfg["rs"]["rupture_thr"] = 1e-6  # faces of every level rupture inside the window
    h = _compare(cfg, BOXES_3, 2, nsteps, _background_arrays(_cfgs(cfg, 3, 2)), exact=False)
    for lev in h.levels:
        assert np.count_nonzero(lev.B[9, G:-G][lev.fvalid] != SENTINEL) >= 8


def _composite_sum(H, cfgs, boxes, r):
    tot = np.zeros(scenarios.NCOMP)
    for l in range(len(boxes)):
        U, _ = H.download(l)
        cfg = cfgs[l]
        mask = ~np.isnan(U[0, G:-G, G:-G])
        if l + 1 < len(boxes):
            for b in boxes[l + 1]:
                mask[b[1] // r:b[3] // r + 1, b[0] // r:b[2] // r + 1] = False
        vals = U[:, G:-G, G:-G][:, mask] - scenarios.background_vector(cfg)[:, None]
        tot += vals.sum(axis=1) * cfg["dx"] ** 2
    return tot


def test_composite_sum_conserved_on_device():
    cfg = scenarios.config_c0(seed=3, nx=48, ny=32, dx=100.0)
    cfg["faces"]["tau"] = np.full(48, 20.0)
    boxes = [[(0, 0, 47, 31)], [(40, 16, 71, 47)], [(96, 48, 127, 79)]]
    cfgs = _cfgs(cfg, 3, 2)
    H = amr.Hierarchy(cfg, boxes, 2, arrays=_bump_arrays(cfgs, 2400.0, 1600.0, 300.0))
    s0 = _composite_sum(H, cfgs, boxes, 2)
    H.run(4)
    H.sync()
    s1 = _composite_sum(H, cfgs, boxes, 2)
    H.close()
    assert np.abs(s1 - s0).max() <= 1e-12 * np.abs(s0).max()


def test_fp32_bulk_hierarchy_close_to_fp64():
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    arrays = _background_arrays(_cfgs(cfg, 3, 2))
    out = {}
    for prec in (capi.PREC_FP64, capi.PREC_FP32_BULK):
        H = amr.Hierarchy(cfg, BOXES_3, 2, prec, arrays=arrays)
        H.run(6)
        H.sync()
        out[prec] = [H.download(l) for l in range(3)]
        H.close()
    for l in range(3):
        B64, B32 = out[capi.PREC_FP64][l][1], out[capi.PREC_FP32_BULK][l][1]
        ok = ~np.isnan(B64[7])
        slip64, slip32 = B64[7][ok], B32[7][ok]
        assert np.max(np.abs(slip32 - slip64)) <= 1e-4 * np.max(np.abs(slip64)) + 1e-12


def test_hierarchy_rejects_misaligned_and_unnested_levels():
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    for boxes in ([[(0, 0, 47, 31)], [(25, 0, 56, 23)]],          # not aligned to the ratio
                  [[(0, 0, 47, 31)], [(24, 0, 55, 23)], [(40, 0, 63, 15)]]):  # level 2 reads outside level 1
        with pytest.raises(capi.TmkError) as ei:
            amr.Hierarchy(cfg, boxes, 2)
        assert ei.value.code == -12


def test_levels_of_a_hierarchy_refuse_direct_steps():
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    H = amr.Hierarchy(cfg, [[(0, 0, 47, 31)], [(24, 0, 55, 23)]], 2)
    with pytest.raises(capi.TmkError):
        H.levels[1].run(1, 0.0, 1e-3)
    H.run(1)
    H.sync()
    H.close()


def _levels_equal(h, H, exact=True, tol=1e-12):
    """Every level of the device hierarchy H against the oracle's h; returns the largest relative
    state error (per component, against its largest magnitude) over the levels."""
    worst = 0.0
    for l, lev in enumerate(h.levels):
        Ug, Bg = H.download(l)
        mask = lev.valid
        Ur, Ugv = lev.U[:, G:-G, G:-G][:, mask], Ug[:, G:-G, G:-G][:, mask]
        Br, Bgv = lev.B[:, G:-G][:, lev.fvalid], Bg[:, G:-G][:, lev.fvalid]
        if exact:
            assert np.array_equal(Ugv, Ur), (l, np.nanmax(np.abs(Ugv - Ur)))
            assert np.array_equal(Bgv, Br), l
        else:
            for c in range(scenarios.NCOMP):
                rel = np.max(np.abs(Ugv[c] - Ur[c])) / max(np.max(np.abs(Ur[c])), 1e-300)
                worst = max(worst, rel)
                assert rel <= tol, (l, c, rel)
            e = fault_errors(Bgv, Br)
            assert np.all(e <= tol), (l, e)
    return worst


def _pair(cfg, boxes, r, arrays):
    h = oamr.Hierarchy(cfg, boxes, r)
    h.init_state(lambda l: arrays(l)[0], lambda l: arrays(l)[1])
    h.set_dt(scenarios.dt_of(cfg))
    H = amr.Hierarchy(cfg, boxes, r, arrays=arrays)
    return h, H


def test_new_layout_of_a_level_matches_oracle():
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    cfgs = _cfgs(cfg, 3, 2)
    arrays = _background_arrays(cfgs)
    h, H = _pair(cfg, BOXES_3, 2, arrays)
    assert h.advance(3) == 0
    H.run(3)
    H.sync()
    # level 1 moves: part of it stays, part is new (interpolated from level 0); level 2 stays nested
    new1 = [(16, 0, 47, 15), (48, 0, 79, 31)]
    h.set_layout(1, new1, *scenarios.level_arrays(cfgs[1]))
    H.set_layout(1, new1)
    _levels_equal(h, H)
    new2 = [(80, 0, 143, 23)]
    h.set_layout(2, new2, *scenarios.level_arrays(cfgs[2]))
    H.set_layout(2, new2)
    _levels_equal(h, H)
    assert h.advance(3) == 0
    H.run(3)
    H.sync()
    _levels_equal(h, H)
    H.close()


def test_regrid_from_device_tags_matches_oracle():
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    cfgs = _cfgs(cfg, 3, 2)
    arrays = _background_arrays(cfgs)
    h, H = _pair(cfg, BOXES_3, 2, arrays)
    assert h.advance(4) == 0
    H.run(4)
    H.sync()
    counts = H.regrid(lambda l: capi.TagParams.make(refine_thresh=1e-4, tag_buffer=2, level=l),
                      fill_ratio=0.7, block_factor=4, max_grid_size=32)
    assert counts[1] > 0 and counts[2] > 0
    for l in (1, 2):
        h.set_layout(l, H.boxes[l], *scenarios.level_arrays(cfgs[l]))
    _levels_equal(h, H)
    # the nucleation patch is refined
    xc = cfg["nx"] // 2 * 4
    assert any(b[0] <= xc <= b[2] and b[1] == 0 for b in H.boxes[2])
    assert h.advance(2) == 0
    H.run(2)
    H.sync()
    _levels_equal(h, H)
    H.close()


def _download_levels(levels, boxes_per_level, cfgs):
    out = []
    for l, lev in enumerate(levels):
        cfg = cfgs[l]
        U = np.full((scenarios.NCOMP, cfg["ny"] + 2 * G, cfg["nx"] + 2 * G), np.nan)
        B = np.full((scenarios.NB[cfg["fault"]], cfg["nx"] + 2 * G), np.nan)
        for k, b in enumerate(boxes_per_level[l]):
            x0, y0, x1, y1 = b
            ub = np.empty((scenarios.NCOMP, y1 - y0 + 1 + 2 * G, x1 - x0 + 1 + 2 * G))
            bb = np.empty((B.shape[0], x1 - x0 + 1 + 2 * G)) if y0 == 0 else None
            lev.download(k, ub, bb)
            U[:, y0 + G:y1 + G + 1, x0 + G:x1 + G + 1] = ub[:, G:-G, G:-G]
            if bb is not None:
                B[:, x0 + G:x1 + G + 1] = bb[:, G:-G]
        out.append((U, B))
    return out


def test_hierarchy_batch_members_equal_alone():
    """Two scenarios of different size and friction law in one hierarchy batch, joining and
    leaving at different steps, each bit-identical to its hierarchy stepped alone."""
    cfgA = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    cfgB = scenarios.config_c0(seed=9, nx=64, ny=32, dx=50.0)
    boxA = [[(0, 0, 47, 31)], [(24, 0, 55, 23), (56, 0, 71, 23)]]
    boxB = [[(0, 0, 63, 31)], [(32, 0, 95, 15)]]
    ctxA, levA = amr.make_levels(cfgA, boxA, 2)
    ctxB, levB = amr.make_levels(cfgB, boxB, 2)
    ab = capi.AmrBatch(2, 2)
    ab.add(levA, 0.0, scenarios.dt_of(cfgA))
    ab.advance(3)
    ab.add(levB, 0.0, scenarios.dt_of(cfgB))
    ab.advance(4)
    ab.sync()
    ab.remove(levA[0])         # A: 7 coarse steps
    ab.advance(2)              # B: 6 coarse steps
    ab.sync()
    ab.remove(levB[0])
    ab.close()
    for cfg, boxes, levels, nsteps in ((cfgA, boxA, levA, 7), (cfgB, boxB, levB, 6)):
        cfgs = _cfgs(cfg, 2, 2)
        got = _download_levels(levels, boxes, cfgs)
        H = amr.Hierarchy(cfg, boxes, 2)
        H.run(nsteps)
        H.sync()
        for l in range(2):
            Ua, Ba = H.download(l)
            assert np.array_equal(np.nan_to_num(got[l][0], nan=7.0), np.nan_to_num(Ua, nan=7.0)), l
            assert np.array_equal(np.nan_to_num(got[l][1], nan=7.0), np.nan_to_num(Ba, nan=7.0)), l
        H.close()
    for lev in levA + levB:
        lev.close()
    for ctx in ctxA + ctxB:
        ctx.close()


def test_hierarchy_batch_member_error_isolated():
    """A member whose fine-level fault data holds a NaN psi reports its own error; the other
    member is bit-identical to running alone and the batch carries on once the failed one
    leaves."""
    good = scenarios.config_c1(seed=5, nx=48, ny=32, dx=25.0)
    bad = scenarios.config_c1(seed=6, nx=48, ny=32, dx=25.0)
    boxes = [[(0, 0, 47, 31)], [(24, 0, 71, 23)]]
    ctxg, levg = amr.make_levels(good, boxes, 2)
    ctxb, levb = amr.make_levels(bad, boxes, 2)
    cfgb1 = scenarios.refine_cfg(bad, 2)
    U1, B1 = scenarios.level_arrays(cfgb1)
    B1[10, G + 40] = np.nan
    levb[1].upload(0, scenarios.box_view(U1, boxes[1][0]), scenarios.box_bview(B1, boxes[1][0]))
    ab = capi.AmrBatch(2, 2)
    ab.add(levg, 0.0, scenarios.dt_of(good))
    ab.add(levb, 0.0, scenarios.dt_of(bad))
    ab.advance(3)
    with pytest.raises(capi.TmkError) as ei:
        ab.sync()
    assert ei.value.code == capi.E_PSI_NAN
    assert ab.member_error(levb[0]) == capi.E_PSI_NAN
    assert ab.member_error(levg[0]) == 0
    ab.remove(levb[0])
    ab.advance(2)
    ab.sync()
    ab.remove(levg[0])
    ab.close()
    got = _download_levels(levg, boxes, _cfgs(good, 2, 2))
    H = amr.Hierarchy(good, boxes, 2)
    H.run(5)
    H.sync()
    for l in range(2):
        Ua, Ba = H.download(l)
        assert np.array_equal(np.nan_to_num(got[l][0], nan=7.0), np.nan_to_num(Ua, nan=7.0)), l
        assert np.array_equal(np.nan_to_num(got[l][1], nan=7.0), np.nan_to_num(Ba, nan=7.0)), l
    H.close()
    for o in levg + levb + ctxg + ctxb:
        o.close()


def test_boundary_plot_file_of_a_hierarchy(tmp_path):
    """The boundary plot file written from a device hierarchy (tetemoko_amd.bdry_plot) holds the
    oracle's fault data bit for bit on every valid fault face of every level (slip-weakening
    rupture, three levels), and a box's x-ghost entry equals its neighbour's valid face."""
    import os
    from h5dump_util import H5DUMP, h5_attr, h5_data, h5_ints
    from tetemoko_amd import bdry_plot
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    cfgs = _cfgs(cfg, 3, 2)
    arrays = _background_arrays(cfgs)
    h = oamr.Hierarchy(cfg, BOXES_3, 2)
    h.init_state(lambda l: arrays(l)[0], lambda l: arrays(l)[1])
    h.set_dt(scenarios.dt_of(cfg))
    assert h.advance(3) == 0
    H = amr.Hierarchy(cfg, BOXES_3, 2, arrays=arrays)
    H.run(3)
    path = bdry_plot.write_hierarchy(str(tmp_path / "run."), H)
    assert path.endswith("run.boundary.000003.2d.hdf5")
    assert h5_attr(path, "/iteration") == 3 and h5_attr(path, "/num_levels") == 3
    assert h5_attr(path, "/time") == H.time
    for l, lev in enumerate(h.levels):
        g = "/level_%d" % l
        assert h5_attr(path, g + "/dx") == cfgs[l]["dx"]
        assert h5_attr(path, g + "/dt") == H.dt0 / 2 ** l
        boxes = np.array(h5_ints(path, g + "/boxes")).reshape(-1, 4)
        assert [tuple(b) for b in boxes] == sorted((b[0], 0, b[2], 0) for b in BOXES_3[l] if b[1] == 0)
        offs = h5_data(path, g + "/data:offsets=0", "<i8", tmp_path)
        data = h5_data(path, g + "/data:datatype=0", "<f8", tmp_path)
        nb = lev.B.shape[0]
        covered = set()
        for b in boxes:
            covered.update(range(b[0], b[2] + 1))
        for k, b in enumerate(boxes):
            n = b[2] - b[0] + 1
            blk = data[offs[k]:offs[k + 1]].reshape(nb, n + 2)
            assert np.array_equal(blk[:, 1:-1], lev.B[:, G + b[0]:G + b[2] + 1]), (l, k)
            for col, i in ((0, b[0] - 1), (-1, b[2] + 1)):
                if i in covered:
                    assert np.array_equal(blk[:, col], lev.B[:, G + i]), (l, k, i)
    H.close()


def test_dt_schedule_ramp_matches_oracle():
    """A run whose coarse dt ramps from initial_cfl by max_dt_growth to the CFL dt
    (amr.dt_schedule) equals the oracle driven by the same dt sequence, bit for bit on every
    level (slip-weakening, three levels).  Consecutive device calls (one per run of equal dt)
    continue the clocks, and a change of dt keeps every level's clock running, as the
    oracle's change_dt does."""
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    arrays = _background_arrays(_cfgs(cfg, 3, 2))
    dts = amr.dt_schedule(6, cfg, cfl=0.9, initial_cfl=0.3, max_dt_growth=1.5)
    assert len(set(dts)) == 4
    h = oamr.Hierarchy(cfg, BOXES_3, 2)
    h.init_state(lambda l: arrays(l)[0], lambda l: arrays(l)[1])
    h.set_dt(dts[0])
    t = 0.0
    for k, dt in enumerate(dts):
        if k and dt != dts[k - 1]:
            h.change_dt(dt)
        assert h.advance(1) == 0
        t = t + dt
    H = amr.Hierarchy(cfg, BOXES_3, 2, arrays=arrays)
    H.run_schedule(dts)
    H.sync()
    assert H.time == t == h.levels[0].time
    for l, lev in enumerate(h.levels):
        Ug, Bg = H.download(l)
        assert np.array_equal(Ug[:, G:-G, G:-G][:, lev.valid], lev.U[:, G:-G, G:-G][:, lev.valid]), l
        assert np.array_equal(Bg[:, G:-G][:, lev.fvalid], lev.B[:, G:-G][:, lev.fvalid]), l
    H.close()


def test_regrid_without_tags_removes_finer_levels_and_tags_restore_them():
    """As Chombo's AMR::regrid: a level without tags ends the hierarchy (its finer levels are
    removed); a later regrid whose tags reach down again re-creates them, properly nested, and
    the hierarchy steps on."""
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    arrays = _background_arrays(_cfgs(cfg, 3, 2))
    H = amr.Hierarchy(cfg, BOXES_3, 2, arrays=arrays)
    H.run(4)

    def tp(active):
        return lambda l: capi.TagParams.make(refine_thresh=1e-4 if l in active else 1e30,
                                             plastic_thresh=1e-5 if l in active else 1e30, tag_buffer=2, level=l)

    counts = H.regrid(tp({0}), fill_ratio=0.7, block_factor=4, max_grid_size=32)
    assert len(counts) == 2 and counts[1] > 0 and len(H.levels) == 2
    H.run(2)
    counts = H.regrid(tp(set()), fill_ratio=0.7, block_factor=4, max_grid_size=32)
    assert counts == [2] and len(H.levels) == 1
    H.run(2)
    counts = H.regrid(tp({0, 1}), fill_ratio=0.7, block_factor=4, max_grid_size=32)
    assert len(counts) == 3 and counts[1] > 0 and counts[2] > 0
    for l in (1, 2):  # every fine box lies inside the next coarser level
        cover = set()
        for b in H.boxes[l - 1]:
            cover.update((i, j) for i in range(b[0], b[2] + 1) for j in range(b[1], b[3] + 1))
        for b in H.boxes[l]:
            assert all((i // 2, j // 2) in cover for i in (b[0], b[2]) for j in (b[1], b[3])), (l, b)
    H.run(2)
    H.sync()
    for l in range(3):
        Ug, _ = H.download(l)
        assert np.isfinite(Ug[:, G:-G, G:-G][~np.isnan(Ug[:, G:-G, G:-G])]).all()
    H.close()


# level 2 in three separate islands (boxes whose ghost regions do not meet)
BOXES_ISL = [[(0, 0, 23, 31), (24, 0, 47, 31)],
             [(16, 0, 79, 31)],
             [(40, 0, 55, 15), (72, 0, 87, 15), (120, 0, 135, 15)]]


def test_separate_fine_islands_build_every_time_and_match_oracle():
    """A finest level of several separate islands: the hierarchy's lists build every time (their
    guard reads the flux-array table per direction, level and island), and a slip-weakening run
    matches the oracle bit for bit on every level."""
    cfg = scenarios.config_c0(seed=2, nx=48, ny=32, dx=100.0)
    arrays = _background_arrays(_cfgs(cfg, 3, 2))
    for _ in range(8):
        H = amr.Hierarchy(cfg, BOXES_ISL, 2, arrays=arrays)
        H.close()
    _compare(cfg, BOXES_ISL, 2, 3, arrays, exact=True)


def test_c2_three_levels_ratio_four_rate_and_state_plastic_tagged_regrid():
    """The multi-level workload on a small base: three levels at ratio 4 (dx 100, 25, 6.25 m),
    rate-and-state with the aging law and Drucker-Prager viscoplasticity (cohesion 1 MPa), tags
    on the velocity gradient, the stress gradient and the fault slip rate, and a regrid between
    steps (AMRLevelLinElast.cpp:577-717, 784-949).  The bulk background carries a deviatoric
    stress just below the yield stress (sxx, szz = -120 -+ 72.95 MPa, leaving the fault's normal
    and shear traction as they are), so that the first waves make cells yield on every level
    inside the short window.  Level 0's device tags equal the oracle's bit for bit.  Every level
    stays within 1e-10 of the oracle before the regrid, right after it and two coarse steps later:
    the last-bit differences of exp/log/asinh on fault faces (1e-16 relative) grow through the
    plastic return, whose Y = tau - Y_s cancels about three digits with the bulk this close to
    yield (measured: 3.2e-12, 3.7e-12 and 5.0e-12 at the three comparisons)."""
    cfg = scenarios.config_c3(seed=6, nx=64, ny=32, dx=100.0)
    cfg["rs"]["rupture_thr"] = 1e-6  # faces of every level rupture inside the window
    cfg["background"] = (-120.0 + 72.95, -120.0, -120.0 - 72.95, 0.0, 0.0, 0.0)
    r = 4
    cfgs = _cfgs(cfg, 3, r)
    boxes = [[(0, 0, 31, 31), (32, 0, 63, 31)], [(96, 0, 159, 31)], [(448, 0, 575, 31)]]
    h, H = _pair(cfg, boxes, r, _background_arrays(cfgs))
    assert h.advance(2) == 0
    H.run(2)
    H.sync()
    errs = [_levels_equal(h, H, exact=False, tol=1e-10)]
    crit = dict(refine_thresh=0.05, stress_thresh=0.02, sliprate_thresh=1e-4, tag_buffer=2)
    full = dict(refine_thresh=0.01, plastic_thresh=1e-5, tag_buffer=2, x_coarsen=(-1e40, 1e40),
                y_coarsen=(-1e40, 1e40), slope_coarsen=0.0, width_coarsen=1e40, x_coarsen2=(-1e40, 1e40),
                y_coarsen2=(-1e40, 1e40), level=0, coarsen2_max_level=3, stress_thresh=0.0, sliprate_thresh=0.0)
    full.update(crit)
    want = tmo.tag_level(h.levels[0].op, h.levels[0].U, h.levels[0].B, full)
    H.amr.close()  # a level tags on its own, outside the hierarchy (as Hierarchy.regrid does)
    got = H.levels[0].tag(capi.TagParams.make(**full), cfg["nx"], cfg["ny"])
    assert np.array_equal(got, want) and want.any()
    counts = H.regrid(lambda l: capi.TagParams.make(level=l, **crit), fill_ratio=0.7, block_factor=4,
                      max_grid_size=32)
    assert len(counts) == 3 and counts[1] >= 2 and counts[2] >= 4, counts
    for l in (1, 2):
        h.set_layout(l, H.boxes[l], *scenarios.level_arrays(cfgs[l]))
    errs.append(_levels_equal(h, H, exact=False, tol=1e-10))
    assert h.advance(2) == 0
    H.run(2)
    H.sync()
    errs.append(_levels_equal(h, H, exact=False, tol=1e-10))
    print("largest relative state error: before the regrid, after it, two steps later:", errs, counts)
    for lev in h.levels:
        U = lev.U[:, G:-G, G:-G]
        assert np.count_nonzero((U[9] > 0) & lev.valid) + np.count_nonzero((U[20] > 0) & lev.valid) > 0
        assert np.count_nonzero(lev.B[9, G:-G][lev.fvalid] != SENTINEL) >= 8
    H.close()
