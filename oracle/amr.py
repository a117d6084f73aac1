"""CPU restatement of the multi-level (AMR) step: coarse-fine ghost fill, flux
registers with reflux, average-down and the subcycled advance.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP AMR path
(tmk_amr_*).  The product package never imports it.

Each level keeps its state as one array over its whole domain plus a 4-cell
frame (22, ny+8, nx+8), and its fault data as (nb, nx+8).  Only the cells of
the level's boxes are real.  A box's ghost cells inside the domain are either
valid cells of another box of the level (the reference's exchange) or filled
from the next coarser level.  The per-box update is the C oracle
(tmo_update_state, the restatement of LinElastPatchGodunov::updateState) on a
ghosted copy of the box, exactly as the single-level oracle driver does.

What each part follows:
  * subcycled advance: Chombo AMR::timeStep (advance level l, then ref_ratio
    steps of level l+1, then postTimeStep of level l), with
    AMRLevelLinElast::advance (AMRLevelLinElast.cpp:372-485) and
    AMRLevelLinElast::postTimeStep (:488-516: reflux with scale -1/dx, then
    average-down of U and of the fault data).
  * the level step: LinElastLevelGodunov::step (LinElastLevelGodunov.cpp:174-398):
    exchange, coarse-fine fill at alpha = (t - tCoarseOld) / (tCoarseNew -
    tCoarseOld) snapped to 0 or 1 within 0.04 dt / ref_ratio (:234-258), box
    updates, flux-register increments (:334-355).
  * fault-data fill: D1PiecewiseLinearFillPatch (D1PiecewiseLinearFillPatch.cpp:
    define :74-278, timeInterp :328-395, fillConstantInterp :398-422,
    computeSimpleSlopesFab :458-517, incrementLinearInterp :558-590).  Its
    limiter box (computeMultiDimSlopes :519-556) is grown and shrunk in every
    direction, so for the one-face-thick fault box it is empty: the fault-data
    slopes are not limited.
  * state fill: Chombo's PiecewiseLinearFillPatch, which is not available here.
    It is restated as the two-dimensional form of the same steps (the D1 class
    is the reference's one-dimensional copy of it): time interpolation,
    piecewise-constant fill, central slopes where both coarse neighbours exist
    and one-sided slopes where one is outside the coarse level, the
    multidimensional limiter (3x3 neighbourhood max/min against half the sum of
    the absolute slopes, skipped on the domain's edge layer of coarse cells,
    as the grow(1)/intersect/grow(-1) box of computeMultiDimSlopes does), then
    the linear increments per direction.  Parity unpinned against Chombo;
    pinned by properties (constants and linear fields reproduced exactly,
    values bounded by the coarse neighbourhood where limited).
  * flux registers: Chombo's LevelFluxRegister, not available here; restated
    from its definition: for each coarse face between an uncovered coarse cell
    and the refined region, reg = -dt_c F_c + sum over fine substeps and the
    ref_ratio fine faces of (dt_f / ref_ratio) F_f, and reflux adds
    (-1/dx_c) * sgn * reg to the uncovered cell (sgn = +1 on its hi face,
    -1 on its lo face).  Pinned by conservation of the composite sum.
  * average-down: CoarseAverage (arithmetic mean of the ref_ratio^2 fine cells,
    summed i fastest then scaled by 1/ref_ratio^2) and D1CoarseAverage
    (D1CoarseAverage.cpp:117-139; D1AVERAGE in tetemoko_chombo_patchfile:142-184).
"""
import numpy as np

import tmo

G = 4
NC = 22


def grow(b, n):
    return (b[0] - n, b[1] - n, b[2] + n, b[3] + n)


def coarsen_box(b, r):
    return (b[0] // r, b[1] // r, b[2] // r, b[3] // r)


def refine_cfg(cfg, r):
    """Level configuration refined by r: cells r x, spacing / r, per-face arrays injected."""
    out = dict(cfg)
    out["nx"], out["ny"], out["dx"] = cfg["nx"] * r, cfg["ny"] * r, cfg["dx"] / r
    faces = cfg.get("faces")
    if faces:
        nf = {}
        for k, v in faces.items():
            nf[k] = v * r if k == "lo" else np.repeat(np.asarray(v), r)
        out["faces"] = nf
    return out


class Level:
    def __init__(self, cfg, boxes, r_coarser=None):
        self.cfg = cfg
        self.op = tmo.OracleParams(cfg)
        self.nx, self.ny, self.dx = cfg["nx"], cfg["ny"], cfg["dx"]
        self.nb = tmo.NB_RS if cfg.get("fault", "rs") == "rs" else tmo.NB_SW
        self.boxes = [tuple(b) for b in boxes]
        self.r = r_coarser
        self.U = np.zeros((NC, self.ny + 2 * G, self.nx + 2 * G))
        self.B = np.zeros((self.nb, self.nx + 2 * G))
        self.Uold = None
        self.Bold = None
        self.time = 0.0
        self.dt = 0.0
        self.valid = np.zeros((self.ny, self.nx), dtype=bool)
        for b in self.boxes:
            self.valid[b[1]:b[3] + 1, b[0]:b[2] + 1] = True
        self.fvalid = np.zeros(self.nx, dtype=bool)   # fault faces of boxes touching j = 0
        for b in self.boxes:
            if b[1] == 0:
                self.fvalid[b[0]:b[2] + 1] = True

    # index helpers: level arrays carry a G-cell frame
    def u_at(self, U, i, j):
        return U[:, j + G, i + G]


def _time_interp(old, new, alpha):
    """D1PiecewiseLinearFillPatch::timeInterp (:328-395): new*alpha + old*(1-alpha)."""
    if alpha == 1.0:
        return new.copy()
    if alpha == 0.0:
        return old.copy()
    return new * alpha + old * (1.0 - alpha)


def fill_state(fine, coarse, alpha):
    """Coarse-fine ghost fill of the fine level state (2-D piecewise-linear fill patch)."""
    r = fine.r
    ncx, ncy = coarse.nx, coarse.ny
    cov_c = coarse.valid                                   # cells of the coarse level
    cov_f = np.zeros((ncy, ncx), dtype=bool)               # coarse cells under the fine level
    for b in fine.boxes:
        cb = coarsen_box(b, r)
        cov_f[cb[1]:cb[3] + 1, cb[0]:cb[2] + 1] = True
    Ct = _time_interp(coarse.Uold, coarse.U, alpha)        # (22, ncy+2G, ncx+2G)

    def in_c(i, j):
        return 0 <= i < ncx and 0 <= j < ncy and cov_c[j, i]

    def cval(i, j):
        # proper nesting: every coarse cell a fill reads is a cell of the coarse level
        assert in_c(i, j), "fine level not properly nested in the coarse level at %r" % ((i, j),)
        return Ct[:, j + G, i + G]

    done = np.zeros((fine.ny, fine.nx), dtype=bool)
    slopes_cache = {}
    for b in fine.boxes:
        gb = grow(b, G)
        for jf in range(max(gb[1], 0), min(gb[3], fine.ny - 1) + 1):
            for i_f in range(max(gb[0], 0), min(gb[2], fine.nx - 1) + 1):
                if fine.valid[jf, i_f] or done[jf, i_f]:
                    continue
                ic, jc = i_f // r, jf // r
                if cov_f[jc, ic]:
                    continue
                done[jf, i_f] = True
                key = (ic, jc)
                if key not in slopes_cache:
                    c0 = cval(ic, jc)
                    s = []
                    for d in range(2):
                        e = (1, 0) if d == 0 else (0, 1)
                        hi_in = in_c(ic + e[0], jc + e[1])
                        lo_in = in_c(ic - e[0], jc - e[1])
                        if lo_in and hi_in:
                            sd = 0.5 * (cval(ic + e[0], jc + e[1]) - cval(ic - e[0], jc - e[1]))
                        elif lo_in:                        # hi neighbour missing: one-sided low
                            sd = c0 - cval(ic - e[0], jc - e[1])
                        elif hi_in:
                            sd = cval(ic + e[0], jc + e[1]) - c0
                        else:                              # both missing: the hi loop runs last
                            sd = cval(ic + e[0], jc + e[1]) - c0
                        s.append(sd)
                    # multidimensional limiter on coarse cells off the domain's edge layer
                    if 1 <= ic <= ncx - 2 and 1 <= jc <= ncy - 2:
                        smax = c0.copy()
                        smin = c0.copy()
                        for jj in (-1, 0, 1):
                            for ii in (-1, 0, 1):
                                v = cval(ic + ii, jc + jj)
                                smax = np.maximum(smax, v)
                                smin = np.minimum(smin, v)
                        dsum = 0.5 * (np.abs(s[0]) + np.abs(s[1]))
                        pos = dsum > 0.0
                        eta = np.minimum(np.minimum(smax - c0, c0 - smin) / np.where(pos, dsum, 1.0), 1.0)
                        s = [np.where(pos, eta * s[0], s[0]), np.where(pos, eta * s[1], s[1])]
                    slopes_cache[key] = (c0, s[0], s[1])
                c0, s0, s1 = slopes_cache[key]
                v = c0.copy()
                off0 = i_f - r * ic
                off1 = jf - r * jc
                v = v + (-0.5 + (off0 + 0.5) / r) * s0
                v = v + (-0.5 + (off1 + 0.5) / r) * s1
                fine.U[:, jf + G, i_f + G] = v


def fill_fault(fine, coarse, alpha):
    """Coarse-fine ghost fill of the fine fault data (D1PiecewiseLinearFillPatch::fillInterp)."""
    r = fine.r
    ncx = coarse.nx
    cov_f = np.zeros(ncx, dtype=bool)
    for b in fine.boxes:
        if b[1] == 0:
            cov_f[b[0] // r:b[2] // r + 1] = True
    Bt = _time_interp(coarse.Bold, coarse.B, alpha)

    def in_c(i):
        return 0 <= i < ncx and coarse.fvalid[i]

    for b in fine.boxes:
        if b[1] != 0:
            continue
        for i_f in range(max(b[0] - G, 0), min(b[2] + G, fine.nx - 1) + 1):
            if fine.fvalid[i_f]:
                continue
            ic = i_f // r
            if cov_f[ic]:
                continue
            c0 = Bt[:, ic + G]
            lo_in, hi_in = in_c(ic - 1), in_c(ic + 1)
            if lo_in and hi_in:
                s = 0.5 * (Bt[:, ic + 1 + G] - Bt[:, ic - 1 + G])
            elif lo_in:
                s = c0 - Bt[:, ic - 1 + G]
            else:
                s = Bt[:, ic + 1 + G] - c0
            # the y-direction increment multiplies a zero slope: it leaves every value as is
            fine.B[:, i_f + G] = c0 + (-0.5 + (i_f - r * ic + 0.5) / r) * s


def regrid_level(new, old, coarse):
    """Data of a new layout of a level (AMRLevelLinElast::regrid, AMRLevelLinElast.cpp:784-951):
    the new level's arrays already hold the initial fill; every new cell takes the old layout's
    value where the old layout covers it, else the coarse level's current data interpolated
    as Chombo's FineInterp does (piecewise constant, then per direction a central slope inside
    the domain, a one-sided slope on its edge layer -- high side at the lo edge, low side at
    the hi edge -- limited by the 3x3 coarse neighbourhood off the edge layer, added with
    coefficient -1/2 + (offset + 1/2)/r).  Fault data: the same in x with D1FineInterp
    (D1FineInterp.cpp:159-265; its limiter box keeps the faces off the x edges)."""
    r = new.r
    ncx, ncy = coarse.nx, coarse.ny
    C = coarse.U

    def cval(i, j):
        assert coarse.valid[j, i], "new level not properly nested at %r" % ((i, j),)
        return C[:, j + G, i + G]

    for b in new.boxes:
        for jf in range(b[1], b[3] + 1):
            for i_f in range(b[0], b[2] + 1):
                if old is not None and old.valid[jf, i_f]:
                    new.U[:, jf + G, i_f + G] = old.U[:, jf + G, i_f + G]
                    continue
                ic, jc = i_f // r, jf // r
                c0 = cval(ic, jc)
                s = []
                for d, (k, n) in enumerate(((ic, ncx), (jc, ncy))):
                    e = (1, 0) if d == 0 else (0, 1)
                    if k == 0:
                        s.append(cval(ic + e[0], jc + e[1]) - c0)
                    elif k == n - 1:
                        s.append(c0 - cval(ic - e[0], jc - e[1]))
                    else:
                        s.append(0.5 * (cval(ic + e[0], jc + e[1]) - cval(ic - e[0], jc - e[1])))
                if 1 <= ic <= ncx - 2 and 1 <= jc <= ncy - 2:
                    smax, smin = c0.copy(), c0.copy()
                    for jj in (-1, 0, 1):
                        for ii in (-1, 0, 1):
                            v = cval(ic + ii, jc + jj)
                            smax = np.maximum(smax, v)
                            smin = np.minimum(smin, v)
                    dsum = 0.5 * (np.abs(s[0]) + np.abs(s[1]))
                    pos = dsum > 0.0
                    eta = np.minimum(np.minimum(smax - c0, c0 - smin) / np.where(pos, dsum, 1.0), 1.0)
                    s = [np.where(pos, eta * s[0], s[0]), np.where(pos, eta * s[1], s[1])]
                v = c0 + (-0.5 + (i_f - r * ic + 0.5) / r) * s[0]
                v = v + (-0.5 + (jf - r * jc + 0.5) / r) * s[1]
                new.U[:, jf + G, i_f + G] = v
    CB = coarse.B

    def bval(i):
        assert coarse.fvalid[i], "new fault data not nested at %r" % (i,)
        return CB[:, i + G]

    for b in new.boxes:
        if b[1] != 0:
            continue
        for i_f in range(b[0], b[2] + 1):
            if old is not None and old.fvalid[i_f]:
                new.B[:, i_f + G] = old.B[:, i_f + G]
                continue
            ic = i_f // r
            c0 = bval(ic)
            if ic == 0:
                sl = bval(ic + 1) - c0
            elif ic == ncx - 1:
                sl = c0 - bval(ic - 1)
            else:
                sl = 0.5 * (bval(ic + 1) - bval(ic - 1))
                lo, hi = bval(ic - 1), bval(ic + 1)
                smax = np.maximum(np.maximum(lo, c0), hi)
                smin = np.minimum(np.minimum(lo, c0), hi)
                dsum = 0.5 * np.abs(sl)
                pos = dsum > 0.0
                eta = np.minimum(np.minimum(smax - c0, c0 - smin) / np.where(pos, dsum, 1.0), 1.0)
                sl = np.where(pos, eta * sl, sl)
            new.B[:, i_f + G] = c0 + (-0.5 + (i_f - r * ic + 0.5) / r) * sl


class Register:
    """Flux register between a coarse level and its finer level (restated LevelFluxRegister)."""

    def __init__(self, coarse, fine):
        r = fine.r
        self.r = r
        cov_f = np.zeros((coarse.ny, coarse.nx), dtype=bool)
        for b in fine.boxes:
            cb = coarsen_box(b, r)
            cov_f[cb[1]:cb[3] + 1, cb[0]:cb[2] + 1] = True
        self.cov_f = cov_f
        # entries: (d, side, i, j) of uncovered cell (i, j) with its side face on the fine region
        self.entries = []
        for d in range(2):
            e = (1, 0) if d == 0 else (0, 1)
            for side in (0, 1):                  # 0: lo face, 1: hi face
                sg = -1 if side == 0 else 1
                for j in range(coarse.ny):
                    for i in range(coarse.nx):
                        if not coarse.valid[j, i] or cov_f[j, i]:
                            continue
                        ni, nj = i + sg * e[0], j + sg * e[1]
                        if 0 <= ni < coarse.nx and 0 <= nj < coarse.ny and cov_f[nj, ni]:
                            self.entries.append((d, side, i, j))
        self.reg = np.zeros((len(self.entries), NC))

    def face_of(self, k):
        """(d, face index i, face index j) of entry k on the coarse level."""
        d, side, i, j = self.entries[k]
        if d == 0:
            return d, i + side, j
        return d, i, j + side

    def increment_coarse(self, Fc, dtc):
        """Fc[d]: dict box -> flux array (22, fny, fnx) over surroundingNodes(box, d) and its box."""
        self.reg[:] = 0.0
        for k in range(len(self.entries)):
            d, side, i, j = self.entries[k]
            _, fi, fj = self.face_of(k)
            F, b = Fc[(i, j)][d]
            self.reg[k] = self.reg[k] - dtc * F[:, fj - b[1], fi - b[0]]

    def increment_fine(self, fine_flux, dtf):
        w = dtf * (1.0 / self.r)
        r = self.r
        for k in range(len(self.entries)):
            d, fi, fj = self.face_of(k)
            for m in range(r):
                if d == 0:
                    ffi, ffj = fi * r, fj * r + m
                else:
                    ffi, ffj = fi * r + m, fj * r
                F, b = fine_flux(d, ffi, ffj)
                self.reg[k] = self.reg[k] + w * F[:, ffj - b[1], ffi - b[0]]

    def reflux(self, coarse):
        inv = -1.0 / coarse.dx
        for k in range(len(self.entries)):
            d, side, i, j = self.entries[k]
            s = inv * (1.0 if side == 1 else -1.0)
            coarse.U[:, j + G, i + G] = coarse.U[:, j + G, i + G] + s * self.reg[k]


def average_down(coarse, fine):
    r = fine.r
    scale = 1.0 / (r * r)
    for b in fine.boxes:
        cb = coarsen_box(b, r)
        for jc in range(cb[1], cb[3] + 1):
            for ic in range(cb[0], cb[2] + 1):
                acc = np.zeros(NC)
                for jj in range(r):
                    for ii in range(r):
                        acc = acc + fine.U[:, jc * r + jj + G, ic * r + ii + G]
                coarse.U[:, jc + G, ic + G] = acc * scale
    bscale = 1.0 / r
    for b in fine.boxes:
        if b[1] != 0:
            continue
        for ic in range(b[0] // r, b[2] // r + 1):
            acc = np.zeros(coarse.nb)
            for ii in range(r):
                acc = acc + fine.B[:, ic * r + ii + G]
            coarse.B[:, ic + G] = acc * bscale


class Hierarchy:
    """Fixed hierarchy: level 0 covers the domain; level l refines level l-1 by r."""

    def __init__(self, cfg0, boxes_per_level, r):
        self.r = r
        self.levels = []
        cfg = cfg0
        for l, boxes in enumerate(boxes_per_level):
            if l > 0:
                cfg = refine_cfg(cfg, r)
            self.levels.append(Level(cfg, boxes, r if l > 0 else None))
        self.regs = [Register(self.levels[l], self.levels[l + 1]) for l in range(len(self.levels) - 1)]
        self.err = 0

    def init_state(self, U_of_level, B_of_level):
        """Set each level's arrays (valid cells and frame) from functions of the level index."""
        for l, lev in enumerate(self.levels):
            lev.U[:] = U_of_level(l)
            lev.B[:] = B_of_level(l)

    def change_dt(self, dt0):
        """A new coarse dt between coarse steps: level l steps dt0 / r^l from now on and every
        level keeps its clock, the sum of its own steps (AMRLevel's m_time)."""
        dt = dt0
        for lev in self.levels:
            lev.dt = dt
            dt = dt / self.r

    def set_dt(self, dt0, t0=0.0):
        dt = dt0
        for lev in self.levels:
            lev.dt = dt
            lev.time = t0
            dt = dt / self.r

    def _step_level(self, l):
        """AMRLevelLinElast::advance + LinElastLevelGodunov::step for level l."""
        lev = self.levels[l]
        lev.Uold = lev.U.copy()
        lev.Bold = lev.B.copy()
        if l > 0:
            crs = self.levels[l - 1]
            t_new = crs.time
            t_old = t_new - crs.dt
            alpha = (lev.time - t_old) / (t_new - t_old)
            eps = 0.04 * lev.dt / self.r
            if abs(alpha) < eps:
                alpha = 0.0
            if abs(1.0 - alpha) < eps:
                alpha = 1.0
            assert 0.0 <= alpha <= 1.0
            fill_state(lev, crs, alpha)
            fill_fault(lev, crs, alpha)
        want_flux = l + 1 < len(self.levels) or l > 0
        Unew = lev.U.copy()
        Bnew = lev.B.copy()
        fluxes = {}
        for b in lev.boxes:
            gb = grow(b, G)
            fab = tmo.FabArray((gb[0], gb[1]), (gb[2], gb[3]), NC,
                               data=np.ascontiguousarray(lev.U[:, gb[1] + G:gb[3] + G + 1, gb[0] + G:gb[2] + G + 1]))
            bfab = None
            if b[1] == 0:
                bfab = tmo.FabArray((b[0] - G, 0), (b[2] + G, 0), lev.nb,
                                    data=np.ascontiguousarray(lev.B[:, None, b[0]:b[2] + 2 * G + 1]))
            F = [None, None]
            if want_flux:
                F[0] = tmo.FabArray((b[0], b[1]), (b[2] + 1, b[3]), NC)
                F[1] = tmo.FabArray((b[0], b[1]), (b[2], b[3] + 1), NC)
            e = tmo.lib().tmo_update_state(fab.fab, bfab.fab if bfab else None,
                                           F[0].fab if F[0] else None, F[1].fab if F[1] else None,
                                           lev.op.ref, tmo.Box.make(*b), lev.time, lev.dt)
            if e:
                self.err = self.err or e
            if lev.op.p.use_plasticity:
                tmo.lib().tmo_plastic_correction(fab.fab, lev.op.ref, lev.dt, tmo.Box.make(*b))
            Unew[:, b[1] + G:b[3] + G + 1, b[0] + G:b[2] + G + 1] = \
                fab.a[:, G:b[3] - b[1] + G + 1, G:b[2] - b[0] + G + 1]
            if bfab is not None:
                Bnew[:, b[0] + G:b[2] + G + 1] = bfab.a[:, 0, G:b[2] - b[0] + G + 1]
            if want_flux:
                fluxes[b] = ((F[0].a, (b[0], b[1])), (F[1].a, (b[0], b[1])))
        lev.U = Unew
        lev.B = Bnew
        if l + 1 < len(self.levels):
            box_of = {}
            for b in lev.boxes:
                for j in range(b[1], b[3] + 1):
                    for i in range(b[0], b[2] + 1):
                        box_of[(i, j)] = fluxes[b]
            self.regs[l].increment_coarse(box_of, lev.dt)
        if l > 0:
            def fine_flux(d, fi, fj, _f=fluxes, _boxes=lev.boxes):
                # the fine box whose face set holds face (fi, fj) on the refined-region side
                for b in _boxes:
                    if d == 0 and b[0] <= fi <= b[2] + 1 and b[1] <= fj <= b[3]:
                        return _f[b][0]
                    if d == 1 and b[0] <= fi <= b[2] and b[1] <= fj <= b[3] + 1:
                        return _f[b][1]
                raise KeyError((d, fi, fj))
            self.regs[l - 1].increment_fine(fine_flux, lev.dt)
        lev.time = lev.time + lev.dt

    def _post(self, l):
        if l + 1 < len(self.levels):
            self.regs[l].reflux(self.levels[l])
            average_down(self.levels[l], self.levels[l + 1])

    def _time_step(self, l):
        self._step_level(l)
        if l + 1 < len(self.levels):
            for _ in range(self.r):
                self._time_step(l + 1)
        self._post(l)

    def set_layout(self, l, boxes, init_U, init_B):
        """A new layout of level l >= 1: arrays from init_U/init_B (the initial fill), data by
        regrid_level from the old layout and level l-1; every level's clock restarts at the
        coarse level's time (as a new device hierarchy does)."""
        old = self.levels[l]
        new = Level(old.cfg, boxes, self.r)
        new.U[:] = init_U
        new.B[:] = init_B
        regrid_level(new, old, self.levels[l - 1])
        new.dt = old.dt
        self.levels[l] = new
        t = self.levels[0].time
        for lev in self.levels:
            lev.time = t
        self.regs = [Register(self.levels[k], self.levels[k + 1]) for k in range(len(self.levels) - 1)]

    def advance(self, nsteps=1):
        for _ in range(nsteps):
            self._time_step(0)
        return self.err
