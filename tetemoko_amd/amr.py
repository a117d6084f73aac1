"""Multi-level rupture runs on one device: a fixed hierarchy of levels stepped with
subcycling (tmk_amr_*).

The host mirror of Chombo's AMR driver over AMRLevelLinElast for a given box layout
per level (amrGodunov.cpp:443-461 set-up; AMRLevelLinElast.cpp:372-516 advance and
postTimeStep).  Level l has its own context (dx / r^l, per-face arrays injected from
level 0) and steps dt0 / r^l.  The caller gives the initial boxes of every level;
regrid() replaces the finer levels' layouts from the device tags of each level
(tmk_level_tag, tmk_make_boxes, tmk_level_regrid; AMR::regrid and
AMRLevelLinElast::regrid, AMRLevelLinElast.cpp:784-951).
"""
import numpy as np

from . import capi, scenarios

G = scenarios.GHOST


def make_levels(cfg0, boxes_per_level, r, precision=capi.PREC_FP64, tile=None):
    """Contexts and levels of a hierarchy (initial state: background and initial fault data),
    not yet stepped together: for a tmk_amr (Hierarchy) or a hierarchy batch (capi.AmrBatch)."""
    cfg = cfg0
    ctxs, levels = [], []
    for l, boxes in enumerate(boxes_per_level):
        if l > 0:
            cfg = scenarios.refine_cfg(cfg, r)
        ctx = capi.Context(cfg, precision)
        if tile:
            ctx.set_tile(*tile)
        lev = capi.Level(ctx, boxes)
        Bf = scenarios.fault_arrays(cfg)
        for k, b in enumerate(boxes):
            lev.fill(k, scenarios.background_vector(cfg), scenarios.box_bview(Bf, b) if b[1] == 0 else None)
        ctxs.append(ctx)
        levels.append(lev)
    return ctxs, levels


def dt_schedule(nsteps, cfg0, cfl=None, initial_cfl=None, max_dt_growth=1.1, fixed_dt=-1.0):
    """Coarse dt of each of nsteps coarse steps as the AMR driver assigns it
    (amrGodunov.cpp:443-461: cfl, initial_cfl, max_dt_growth, fixed_dt).

    Every level's dt is a CFL number times dx_l / max(cs, cp) (AMRLevelLinElast.cpp:479-480,
    :1892-1901), and dx_l r^l = dx_0, so the coarse dt the AMR driver takes as the minimum over
    levels is the same number times dx_0 / max(cs, cp): the first step uses initial_cfl, every
    later one min(cfl dx_0 / c, max_dt_growth times the previous dt); a positive fixed_dt
    replaces all of that.  The wave speed is constant, so the schedule ramps geometrically
    from the initial dt to the CFL dt and stays there.  Defaults of cfl and initial_cfl: the
    configuration's cfl."""
    if fixed_dt is not None and fixed_dt > 0:
        return [float(fixed_dt)] * nsteps
    c = max(cfg0["cs"], cfg0["cp"])
    cfl = cfg0.get("cfl", 0.9) if cfl is None else cfl
    initial_cfl = cfl if initial_cfl is None else initial_cfl
    dt, out = initial_cfl * cfg0["dx"] / c, []
    for _ in range(nsteps):
        out.append(dt)
        dt = min(cfl * cfg0["dx"] / c, max_dt_growth * dt)
    return out


class Hierarchy:
    def __init__(self, cfg0, boxes_per_level, r, precision=capi.PREC_FP64, tile=None, arrays=None):
        """arrays(l) -> (U, B) level arrays over the whole level domain with ghosts (the
        initial state); default: the background state and the initial fault data."""
        self.r = r
        self.cfgs = [cfg0]
        for _ in range(1, len(boxes_per_level)):
            self.cfgs.append(scenarios.refine_cfg(self.cfgs[-1], r))
        self.boxes = [[tuple(b) for b in bl] for bl in boxes_per_level]
        self.ctxs, self.levels = [], []
        for l, cfg in enumerate(self.cfgs):
            ctx = capi.Context(cfg, precision)
            if tile:
                ctx.set_tile(*tile)
            self.ctxs.append(ctx)
            self.levels.append(self._new_level(l, self.boxes[l], arrays))
        self.amr = capi.Amr(self.levels, r)
        self.dt0 = scenarios.dt_of(cfg0)
        self.time = 0.0
        self.steps = 0

    def _new_level(self, l, boxes, arrays=None):
        cfg = self.cfgs[l]
        lev = capi.Level(self.ctxs[l], boxes)
        if arrays is None:
            Bf = scenarios.fault_arrays(cfg)
            for k, b in enumerate(boxes):
                lev.fill(k, scenarios.background_vector(cfg), scenarios.box_bview(Bf, b) if b[1] == 0 else None)
        else:
            U, Bf = arrays(l)
            for k, b in enumerate(boxes):
                lev.upload(k, scenarios.box_view(U, b), scenarios.box_bview(Bf, b) if b[1] == 0 else None)
        return lev

    def set_layout(self, l, boxes):
        """Give level l >= 1 a new box layout (boxes in its cells): its data come from the old
        layout where it covers them and from level l-1 elsewhere (AMRLevelLinElast::regrid)."""
        self.amr.close()
        new = self._new_level(l, [tuple(b) for b in boxes])
        capi.level_regrid(new, self.levels[l], self.levels[l - 1], self.r)
        self.levels[l].close()
        self.levels[l] = new
        self.boxes[l] = [tuple(b) for b in boxes]
        self.amr = capi.Amr(self.levels, self.r)

    def regrid(self, tag_params, fill_ratio=0.75, block_factor=8, max_grid_size=64):
        """AMR::regrid from level 0: for l = 0, 1, ... the tags of level l (tag_params(l) ->
        capi.TagParams) give the boxes of level l+1 (Berger-Rigoutsos, properly nested in
        level l), whose data come from its old layout where it had one and from level l.  As in
        Chombo's AMR::regrid, a level without tags ends the hierarchy there: its finer levels are
        removed, and a later regrid whose tags reach down again re-creates them (up to the
        number of levels the hierarchy was made with).  Returns the box counts per level."""
        self.amr.close()
        l = 0
        while l < len(self.cfgs) - 1 and l < len(self.levels):
            cfg = self.cfgs[l]
            tags = self.levels[l].tag(tag_params(l), cfg["nx"], cfg["ny"])
            boxes = []
            if tags.any():
                boxes = capi.make_boxes(tags, self.boxes[l] if l > 0 else None, ref_ratio=self.r,
                                        block_factor=block_factor, max_grid_size=max_grid_size, fill_ratio=fill_ratio)
            if not boxes:
                for lev in self.levels[l + 1:]:
                    lev.close()
                del self.levels[l + 1:]
                del self.boxes[l + 1:]
                break
            new = self._new_level(l + 1, boxes)
            old = self.levels[l + 1] if l + 1 < len(self.levels) else None
            capi.level_regrid(new, old, self.levels[l], self.r)
            if old is not None:
                old.close()
                self.levels[l + 1] = new
                self.boxes[l + 1] = boxes
            else:
                self.levels.append(new)
                self.boxes.append(boxes)
            l += 1
        self.amr = capi.Amr(self.levels, self.r)
        return [len(b) for b in self.boxes]

    def run(self, nsteps):
        self.amr.run(nsteps, self.time, self.dt0)
        for _ in range(nsteps):
            self.time = self.time + self.dt0
        self.steps += nsteps

    def run_schedule(self, dts):
        """One coarse step per entry of dts (coarse dt of each step, e.g. dt_schedule(...));
        runs of equal dt go to the device as one call (two coarse steps per graph)."""
        k = 0
        while k < len(dts):
            n = 1
            while k + n < len(dts) and dts[k + n] == dts[k]:
                n += 1
            self.dt0 = float(dts[k])
            self.run(n)
            k += n

    def run_timed(self, nsteps):
        ms = self.amr.run_timed(nsteps, self.time, self.dt0)
        for _ in range(nsteps):
            self.time = self.time + self.dt0
        self.steps += nsteps
        return ms

    def sync(self):
        self.amr.sync()

    def cell_updates_per_coarse_step(self):
        """sum over levels of valid cells x steps of that level per coarse step"""
        tot = 0
        for l, bl in enumerate(self.boxes):
            cells = sum((b[2] - b[0] + 1) * (b[3] - b[1] + 1) for b in bl)
            tot += cells * self.r ** l
        return tot

    def download(self, l):
        """(U, B) level arrays of level l: the boxes' valid cells, NaN elsewhere."""
        cfg = self.cfgs[l]
        nx, ny = cfg["nx"], cfg["ny"]
        nb = scenarios.NB[cfg["fault"]]
        U = np.full((scenarios.NCOMP, ny + 2 * G, nx + 2 * G), np.nan)
        B = np.full((nb, nx + 2 * G), np.nan)
        lev = self.levels[l]
        for k, b in enumerate(self.boxes[l]):
            x0, y0, x1, y1 = b
            ub = np.empty((scenarios.NCOMP, y1 - y0 + 1 + 2 * G, x1 - x0 + 1 + 2 * G))
            bb = np.empty((nb, x1 - x0 + 1 + 2 * G)) if y0 == 0 else None
            lev.download(k, ub, bb)
            U[:, y0 + G:y1 + G + 1, x0 + G:x1 + G + 1] = ub[:, G:-G, G:-G]
            if bb is not None:
                B[:, x0 + G:x1 + G + 1] = bb[:, G:-G]
        return U, B

    def close(self):
        self.amr.close()
        for lev in self.levels:
            lev.close()
        for ctx in self.ctxs:
            ctx.close()
