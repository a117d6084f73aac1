/*
 * tetemoko_hip.h -- C ABI of the MI355X (gfx950) per-patch Godunov update.
 *
 * The library replaces the Fortran kernels behind tetemoko's Chombo plugin
 * surface.  Every entry point takes plain pointers, sizes and index boxes;
 * no Chombo, HIP or PyTorch types appear in the signatures.  The physics
 * parameters, which the reference keeps in Fortran COMMON blocks
 * (/IBC/, /PLASTIC/, /rs/, /slipweak/), are carried per context instead, so
 * several scenarios can live in one process.
 *
 * Storage convention (Chombo FArrayBox): an array over an inclusive index box
 * is stored in Fortran order, x fastest, then y, then component.
 * State: 22 components (srcLinElast/LEINDEX.H:12-47).  Rate-and-state fault
 * data: 11 components (ibcLinElast/RSINDEX.H:12-22).  Slip-weakening fault
 * data: 9 components (ibcLinElast/SWIBC.H:129-142).
 */
#ifndef TETEMOKO_HIP_H
#define TETEMOKO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TMK_NCOMP 22
#define TMK_NBDRY_RS 11
#define TMK_NBDRY_SW 9
#define TMK_MAXPATCH 8
#define TMK_GHOST 4

/* error codes; the reference aborts (MayDay::Error / Fortran stop) on the first four */
#define TMK_OK 0
#define TMK_E_PSI_NAN (-1)     /* RSIBCF.ChF:133-145 "dPsi is nan" */
#define TMK_E_BAD_BC (-2)      /* RSIBC.cpp:269 "Invalid Boundary Type" */
#define TMK_E_PRED_ORDER (-3)  /* LinElastPatchGodunov.cpp:313, :543, :645 */
#define TMK_E_NO_BDRY (-4)     /* fault face reached without fault data (RSIBC::setBdryData not called) */
#define TMK_E_BOX (-5)         /* an array box does not cover the cells the update reads or writes */
#define TMK_E_ARG (-6)         /* invalid argument */
#define TMK_E_HIP (-7)         /* HIP runtime failure */
#define TMK_E_UNSUPPORTED (-8) /* option the reference aborts on (flattening, artificial viscosity, high-order limiter) */
#define TMK_E_NOMEM (-9)       /* device memory exhausted */
#define TMK_E_OVERFLOW (-10)   /* station records dropped: buffer full before it was read */
#define TMK_E_AMR_TIME (-11)   /* LinElastLevelGodunov.cpp:250-258 "alpha < 0.0" / "alpha > 1.0" */
#define TMK_E_NESTING (-12)    /* a finer level is not aligned to the ratio or not properly nested */
#define TMK_E_IO (-13)         /* file output failed (boundary plot files, libtmk_h5.so) */
#define TMK_E_SCRATCH (-14)    /* tmk_patch_update_state_noalloc: the call needs more than its fixed scratch */

#define TMK_FAULT_RS 0
#define TMK_FAULT_SW 1

#define TMK_PRED_CTU 0
#define TMK_PRED_PLM 1
#define TMK_PRED_PPM 2

/* arithmetic of the bulk update: FP64 everywhere, or FP32 state and bulk
 * arithmetic with the fault solve, state evolution, slip and time in FP64 */
#define TMK_PREC_FP64 0
#define TMK_PREC_FP32_BULK 1

typedef struct { int lo[2]; int hi[2]; } tmk_box;

/* Replaces the COMMON blocks and the IBC constructor arguments
 * (IBCCommon.fh:4-6, PlasticCommon.fh:4-5, RSCommon.fh:4-5, SWCommon.fh:4-5,
 * RSIBC.cpp:17-49, SWIBC.cpp:15-50, amrGodunov.cpp:194-759). */
typedef struct {
    /* material: LINELASTSETF inputs (LinElastPhysicsF.ChF:1149-1158) */
    double cs, cp, G;
    double sxx0, syy0, szz0, sxy0, sxz0, syz0;
    /* plasticity: PLASTICSETF (PlasticPhysicsF.ChF:6-25); p_b (cohesion) and p_h (hardening) are 0 in the reference */
    int use_plasticity;
    double p_mu, p_beta, p_eta, p_b, p_h;
    /* boundary type per [dir*2 + side] (amrGodunov.cpp:600-611): 0 outflow, 1 free surface; lo-y is the fault */
    int bc[4];
    int fault_kind;
    /* rate-and-state (RSSETF, RSIBCF.ChF:8-55) */
    double r0, nuc_time, nsig, rs_a, rs_b, rs_V0, rs_f0, rs_L, rs_fw, rs_Vw, rs_fexp;
    double x0, y0, rs_rupture_thr, ramp_x, ramp_w, ramp_a, ramp_Vw;
    double rs_psi0;
    int state_law; /* 0: RSSETBND psi update (RSIBCF.ChF:126-132); 1: aging law (extension) */
    /* slip weakening (SWSETF, SWIBCF.ChF:6-27, and SWIBC members) */
    double fs, fd, dw, sm_value, sw_rupture_thr;
    int num_patches;
    double xc[TMK_MAXPATCH], xw[TMK_MAXPATCH], zc[TMK_MAXPATCH], zw[TMK_MAXPATCH], tau_p[TMK_MAXPATCH];
    double fric_center[2], fric_width[2], outside_f;
    /* per-face extensions, host arrays indexed by (face i - face_lo), NULL = reference scalar;
     * copied to the device by tmk_ctx_create */
    int face_lo, nface;
    const double *f_a, *f_b, *f_L, *f_fs, *f_fd, *f_dc, *f_syy, *f_tau;
    /* numerics (amrGodunov.cpp:298-355) */
    int normal_pred, fourth_order, prim_limiting, char_limiting;
    double dx;
    tmk_box domain;
} tmk_params;

typedef struct tmk_ctx tmk_ctx;
typedef struct tmk_level tmk_level;
typedef struct tmk_batch tmk_batch;

/* Context = one scenario's parameters + stream + device scratch.
 * Replaces LEPhysIBC::setFortranCommonLE/Plastic (LEPhysIBC.cpp:13-36) and the
 * RSIBC/SWIBC constructors (RSIBC.cpp:17-49, SWIBC.cpp:15-50). */
int tmk_ctx_create(tmk_ctx **ctx, const tmk_params *p, int precision);
int tmk_ctx_destroy(tmk_ctx *ctx);
const char *tmk_strerror(int code);
/* 1 when the library was built for gfx950 and a device is visible */
int tmk_device_ok(void);
/* free and total device memory (admission control of the serving loop); free_bytes includes
 * the blocks the level-array cache holds on the current device */
int tmk_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Level-array cache (no counterpart in the reference, whose levels live for a whole run; for
 * serving loops that create and destroy levels continuously).  Island, fault-row and owner
 * arrays of levels, hierarchy members' flux arrays and face masks, and contexts' per-face
 * parameter arrays are taken from a per-device cache of blocks in size classes (at most 1/8
 * above the request) and go back to it when their owner is destroyed.  tmk_cache_info reports
 * the current device's idle cached bytes, the bytes in use by live objects, and the limit on
 * cached bytes per device (default 32 GiB); tmk_cache_set_limit sets that limit and frees idle
 * blocks above it on every device (0 turns caching off); tmk_cache_trim frees every idle block
 * of the current device. */
int tmk_cache_info(size_t *cached_bytes, size_t *live_bytes, size_t *limit_bytes);
int tmk_cache_set_limit(size_t bytes);
int tmk_cache_trim(void);

/* LinElastPatchGodunov::updateState (LinElastPatchGodunov.cpp:101-129), with
 * LinElastPhysics::computeUpdate (LinElastPhysics.cpp:135-179) and the IBC
 * calls it makes (RSIBC::primBC, RSIBC::updateBoundary).  Host FArrayBox data:
 *   U   : state on U_box (valid grown by the ghost cells), 22 comps, updated in place
 *   F0/F1: optional fluxes on surroundingNodes(valid, d), 22 comps (NULL to skip)
 *   B   : fault data on B_box (j = 0 faces, x-ghosts), NULL when valid does not touch lo-y
 *   max_wave_speed: MAXWAVESPEEDF (LinElastPhysicsF.ChF:10-21) */
int tmk_patch_update_state(tmk_ctx *ctx, tmk_box valid, double *U, tmk_box U_box, int ncomp,
                           double *F0, double *F1, double *B, tmk_box B_box, int nbdry,
                           double time, double dt, double *max_wave_speed);

/* tmk_patch_update_state with no device allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, checks, results and error codes as
 * tmk_patch_update_state (the same kernels on the same tiles), but every device buffer of the
 * call -- both state copies, the fault data, the fluxes, the FP64 staging of the FP32 bulk and the
 * launch tables -- lives in one fixed array of TMK_NOALLOC_SCRATCH_BYTES per device, held from
 * the first context created on that device until the process ends.  TMK_E_SCRATCH when the call
 * needs more than that (for FP64 with both fluxes, boxes up to about 128 x 128 cells fit) or
 * more than 1024 tiles.  Every call on a device shares that array: calls that overlap on one
 * device (from several threads, or contexts used concurrently) race and must be serialised by
 * the caller. */
#define TMK_NOALLOC_SCRATCH_BYTES (16u << 20)
int tmk_patch_update_state_noalloc(tmk_ctx *ctx, tmk_box valid, double *U, tmk_box U_box, int ncomp,
                                   double *F0, double *F1, double *B, tmk_box B_box, int nbdry,
                                   double time, double dt, double *max_wave_speed);

/* LinElastPatchGodunov::plasticUpdate (LinElastPatchGodunov.cpp:161-170) ->
 * PLASTICCORRECTION (PlasticPhysicsF.ChF:30-69) on the valid box. */
int tmk_patch_plastic_update(tmk_ctx *ctx, tmk_box valid, double *U, tmk_box U_box, int ncomp, double dt);

/* Device-resident level: LinElastLevelGodunov (LinElastLevelGodunov.cpp:45-398).
 * boxes: disjoint valid boxes of one level.  Fault data is allocated for every
 * box whose lo-y edge lies on the domain lo-y boundary.  TMK_E_NOMEM also when one state
 * plane of an island (boxes whose ghost regions touch share one) would exceed 4 GiB: the
 * kernels address planes with 32-bit offsets (536 M FP64 cells, more than the device holds
 * at three copies of 22 planes). */
int tmk_level_create(tmk_ctx *ctx, int nboxes, const tmk_box *boxes, tmk_level **out);
/* device bytes tmk_level_create would allocate for these boxes */
int tmk_level_footprint(const tmk_ctx *ctx, int nboxes, const tmk_box *boxes, size_t *bytes);
/* the same before any context exists, from the parameters a context would be created with
 * (admission control: no reference counterpart; the reference sizes one run per process) */
int tmk_level_footprint_params(const tmk_params *params, int precision, int nboxes, const tmk_box *boxes,
                               size_t *bytes);
int tmk_level_destroy(tmk_level *lev);
/* copy one box's ghosted state (grow(box,4), 22 comps) and fault data
 * (x in [lo-4, hi+4], j = 0, nbdry comps; NULL if the box has none) to/from the device */
int tmk_level_upload(tmk_level *lev, int ibox, const double *U, const double *B);
int tmk_level_download(tmk_level *lev, int ibox, double *U, double *B);
/* fill one box's ghosted state with a uniform 22-component state (LEINITF,
 * LinElastPhysicsF.ChF:1215-1263) on the device and upload its fault data */
int tmk_level_fill(tmk_level *lev, int ibox, const double *state22, const double *B);
/* copy one box's fault data (x in [lo-4, hi+4], nbdry comps) to the host */
int tmk_level_download_fault(tmk_level *lev, int ibox, double *B);
/* LinElastLevelGodunov::step (LinElastLevelGodunov.cpp:174-398) followed by
 * LinElastLevelGodunov::plasticUpdate (:643-667) when plasticity is on:
 * same-level ghost exchange, one launch over every box of the level, fault data
 * update.  dt_new = dx / max(cs,cp) (LinElastLevelGodunov.cpp:377). */
int tmk_level_step(tmk_level *lev, double time, double dt, double *dt_new);
/* nsteps level steps with time advancing by dt, captured once as a hipGraph */
int tmk_level_run(tmk_level *lev, int nsteps, double time, double dt);
/* tmk_level_run bracketed by HIP events recorded on the level's own stream;
 * waits for completion and returns the elapsed milliseconds in *ms */
int tmk_level_run_timed(tmk_level *lev, int nsteps, double time, double dt, double *ms);
/* nsteps eager level steps (host-supplied times, no graph) bracketed by HIP
 * events on the level's stream; *ms_per_step = mean elapsed time per step.  A step
 * of a single-box level is the fused update (one launch over the interior tiles and,
 * concurrently on a side stream, one over the tiles that meet the domain boundary
 * or the fault) plus the clock advance. */
int tmk_level_profile(tmk_level *lev, int nsteps, double time, double dt, double *ms_per_step);

/* Per-kernel timing of level steps.  Measurement aid with no counterpart in the
 * reference.  Runs nsteps steps from (time, dt), captured as one graph and replayed once
 * (eagerly where capture or event timing inside a graph is unavailable; in_graph says
 * which).  Each launch is bracketed by HIP events on the stream it runs on: the
 * interior-tile update on the level's stream, the boundary-tile update on the side
 * stream.  Reports means in ms. */
typedef struct {
    double ms_step;      /* whole step (both update launches, exchanges, clock) */
    double ms_interior;  /* interior-tile update launch */
    double ms_boundary;  /* boundary-tile update launch (0 when there is none) */
    long long cells_interior, cells_boundary; /* valid cells each launch updates */
    int tiles_interior, tiles_boundary;
    int in_graph;        /* 1: timed inside a replayed graph; 0: eager steps */
} tmk_profile;
int tmk_level_profile_kernels(tmk_level *lev, int nsteps, double time, double dt, tmk_profile *out);
/* synchronize and return the first error any launch reported */
int tmk_level_sync(tmk_level *lev);

/* Refinement tagging of a level's current state: AMRLevelLinElast::tagCells
 * (AMRLevelLinElast.cpp:577-717) with the IBC nucleation override
 * (RSIBC.cpp:341-377, SWIBC.cpp:336-396), tag buffer growth and domain clip.
 * stress_thresh / sliprate_thresh > 0 add the stress-gradient (sxx, syy, sxy)
 * and fault slip-rate criteria (extensions); <= 0 disables them.
 * tags: (domain nx) * (domain ny) bytes, x fastest, 1 = tagged. */
typedef struct {
    double refine_thresh, plastic_thresh;
    int tag_buffer;
    double x_coarsen[2], y_coarsen[2], slope_coarsen, width_coarsen;
    double x_coarsen2[2], y_coarsen2[2];
    int level, coarsen2_max_level;
    double stress_thresh, sliprate_thresh;
} tmk_tagparams;
int tmk_level_tag(tmk_level *lev, const tmk_tagparams *tp, unsigned char *tags);

/* Batches: the boxes of every member level (one scenario each, own parameters,
 * clock and step) are updated in one shared fused launch per step.  Members
 * join and leave between steps (continuous batching); a member's results are
 * bit-identical to stepping it alone.  Members must share precision, tile
 * shape, fault law and anti-plane mode.  No reference counterpart: the
 * reference runs one scenario per process (amrGodunov.cpp:81-166). */
int tmk_batch_create(tmk_batch **out);
int tmk_batch_destroy(tmk_batch *b);
/* join with the member's clock (time of its next step) and step dt */
int tmk_batch_add(tmk_batch *b, tmk_level *lev, double time, double dt);
/* leave; the level can then be downloaded or stepped alone */
int tmk_batch_remove(tmk_batch *b, tmk_level *lev);
/* advance every member by nsteps of its own dt */
int tmk_batch_run(tmk_batch *b, int nsteps);
/* synchronize; first error reported by any member (0 = none) */
int tmk_batch_sync(tmk_batch *b);
/* without waiting: *done = 1 when every step enqueued on the batch has completed, else 0.
 * No counterpart in the reference: lets a serving loop tend whichever batch has finished
 * its chunk while the others run. */
int tmk_batch_done(tmk_batch *b, int *done);
/* synchronize; *code = the first error member lev reported (0 = none).  No
 * counterpart in the reference, which runs one scenario per process: a serving loop
 * uses it to retire only the scenario that failed (a member's error does not touch
 * the others' data). */
int tmk_batch_member_error(tmk_batch *b, tmk_level *lev, int *code);

/* Fault stations of a level (AMRLevelLinElast fault stations: placement
 * AMRLevelLinElast.cpp:258-278, record AMRLevelLinElast.cpp:1758-1810).  Station k sits at
 * x[k] metres from the domain's lo x edge (the reference's x_fault_stations plus its domain
 * centre): face i = floor(x/dx - 0.5), weight d = x - (i + 0.5) dx.  After every level step of
 * tmk_level_step/run the device appends one record per station -- the new time, then
 * B[i] (1 - d) + B[i+1] d for each fault-data component, the reference's expression with d in
 * metres -- inside captured step graphs as well.  capacity = records kept between reads.
 * Stations of a level that is in a batch record nothing.  A station whose face lies on no
 * fault box is TMK_E_ARG. */
typedef struct tmk_stations tmk_stations;
int tmk_stations_create(tmk_level *lev, int nstations, const double *x, int capacity, tmk_stations **out);
/* waits for the level, copies the records written since the last read into
 * records[nrecords][nstations][1 + numBdryVars] (max_records >= capacity) and clears them;
 * TMK_E_OVERFLOW when records were dropped (the kept ones are still copied) */
int tmk_stations_read(tmk_stations *st, double *records, int max_records, int *nrecords);
int tmk_stations_destroy(tmk_stations *st);

/* Hierarchies of levels: the subcycled multi-level step.  Replaces the level loop
 * of Chombo's AMR::timeStep over AMRLevelLinElast::advance and ::postTimeStep
 * (AMRLevelLinElast.cpp:372-516) and the coarse-fine parts of
 * LinElastLevelGodunov::step (LinElastLevelGodunov.cpp:226-264 ghost fill of U and of
 * the fault data from the coarser level at the interpolated time, :334-355 flux
 * registers), on one device stream, with two coarse steps captured as one graph.
 * levels[0] covers its domain (lo corner 0, 0); levels[l] has its own context with
 * the domain refined by ref_ratio and dx / ref_ratio, boxes aligned to ref_ratio and
 * nested in levels[l-1] so that every coarse cell the fill reads (the 3x3
 * neighbourhood of the coarse parent of each fine ghost cell inside the domain) is a
 * cell of levels[l-1]; otherwise TMK_E_NESTING.  While a level belongs to a hierarchy
 * it cannot be stepped, uploaded or batched on its own; downloads wait for the
 * hierarchy.  One coarse step: level 0 steps dt0, then ref_ratio steps of level 1,
 * each followed by its own finer steps, then reflux (-1/dx) and average-down of the
 * state and the fault data into the coarser level. */
typedef struct tmk_amr tmk_amr;
int tmk_amr_create(int nlevels, tmk_level *const *levels, int ref_ratio, tmk_amr **out);
int tmk_amr_destroy(tmk_amr *amr);
/* one-scenario hierarchy: nsteps coarse steps of dt0 (level l steps dt0 / ref_ratio^l).  A time
 * other than the hierarchy's current coarse time (the first call's, or the previous calls'
 * start plus their steps) sets every level's clock to it; a new dt0 at the current time keeps
 * the clocks running, each the sum of its level's steps (AMRLevel's m_time). */
int tmk_amr_run(tmk_amr *amr, int nsteps, double time, double dt0);
/* the same, timed with HIP events on the hierarchy's stream (waits) */
int tmk_amr_run_timed(tmk_amr *amr, int nsteps, double time, double dt0, double *ms);
/* waits; the first error any member reported (0 = none); a member's error stays with it
 * until it leaves (tmk_amr_member_error names the member) */
int tmk_amr_sync(tmk_amr *amr);
/* without waiting: *done = 1 when every step enqueued on the hierarchy (or hierarchy batch)
 * has completed, else 0 (no counterpart in the reference, as tmk_batch_done) */
int tmk_amr_done(tmk_amr *amr, int *done);

/* Hierarchy batches: scenarios with the same number of levels and ratio whose level-l boxes
 * share the level-l update launches (and whose fills, registers and averages share theirs),
 * each with its own parameters, clock and dt.  Members join and leave between steps.  No
 * reference counterpart: the reference runs one scenario per process
 * (amrGodunov.cpp:81-166). */
int tmk_amr_batch_create(int nlevels, int ref_ratio, tmk_amr **out);
/* join with levels[0 .. nlevels-1] (coarsest first), the clock of the next coarse step and dt0 */
int tmk_amr_add(tmk_amr *amr, tmk_level *const *levels, double time, double dt0);
/* leave; the member's levels are plain levels again */
int tmk_amr_remove(tmk_amr *amr, tmk_level *level0);
/* nsteps coarse steps of every member from its own clock */
int tmk_amr_advance(tmk_amr *amr, int nsteps);
/* waits; *code = the first error the member reported (0 = none); the others are untouched */
int tmk_amr_member_error(tmk_amr *amr, tmk_level *level0, int *code);

/* Regrid, step 1: boxes of the next finer level from the tags of a level (Chombo's
 * MeshRefine::regrid with BRMeshRefine, as AMR::regrid drives it; deck parameters
 * fill_ratio, block_factor, max_grid_size, ref_ratio, amrGodunov.cpp:263-320).  Tags are
 * nx x ny bytes over the level's domain (tmk_level_tag).  Blocks of block_factor/ref_ratio
 * cells of the level are clustered with the Berger-Rigoutsos method, keeping every box
 * inside the level's proper nesting domain (its boxes shrunk by nest_buffer cells, the
 * domain edge excepted; ncoarse = 0: the whole domain); the boxes are returned in the
 * finer level's cells, multiples of block_factor, at most max_grid_size per side.
 * TMK_E_OVERFLOW with *nout = the count when more than max_out boxes result.  Host only.
 * Chombo's BRMeshRefine is not available here: layouts are this library's own. */
typedef struct {
    double fill_ratio;   /* accepted fraction of tagged cells in a box, (0, 1] */
    int block_factor;    /* finer-level boxes are multiples of this, a multiple of ref_ratio */
    int max_grid_size;   /* finer-level boxes have at most this many cells per side */
    int ref_ratio;
    int nest_buffer;     /* cells of this level around a finer box that must belong to it;
                            <= 0: the coarse-fine fill's reach, ceil(4 / ref_ratio) + 1 */
} tmk_meshparams;
int tmk_make_boxes(int nx, int ny, const unsigned char *tags, int ncoarse, const tmk_box *coarse_boxes,
                   const tmk_meshparams *mp, int max_out, tmk_box *out, int *nout);

/* Regrid, step 2: the data of a new layout of a level (AMRLevelLinElast::regrid,
 * AMRLevelLinElast.cpp:784-951).  The new level's cells take the old level's current data
 * where the old layout covers them and otherwise are interpolated from the coarser level's
 * current data: state as Chombo's FineInterp (piecewise linear, central slopes inside the
 * domain and one-sided at its edges, multidimensionally limited off the domain's edge
 * layer), fault data as D1FineInterp (D1FineInterp.cpp:61-267).  old may be NULL.  The new
 * level's other cells keep what tmk_level_fill/upload put there. */
int tmk_level_regrid(tmk_level *fresh, tmk_level *old, tmk_level *coarse, int ref_ratio);

/* Workgroup tile shape of the update kernels (autotuner hook): bx x by threads
 * per side-pair, (bx - 2) x (by - 2) valid cells.  Shapes: 16x16, 32x8, 32x12, 32x16,
 * 32x32, 64x4, 64x8, 64x16.  0, 0 = the library default for the context's precision. */
int tmk_set_tile(tmk_ctx *ctx, int bx, int by);

/* Row-marching interior update (autotuner hook; no reference counterpart: the reference updates
 * box by box, LinElastLevelGodunov.cpp:284-358).  rows > 0: levels created from this context
 * update the cells whose +-3 stencil lies inside the domain in strips of 58 columns x rows,
 * one wavefront per strip and side marching up the rows with no LDS (mode II, PPM with primitive
 * limiting; other numerics keep the tile kernels); the cells along the domain boundary run the
 * general kernel.  0 = the tile kernels of tmk_set_tile.  Results are those of the tile kernels
 * (bit-identical in FP64).  New contexts start at 32 rows. */
int tmk_set_march(tmk_ctx *ctx, int rows);

/* A level's own update shape (no reference counterpart; levels start with their context's):
 * march_rows as tmk_set_march, (bx, by) as tmk_set_tile.  Not while the level is in a batch or
 * hierarchy; members of one batch or hierarchy level need equal shapes.  get reports the
 * shape the update runs (march_rows 0 where the numerics keep the tile kernels). */
int tmk_level_set_update_shape(tmk_level *lev, int march_rows, int bx, int by);
int tmk_level_get_update_shape(const tmk_level *lev, int *march_rows, int *bx, int *by);
/* The order of a level's interior workgroups (no reference counterpart, whose box loop is
 * serial, LinElastLevelGodunov.cpp:284-358): xcd = 1 (the default) gives each of the device's
 * 8 XCDs a contiguous band of the interior tiles or strips, so that the halos neighbouring
 * workgroups share stay in that XCD's L2; 0 keeps row-major order.  Results do not depend on
 * it.  A batch or hierarchy level takes its first member's order.  Not while the level is in a
 * batch or hierarchy. */
int tmk_level_set_tile_order(tmk_level *lev, int xcd);

/* Autotuning of a level's update shape (no reference counterpart).  Times nsteps
 * graph-replayed steps from (time, dt), after two untimed ones, for every strip height of the
 * row-marching update (4 .. 64 rows, where the numerics allow it) and every tile shape, each
 * from the level's current state, then the fastest of them with the other order of the
 * interior workgroups (tmk_level_set_tile_order); sets the fastest shape and order and returns
 * the level to the state it started from (a device copy of the state is held meanwhile).  Not
 * while the level is in a batch or hierarchy or has fault stations. */
#define TMK_TUNE_MAX 32
typedef struct {
    int march_rows, bx, by;  /* the shape chosen and set */
    double ms_per_step;      /* its mean step time */
    int ntried;
    int march_rows_tried[TMK_TUNE_MAX], bx_tried[TMK_TUNE_MAX], by_tried[TMK_TUNE_MAX];
    double ms_tried[TMK_TUNE_MAX];
    int xcd;                 /* the order chosen and set (tmk_level_set_tile_order) */
    int xcd_tried[TMK_TUNE_MAX];
} tmk_tuning;
int tmk_level_autotune(tmk_level *lev, double time, double dt, int nsteps, tmk_tuning *out);

#ifdef __cplusplus
}
#endif
#endif
