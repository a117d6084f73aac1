/*
 * tmo.h -- CPU oracle for the tetemoko per-patch Godunov update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the pytest suite,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it to
 * compare against the HIP product path.  Nothing in tetemoko_amd/ includes,
 * links or calls it, and it must never be used as a fallback for the product.
 *
 * It is a plain-C restatement of the reference algorithm, written stage by
 * stage in the same order as the reference C++/Fortran so that the operation
 * order (and therefore rounding) of every value matches:
 *   srcLinElast/LinElastPatchGodunov.cpp:101-702   (updateState, computeWHalf, predictors)
 *   srcLinElast/LinElastPhysics.cpp:135-554        (computeUpdate, riemann, quasilinear update)
 *   srcLinElast/LinElastPhysicsF.ChF               (all elastic kernels)
 *   srcLinElast/PlasticPhysicsF.ChF                (Drucker-Prager return)
 *   ibcLinElast/RSIBC.cpp:115-336, RSIBCF.ChF      (rate-and-state fault)
 *   ibcLinElast/SWIBC.cpp:121-331, SWIBCF.ChF      (slip-weakening fault, side-2 index corrected)
 *   srcLinElast/LinElastLevelGodunov.cpp:174-398,643-667 (level step)
 *   srcLinElast/AMRLevelLinElast.cpp:577-717       (tagging)
 * Chombo's GodunovUtilities (PPM/PLM, slope limiters) are not available; the
 * functions marked "restated from the published method" follow Colella &
 * Woodward (1984) / Colella (1990) and are parity-unpinned against Chombo.
 *
 * Storage follows Chombo's FArrayBox: Fortran order over an inclusive index
 * box, x fastest, then y, then component.
 */
#ifndef TMO_H
#define TMO_H

#ifdef __cplusplus
extern "C" {
#endif

#define TMO_NCOMP 22
#define TMO_NB_RS 11
#define TMO_NB_SW 9
#define TMO_MAXPATCH 8

typedef struct { int lo[2]; int hi[2]; } tmo_box;

typedef struct {
    double *d;
    tmo_box b;
    int nc;
} tmo_fab;

typedef struct {
    /* /IBC/ common (IBCCommon.fh:4-6), derived by tmo_set_material */
    double cs, cp, G, rho, lam, K, Zp, Zs;
    double background[TMO_NCOMP];
    /* /PLASTIC/ common (PlasticCommon.fh:4-5); p_b is the cohesion extension */
    double p_mu, p_beta, p_eta, p_b, p_h;
    int use_plasticity;
    /* boundary type per [dir*2 + side]: 0 outflow, 1 free surface.  The lo-y face is the fault. */
    int bc[4];
    int fault_kind; /* 0 rate-and-state (RSIBC), 1 slip-weakening (SWIBC) */
    /* /rs/ common (RSCommon.fh:4-5) + RSIBC::m_psi */
    double r0, nucTime, nsig, rs_a, rs_b, rs_V0, rs_f0, rs_L, rs_fw, rs_Vw, rs_fExp;
    double x0, y0, rs_rupture_thr, ramp_x, ramp_w, ramp_a, ramp_Vw;
    double rs_psi0;
    int state_law; /* 0: reference slip-law form (RSIBCF.ChF:127-132); 1: aging law (extension) */
    /* /slipweak/ common (SWCommon.fh:4-5) + SWIBC members */
    double fs, fd, dw, smValue, sw_rupture_thr;
    int num_patches;
    double xc[TMO_MAXPATCH], xw[TMO_MAXPATCH], zc[TMO_MAXPATCH], zw[TMO_MAXPATCH], tau_p[TMO_MAXPATCH];
    double fric_center[2], fric_width[2], outside_f;
    /* Per-face extension arrays indexed by (face i - face_lo), NULL = reference scalar. */
    int face_lo, nface;
    const double *f_a, *f_b, *f_L;      /* rate-and-state a, b, L */
    const double *f_fs, *f_fd, *f_dc;   /* slip-weakening mu_s, mu_d, Dc */
    const double *f_syy;                /* fault normal-stress offset added to syy at the fault */
    const double *f_tau;                /* fault shear-stress offset added to tau_0 */
    /* numerics (amrGodunov.cpp:298-355) */
    int normal_pred;  /* 0 CTU, 1 PLM, 2 PPM */
    int fourth_order, prim_limiting, char_limiting;
    double dx;
    tmo_box domain;
} tmo_params;

/* LINELASTSETF (LinElastPhysicsF.ChF:1149-1210) */
void tmo_set_material(tmo_params *p, double cs, double cp, double G,
                      double sxx0, double syy0, double szz0,
                      double sxy0, double sxz0, double syz0);

/* --- pointwise / box kernels (box arguments are face or cell index boxes) --- */
void tmo_charanalysis(tmo_fab *dW, const tmo_params *p, int idir, tmo_box box);
void tmo_charsynthesis(tmo_fab *dW, const tmo_params *p, int idir, tmo_box box);
void tmo_charvalues(tmo_fab *lambda, const tmo_params *p, tmo_box box);
void tmo_riemann_kernel(tmo_fab *Wg, const tmo_fab *WL, const tmo_fab *WR,
                        const tmo_params *p, int idir, tmo_box box);
void tmo_getflux(tmo_fab *F, const tmo_fab *Wg, const tmo_params *p, int idir, tmo_box box);
void tmo_getadwdx(tmo_fab *A, const tmo_fab *WHalf, const tmo_params *p,
                  double scale, int idir, tmo_box box);
void tmo_outbc(tmo_fab *Wg, const tmo_fab *Win, const tmo_params *p, int lohisign, int idir, tmo_box box);
void tmo_freebc(tmo_fab *Wg, const tmo_fab *Win, const tmo_params *p, int lohisign, int idir, tmo_box box);
void tmo_rsfaultbc(tmo_fab *Wg, const tmo_fab *Win, const tmo_fab *B, const tmo_params *p, tmo_box box);
double tmo_rsnewton(const tmo_params *p, double phi, double psi, double sn, double a, double *V, int *iters);
void tmo_rssetbnd(tmo_fab *Bout, const tmo_fab *Bin, const tmo_fab *U, const tmo_params *p,
                  double dt, double time, tmo_box box);
void tmo_swfaultbc(tmo_fab *Wg, const tmo_fab *Win, const tmo_fab *B, const tmo_params *p, tmo_box box);
void tmo_swsetbnd(tmo_fab *Bout, const tmo_fab *Bin, const tmo_fab *U, const tmo_params *p,
                  double dt, double time, tmo_box box);
void tmo_plasticflow(const tmo_params *p, double dt, double *sxx, double *syy, double *szz,
                     double *sxy, double *sxz, double *syz, double *lam, double *gam);
void tmo_plastic_correction(tmo_fab *U, const tmo_params *p, double dt, tmo_box box);

/* restated GodunovUtilities (parity-unpinned against Chombo) */
void tmo_vanleer_slopes(tmo_fab *dW, const tmo_fab *W, int nc, int limit, int dir,
                        tmo_box box, tmo_box domain);
void tmo_ppm_face_values(tmo_fab *WFace, const tmo_fab *W, int nc, int limit, int dir,
                         tmo_box facebox, tmo_box domain);
/* the remaining restated GodunovUtilities stages, exported for the pin tests; each works on
 * all TMO_NCOMP components of its arrays over box */
void tmo_fourth_order_slopes(tmo_fab *dW, const tmo_fab *W, const tmo_fab *dWvL, int dir,
                             tmo_box box, tmo_box domain);
void tmo_ppm_limiter(tmo_fab *dWm, tmo_fab *dWp, tmo_box box);
void tmo_ppm_normal_pred(tmo_fab *dWm, tmo_fab *dWp, const tmo_fab *lambda, double dtbydx, tmo_box box);
void tmo_plm_normal_pred(tmo_fab *dWm, tmo_fab *dWp, const tmo_fab *dW, const tmo_fab *lambda,
                         double dtbydx, tmo_box box);

/* --- patch integrator: LinElastPatchGodunov::updateState ---
 * U   : ghosted state (22 comps), updated in place over U->b
 * B   : fault data for this box (NULL when the box does not touch the lo-y boundary)
 * F0,F1: optional flux outputs on surroundingNodes(box,d) (may be NULL)
 * returns 0, or a negative error code (-1 NaN psi, -2 bad BC, -3 bad predictor)     */
int tmo_update_state(tmo_fab *U, tmo_fab *B, tmo_fab *F0, tmo_fab *F1,
                     const tmo_params *p, tmo_box box, double time, double dt);

/* AMRLevelLinElast tagging gradient: GETFULLGRADF/MAXF in both directions */
void tmo_fullgrad(tmo_fab *grad, const tmo_fab *U, tmo_box box, tmo_box domain);

/* Refinement tagging (AMRLevelLinElast::tagCells, AMRLevelLinElast.cpp:577-717, and the IBC
 * override RSIBC::tagCellsInit / SWIBC::tagCellsInit) on a uniform level: U ghosted by 4,
 * B the level's fault data; tags is nx*ny bytes, x fastest.  stress_thresh and
 * sliprate_thresh (> 0) add the stress-gradient and fault slip-rate criteria (extensions). */
typedef struct {
    /* stress_thresh: undivided difference of sxx, syy, sxy (side 1); sliprate_thresh: fault V */
    double refine_thresh, plastic_thresh;
    int tag_buffer;
    double x_coarsen[2], y_coarsen[2], slope_coarsen, width_coarsen;
    double x_coarsen2[2], y_coarsen2[2];
    int level, coarsen2_max_level;
    double stress_thresh, sliprate_thresh;
} tmo_tagparams;
int tmo_tag_level(const double *U, const double *B, const tmo_params *p, const tmo_tagparams *tp,
                  int nx, int ny, unsigned char *tags);

/* Uniform-grid level driver used by the tests and the CPU baseline:
 * nx x ny valid cells at (0,0), a single box, U ghosted by 4, B with 4 x-ghosts. */
int tmo_level_steps(double *U, double *B, const tmo_params *p, int nx, int ny,
                    double time, double dt, int nsteps, int nthreads);
/* same, with the level split into boxes of at most maxbox x maxbox cells */
int tmo_level_steps_boxed(double *U, double *B, const tmo_params *p, int nx, int ny,
                          double time, double dt, int nsteps, int maxbox, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
