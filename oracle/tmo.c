/*
 * tmo.c -- CPU oracle (test infrastructure only; see tmo.h).
 *
 * Build: oracle/Makefile, gcc -O2 -ffp-contract=off (no FMA contraction so
 * that every expression rounds exactly as written).  Each routine names the
 * reference routine it restates and keeps its operation order.
 */
#include "tmo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------- Chombo index-box helpers (Box semantics, 2-D) ---------------- */

static tmo_box bx(int lo0, int lo1, int hi0, int hi1)
{
    tmo_box b;
    b.lo[0] = lo0; b.lo[1] = lo1; b.hi[0] = hi0; b.hi[1] = hi1;
    return b;
}
static int bempty(tmo_box b) { return b.hi[0] < b.lo[0] || b.hi[1] < b.lo[1]; }
static tmo_box bgrow(tmo_box b, int n) { b.lo[0] -= n; b.lo[1] -= n; b.hi[0] += n; b.hi[1] += n; return b; }
static tmo_box bgrowd(tmo_box b, int d, int n) { b.lo[d] -= n; b.hi[d] += n; return b; }
static tmo_box band(tmo_box a, tmo_box b)
{
    for (int d = 0; d < 2; ++d) {
        if (b.lo[d] > a.lo[d]) a.lo[d] = b.lo[d];
        if (b.hi[d] < a.hi[d]) a.hi[d] = b.hi[d];
    }
    return a;
}
/* surroundingNodes(d): cells lo..hi -> faces lo..hi+1 */
static tmo_box bnodes(tmo_box b, int d) { b.hi[d] += 1; return b; }
static int bcontains(tmo_box outer, tmo_box inner)
{
    for (int d = 0; d < 2; ++d)
        if (inner.lo[d] < outer.lo[d] || inner.hi[d] > outer.hi[d]) return 0;
    return 1;
}

static size_t fab_npts(const tmo_fab *f)
{
    return (size_t)(f->b.hi[0] - f->b.lo[0] + 1) * (size_t)(f->b.hi[1] - f->b.lo[1] + 1);
}
static double *fab_new(tmo_fab *f, tmo_box b, int nc, double fill)
{
    f->b = b; f->nc = nc;
    size_t n = fab_npts(f) * (size_t)nc;
    f->d = (double *)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; ++k) f->d[k] = fill;
    return f->d;
}
static void fab_free(tmo_fab *f) { free(f->d); f->d = NULL; }

/* Fortran-order element access */
static inline double *AT(const tmo_fab *f, int i, int j, int c)
{
    const int nx = f->b.hi[0] - f->b.lo[0] + 1;
    const int ny = f->b.hi[1] - f->b.lo[1] + 1;
    return f->d + (size_t)(i - f->b.lo[0]) + (size_t)(j - f->b.lo[1]) * nx + (size_t)c * nx * ny;
}

/* a view whose box is shifted by s cells in direction d (same storage) */
static tmo_fab fab_shift(const tmo_fab *f, int d, int s)
{
    tmo_fab g = *f;
    g.b.lo[d] += s; g.b.hi[d] += s;
    return g;
}

#define NAN_FILL (NAN)

/* ---------------- parameters ---------------- */

void tmo_set_material(tmo_params *p, double cs, double cp, double G,
                      double sxx0, double syy0, double szz0,
                      double sxy0, double sxz0, double syz0)
{
    /* LinElastPhysicsF.ChF:1165-1207 */
    p->cs = cs; p->cp = cp; p->G = G;
    p->rho = G / (cs * cs);
    p->lam = (cp * cp) * p->rho - 2.0 * G;
    p->Zs = p->rho * cs;
    p->Zp = p->rho * cp;
    p->K = p->lam + 2.0 * G / 3.0;
    const double s[9] = {0.0, 0.0, 0.0, sxx0, syy0, szz0, sxy0, sxz0, syz0};
    for (int side = 0; side < 2; ++side) {
        for (int c = 0; c < 9; ++c) p->background[11 * side + c] = s[c];
        p->background[11 * side + 9] = 0.0;
        p->background[11 * side + 10] = 0.0;
    }
}

/* direction permutation shared by every elastic kernel (LinElastPhysicsF.ChF:278-301) */
typedef struct { int cx, cy, cz, cxy, cxz, cyz; double rv; } perm_t;
static perm_t perm(int idir)
{
    perm_t q;
    q.cx = idir; q.cy = (idir + 1) % 3; q.cz = (idir + 2) % 3;
    if (idir == 0) { q.cxy = 6; q.cxz = 7; q.cyz = 8; q.rv = 1.0; }
    else if (idir == 1) { q.cxy = 8; q.cxz = 6; q.cyz = 7; q.rv = -1.0; }
    else { q.cxy = 7; q.cxz = 8; q.cyz = 6; q.rv = 1.0; }
    return q;
}

#define OFF 11

/* ---------------- elastic kernels ---------------- */

/* CHARANALYSISF (LinElastPhysicsF.ChF:359-452) */
void tmo_charanalysis(tmo_fab *dW, const tmo_params *p, int idir, tmo_box box)
{
    const perm_t q = perm(idir);
    const double lam = p->lam, G = p->G, Zp = p->Zp, Zs = p->Zs;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            double v[2][11];
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
                const double r = s ? q.rv : 1.0;
                double vx = *AT(dW, i, j, o + q.cx), vy = *AT(dW, i, j, o + q.cy), vz = *AT(dW, i, j, o + q.cz);
                if (s) { vx = r * vx; vy = r * vy; vz = r * vz; }
                const double sxx = *AT(dW, i, j, o + q.cx + 3), syy = *AT(dW, i, j, o + q.cy + 3),
                             szz = *AT(dW, i, j, o + q.cz + 3), sxy = *AT(dW, i, j, o + q.cxy),
                             sxz = *AT(dW, i, j, o + q.cxz), syz = *AT(dW, i, j, o + q.cyz),
                             lm = *AT(dW, i, j, o + 9), gm = *AT(dW, i, j, o + 10);
                v[s][0] = sxx + vx * Zp;
                v[s][1] = sxy + vy * Zs;
                v[s][2] = sxz + vz * Zs;
                v[s][3] = syy - sxx * lam / (2.0 * G + lam);
                v[s][4] = szz - sxx * lam / (2.0 * G + lam);
                v[s][5] = syz;
                v[s][6] = lm;
                v[s][7] = gm;
                v[s][8] = sxz - vz * Zs;
                v[s][9] = sxy - vy * Zs;
                v[s][10] = sxx - vx * Zp;
            }
            for (int k = 0; k < 11; ++k) {
                *AT(dW, i, j, 2 * k) = v[0][k];
                *AT(dW, i, j, 2 * k + 1) = v[1][k];
            }
        }
}

/* CHARSYNTHESISF (LinElastPhysicsF.ChF:458-555) */
void tmo_charsynthesis(tmo_fab *dW, const tmo_params *p, int idir, tmo_box box)
{
    const perm_t q = perm(idir);
    const double lam = p->lam, G = p->G, Zp = p->Zp, Zs = p->Zs;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            double w[2][11];
            for (int k = 0; k < 11; ++k) {
                w[0][k] = *AT(dW, i, j, 2 * k);
                w[1][k] = *AT(dW, i, j, 2 * k + 1);
            }
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
                const double wxxm = w[s][0], wxym = w[s][1], wxzm = w[s][2], wyy = w[s][3], wzz = w[s][4],
                             wyz = w[s][5], lm = w[s][6], gm = w[s][7], wxzp = w[s][8], wxyp = w[s][9],
                             wxxp = w[s][10];
                const double sxx = (wxxm + wxxp) / 2.0;
                if (s == 0) {
                    *AT(dW, i, j, o + q.cx) = (wxxm - wxxp) / 2.0 / Zp;
                    *AT(dW, i, j, o + q.cy) = (wxym - wxyp) / 2.0 / Zs;
                    *AT(dW, i, j, o + q.cz) = (wxzm - wxzp) / 2.0 / Zs;
                } else {
                    *AT(dW, i, j, o + q.cx) = q.rv * (wxxm - wxxp) / 2.0 / Zp;
                    *AT(dW, i, j, o + q.cy) = q.rv * (wxym - wxyp) / 2.0 / Zs;
                    *AT(dW, i, j, o + q.cz) = q.rv * (wxzm - wxzp) / 2.0 / Zs;
                }
                *AT(dW, i, j, o + q.cx + 3) = sxx;
                *AT(dW, i, j, o + q.cy + 3) = wyy + sxx * lam / (2.0 * G + lam);
                *AT(dW, i, j, o + q.cz + 3) = wzz + sxx * lam / (2.0 * G + lam);
                *AT(dW, i, j, o + q.cxy) = (wxym + wxyp) / 2.0;
                *AT(dW, i, j, o + q.cxz) = (wxzm + wxzp) / 2.0;
                *AT(dW, i, j, o + q.cyz) = wyz;
                *AT(dW, i, j, o + 9) = lm;
                *AT(dW, i, j, o + 10) = gm;
            }
        }
}

/* CHARVALUESF (LinElastPhysicsF.ChF:561-595) */
void tmo_charvalues(tmo_fab *lambda, const tmo_params *p, tmo_box box)
{
    double ev[22];
    for (int k = 0; k < 22; ++k) ev[k] = 0.0;
    ev[0] = ev[1] = -p->cp;
    ev[2] = ev[3] = ev[4] = ev[5] = -p->cs;
    ev[16] = ev[17] = ev[18] = ev[19] = p->cs;
    ev[20] = ev[21] = p->cp;
    for (int c = 0; c < 22; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i) *AT(lambda, i, j, c) = ev[c];
}

/* RIEMANNF (LinElastPhysicsF.ChF:601-765) */
void tmo_riemann_kernel(tmo_fab *Wg, const tmo_fab *WL, const tmo_fab *WR,
                        const tmo_params *p, int idir, tmo_box box)
{
    const perm_t q = perm(idir);
    const double lam = p->lam, G = p->G, Zp = p->Zp, Zs = p->Zs;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i)
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
                const double rv = s ? q.rv : 1.0;
                const double Rvx = *AT(WR, i, j, o + q.cx), Rvy = *AT(WR, i, j, o + q.cy),
                             Rvz = *AT(WR, i, j, o + q.cz), Rsxx = *AT(WR, i, j, o + q.cx + 3),
                             Rsyy = *AT(WR, i, j, o + q.cy + 3), Rszz = *AT(WR, i, j, o + q.cz + 3),
                             Rsxy = *AT(WR, i, j, o + q.cxy), Rsxz = *AT(WR, i, j, o + q.cxz),
                             Rsyz = *AT(WR, i, j, o + q.cyz), Rlam = *AT(WR, i, j, o + 9),
                             Rgam = *AT(WR, i, j, o + 10);
                const double Lvx = *AT(WL, i, j, o + q.cx), Lvy = *AT(WL, i, j, o + q.cy),
                             Lvz = *AT(WL, i, j, o + q.cz), Lsxx = *AT(WL, i, j, o + q.cx + 3),
                             Lsyy = *AT(WL, i, j, o + q.cy + 3), Lszz = *AT(WL, i, j, o + q.cz + 3),
                             Lsxy = *AT(WL, i, j, o + q.cxy), Lsxz = *AT(WL, i, j, o + q.cxz),
                             Lsyz = *AT(WL, i, j, o + q.cyz), Llam = *AT(WL, i, j, o + 9),
                             Lgam = *AT(WL, i, j, o + 10);
                double wxxp, wxyp, wxzp, wxxm, wxym, wxzm;
                if (s == 0) {
                    wxxp = Lsxx - Zp * Lvx; wxyp = Lsxy - Zs * Lvy; wxzp = Lsxz - Zs * Lvz;
                    wxxm = Rsxx + Zp * Rvx; wxym = Rsxy + Zs * Rvy; wxzm = Rsxz + Zs * Rvz;
                } else {
                    wxxp = Lsxx - rv * Zp * Lvx; wxyp = Lsxy - rv * Zs * Lvy; wxzp = Lsxz - rv * Zs * Lvz;
                    wxxm = Rsxx + rv * Zp * Rvx; wxym = Rsxy + rv * Zs * Rvy; wxzm = Rsxz + rv * Zs * Rvz;
                }
                double wyy = Rsyy - Rsxx * lam / (2.0 * G + lam);
                double wzz = Rszz - Rsxx * lam / (2.0 * G + lam);
                double wyz = Rsyz;
                wyy = (wyy + Lsyy - Lsxx * lam / (2.0 * G + lam)) / 2.0;
                wzz = (wzz + Lszz - Lsxx * lam / (2.0 * G + lam)) / 2.0;
                wyz = (wyz + Lsyz) / 2.0;
                const double sxx = (wxxm + wxxp) / 2.0;
                if (s == 0) {
                    *AT(Wg, i, j, o + q.cx) = (wxxm - wxxp) / 2.0 / Zp;
                    *AT(Wg, i, j, o + q.cy) = (wxym - wxyp) / 2.0 / Zs;
                    *AT(Wg, i, j, o + q.cz) = (wxzm - wxzp) / 2.0 / Zs;
                } else {
                    *AT(Wg, i, j, o + q.cx) = rv * (wxxm - wxxp) / 2.0 / Zp;
                    *AT(Wg, i, j, o + q.cy) = rv * (wxym - wxyp) / 2.0 / Zs;
                    *AT(Wg, i, j, o + q.cz) = rv * (wxzm - wxzp) / 2.0 / Zs;
                }
                *AT(Wg, i, j, o + q.cx + 3) = (wxxm + wxxp) / 2.0;
                *AT(Wg, i, j, o + q.cy + 3) = wyy + sxx * lam / (2.0 * G + lam);
                *AT(Wg, i, j, o + q.cz + 3) = wzz + sxx * lam / (2.0 * G + lam);
                *AT(Wg, i, j, o + q.cxy) = (wxym + wxyp) / 2.0;
                *AT(Wg, i, j, o + q.cxz) = (wxzm + wxzp) / 2.0;
                *AT(Wg, i, j, o + q.cyz) = wyz;
                *AT(Wg, i, j, o + 9) = (Rlam + Llam) / 2.0;
                *AT(Wg, i, j, o + 10) = (Rgam + Lgam) / 2.0;
            }
}

/* GETFLUXF (LinElastPhysicsF.ChF:265-353) */
void tmo_getflux(tmo_fab *F, const tmo_fab *Wg, const tmo_params *p, int idir, tmo_box box)
{
    const perm_t q = perm(idir);
    const double lam = p->lam, G = p->G, rho = p->rho;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i)
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
                const double vx = *AT(Wg, i, j, o + q.cx), vy = *AT(Wg, i, j, o + q.cy),
                             vz = *AT(Wg, i, j, o + q.cz), sxx = *AT(Wg, i, j, o + q.cx + 3),
                             sxy = *AT(Wg, i, j, o + q.cxy), sxz = *AT(Wg, i, j, o + q.cxz);
                if (s == 0) {
                    *AT(F, i, j, o + q.cx) = -sxx / rho;
                    *AT(F, i, j, o + q.cy) = -sxy / rho;
                    *AT(F, i, j, o + q.cz) = -sxz / rho;
                    *AT(F, i, j, o + q.cx + 3) = -vx * (lam + 2.0 * G);
                    *AT(F, i, j, o + q.cy + 3) = -vx * lam;
                    *AT(F, i, j, o + q.cz + 3) = -vx * lam;
                    *AT(F, i, j, o + q.cxy) = -vy * G;
                    *AT(F, i, j, o + q.cxz) = -vz * G;
                } else {
                    const double rv = q.rv;
                    *AT(F, i, j, o + q.cx) = -(rv * sxx / rho);
                    *AT(F, i, j, o + q.cy) = -(rv * sxy / rho);
                    *AT(F, i, j, o + q.cz) = -(rv * sxz / rho);
                    *AT(F, i, j, o + q.cx + 3) = -(rv * vx * (lam + 2.0 * G));
                    *AT(F, i, j, o + q.cy + 3) = -(rv * vx * lam);
                    *AT(F, i, j, o + q.cz + 3) = -(rv * vx * lam);
                    *AT(F, i, j, o + q.cxy) = -(rv * vy * G);
                    *AT(F, i, j, o + q.cxz) = -(rv * vz * G);
                }
                *AT(F, i, j, o + q.cyz) = 0.0;
                *AT(F, i, j, o + 9) = 0.0;
                *AT(F, i, j, o + 10) = 0.0;
            }
}

/* GETADWDXF (LinElastPhysicsF.ChF:791-901) */
void tmo_getadwdx(tmo_fab *A, const tmo_fab *WHalf, const tmo_params *p,
                  double scale, int idir, tmo_box box)
{
    const perm_t q = perm(idir);
    const double lam = p->lam, G = p->G, rho = p->rho;
    const int io = idir == 0, jo = idir == 1;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i)
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
#define DIF(c) (*AT(WHalf, i + io, j + jo, o + (c)) - *AT(WHalf, i, j, o + (c)))
                const double vx = DIF(q.cx), vy = DIF(q.cy), vz = DIF(q.cz);
                const double sxx = DIF(q.cx + 3), sxy = DIF(q.cxy), sxz = DIF(q.cxz);
#undef DIF
                const double sc = s ? q.rv * scale : scale;
                *AT(A, i, j, o + q.cx) = -(sc * sxx / rho);
                *AT(A, i, j, o + q.cy) = -(sc * sxy / rho);
                *AT(A, i, j, o + q.cz) = -(sc * sxz / rho);
                *AT(A, i, j, o + q.cx + 3) = -(sc * vx * (lam + 2.0 * G));
                *AT(A, i, j, o + q.cy + 3) = -(sc * vx * lam);
                *AT(A, i, j, o + q.cz + 3) = -(sc * vx * lam);
                *AT(A, i, j, o + q.cxy) = -(sc * vy * G);
                *AT(A, i, j, o + q.cxz) = -(sc * vz * G);
                *AT(A, i, j, o + q.cyz) = -(sc * 0.0);
                *AT(A, i, j, o + 9) = -(sc * 0.0);
                *AT(A, i, j, o + 10) = -(sc * 0.0);
            }
}

/* boundary-state assembly shared by LINELASTOUTBCF and LINELASTFREEBCF
 * (LinElastPhysicsF.ChF:968-1014 and :1101-1140) */
static void bc_assemble(tmo_fab *Wg, const tmo_fab *Win, const tmo_params *p, const perm_t *q,
                        int i, int j, int s,
                        double wmx, double wmy, double wmz, double wpx, double wpy, double wpz)
{
    const int o = s * OFF;
    const double lam = p->lam, G = p->G, Zp = p->Zp, Zs = p->Zs;
    const double sxx = (wmx + wpx) / 2.0;
    const double SHF = -((*AT(Win, i, j, o + q->cx) - sxx) * lam / (lam + 2.0 * G));
    if (s == 0) {
        *AT(Wg, i, j, o + q->cx) = (wmx - wpx) / (2.0 * Zp);
        *AT(Wg, i, j, o + q->cy) = (wmy - wpy) / (2.0 * Zs);
        *AT(Wg, i, j, o + q->cz) = (wmz - wpz) / (2.0 * Zs);
    } else {
        *AT(Wg, i, j, o + q->cx) = q->rv * (wmx - wpx) / (2.0 * Zp);
        *AT(Wg, i, j, o + q->cy) = q->rv * (wmy - wpy) / (2.0 * Zs);
        *AT(Wg, i, j, o + q->cz) = q->rv * (wmz - wpz) / (2.0 * Zs);
    }
    *AT(Wg, i, j, o + q->cx + 3) = sxx;
    *AT(Wg, i, j, o + q->cy + 3) = *AT(Win, i, j, o + q->cy + 3) + SHF;
    *AT(Wg, i, j, o + q->cz + 3) = *AT(Win, i, j, o + q->cz + 3) + SHF;
    *AT(Wg, i, j, o + q->cxy) = (wmy + wpy) / 2.0;
    *AT(Wg, i, j, o + q->cxz) = (wmz + wpz) / 2.0;
    *AT(Wg, i, j, o + q->cyz) = *AT(Win, i, j, o + q->cyz);
    *AT(Wg, i, j, o + 9) = *AT(Win, i, j, o + 9);
    *AT(Wg, i, j, o + 10) = *AT(Win, i, j, o + 10);
}

/* LINELASTOUTBCF (LinElastPhysicsF.ChF:1025-1144): absorbing boundary */
void tmo_outbc(tmo_fab *Wg, const tmo_fab *Win, const tmo_params *p, int lohisign, int idir, tmo_box box)
{
    const perm_t q = perm(idir);
    const double Zp = p->Zp, Zs = p->Zs;
    const double *bg = p->background;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i)
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
                const double rv = s ? q.rv : 1.0;
                double wpx, wpy, wpz, wmx, wmy, wmz;
                if (s == 0) {
                    wpx = bg[o + q.cx + 3] - Zp * bg[o + q.cx];
                    wpy = bg[o + q.cxy] - Zs * bg[o + q.cy];
                    wpz = bg[o + q.cxz] - Zs * bg[o + q.cz];
                    wmx = bg[o + q.cx + 3] + Zp * bg[o + q.cx];
                    wmy = bg[o + q.cxy] + Zs * bg[o + q.cy];
                    wmz = bg[o + q.cxz] + Zs * bg[o + q.cz];
                } else {
                    wpx = bg[o + q.cx + 3] - rv * Zp * bg[o + q.cx];
                    wpy = bg[o + q.cxy] - rv * Zs * bg[o + q.cy];
                    wpz = bg[o + q.cxz] - rv * Zs * bg[o + q.cz];
                    wmx = bg[o + q.cx + 3] + rv * Zp * bg[o + q.cx];
                    wmy = bg[o + q.cxy] + rv * Zs * bg[o + q.cy];
                    wmz = bg[o + q.cxz] + rv * Zs * bg[o + q.cz];
                }
                const double vx = *AT(Win, i, j, o + q.cx), vy = *AT(Win, i, j, o + q.cy),
                             vz = *AT(Win, i, j, o + q.cz), sxx = *AT(Win, i, j, o + q.cx + 3),
                             sxy = *AT(Win, i, j, o + q.cxy), sxz = *AT(Win, i, j, o + q.cxz);
                if (lohisign == -1) {
                    if (s == 0) { wmx = sxx + Zp * vx; wmy = sxy + Zs * vy; wmz = sxz + Zs * vz; }
                    else { wmx = sxx + rv * Zp * vx; wmy = sxy + rv * Zs * vy; wmz = sxz + rv * Zs * vz; }
                } else {
                    if (s == 0) { wpx = sxx - Zp * vx; wpy = sxy - Zs * vy; wpz = sxz - Zs * vz; }
                    else { wpx = sxx - rv * Zp * vx; wpy = sxy - rv * Zs * vy; wpz = sxz - rv * Zs * vz; }
                }
                bc_assemble(Wg, Win, p, &q, i, j, s, wmx, wmy, wmz, wpx, wpy, wpz);
            }
}

/* LINELASTFREEBCF (LinElastPhysicsF.ChF:906-1020): traction-free boundary */
void tmo_freebc(tmo_fab *Wg, const tmo_fab *Win, const tmo_params *p, int lohisign, int idir, tmo_box box)
{
    const perm_t q = perm(idir);
    const double Zp = p->Zp, Zs = p->Zs;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i)
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
                const double rv = s ? q.rv : 1.0;
                const double vx = *AT(Win, i, j, o + q.cx), vy = *AT(Win, i, j, o + q.cy),
                             vz = *AT(Win, i, j, o + q.cz), sxx = *AT(Win, i, j, o + q.cx + 3),
                             sxy = *AT(Win, i, j, o + q.cxy), sxz = *AT(Win, i, j, o + q.cxz);
                double wpx, wpy, wpz, wmx, wmy, wmz;
                if (lohisign == -1) {
                    if (s == 0) { wmx = sxx + Zp * vx; wmy = sxy + Zs * vy; wmz = sxz + Zs * vz; }
                    else { wmx = sxx + rv * Zp * vx; wmy = sxy + rv * Zs * vy; wmz = sxz + rv * Zs * vz; }
                    wpx = -wmx; wpy = -wmy; wpz = -wmz;
                } else {
                    if (s == 0) { wpx = sxx - Zp * vx; wpy = sxy - Zs * vy; wpz = sxz - Zs * vz; }
                    else { wpx = sxx - rv * Zp * vx; wpy = sxy - rv * Zs * vy; wpz = sxz - rv * Zs * vz; }
                    wmx = -wpx; wmy = -wpy; wmz = -wpz;
                }
                bc_assemble(Wg, Win, p, &q, i, j, s, wmx, wmy, wmz, wpx, wpy, wpz);
            }
}

/* per-face extension lookup (NULL array -> reference scalar) */
static double face_val(const double *arr, const tmo_params *p, int i, double dflt)
{
    if (!arr) return dflt;
    int k = i - p->face_lo;
    if (k < 0) k = 0;
    if (k >= p->nface) k = p->nface - 1;
    return arr[k];
}

/* GAUSSIANFUNC (LinElastPhysicsF.ChF:1291-1298) */
static double gaussianfunc(double w, double r) { return exp(-(0.5 * ((r / w) * (r / w)))); }
/* SMOOTHFUNC (LinElastPhysicsF.ChF:1303-1315) */
static double smoothfunc(double w, double r)
{
    if (r >= w) return 0.0;
    if (r <= 0.0) return 1.0;
    return exp(r * r / (r * r - w * w));
}
/* BOXCARFUNC (LinElastPhysicsF.ChF:1334-1344) */
static double boxcarfunc(double w, double r) { return (r >= w) ? 0.0 : 1.0; }

/* RSNEWTON (RSIBCF.ChF:336-401) */
double tmo_rsnewton(const tmo_params *p, double phi, double psi, double sn, double c_rs_a, double *Vio, int *iters)
{
    const double atolG = DBL_EPSILON * 1e4, atolV = 0.0, rtolV = 0.0;
    const int IMAX = 100;
    double V = *Vio;
    double Vmin = 0.0, Vmax = phi / p->Zs;
    if (V < Vmin || V > Vmax) V = (Vmin + Vmax) / 2.0;
    const double alpha = sn * c_rs_a;
    const double beta = exp(psi / c_rs_a) / (2.0 * p->rs_V0);
    const double eta = 0.5 * p->Zs;
    double fval = alpha * asinh(beta * V) + eta * V - phi;
    int it = 0;
    for (int i = 1; i <= IMAX; ++i) {
        it = i;
        if (fval < 0.0) Vmin = V; else Vmax = V;
        const double dfval = alpha * beta / sqrt(beta * beta * (V * V) + 1.0) + eta;
        const double dV = -(fval / dfval);
        V = V + dV;
        if (V < Vmin || V > Vmax) V = (Vmin + Vmax) / 2.0;
        fval = alpha * asinh(beta * V) + eta * V - phi;
        if (fabs(fval) <= atolG) break;
        if (fabs(dV) <= atolV + rtolV * (fabs(V) + fabs(dV))) break;
    }
    if (iters) *iters = it;
    *Vio = V;
    return alpha * asinh(beta * V);
}

/* RSFAULTBCF (RSIBCF.ChF:158-323); B supplies psi (comp 10) */
void tmo_rsfaultbc(tmo_fab *Wg, const tmo_fab *Win, const tmo_fab *B, const tmo_params *p, tmo_box box)
{
    const double Zp = p->Zp, Zs = p->Zs, lam = p->lam, G = p->G;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            const double x = (i + 0.5) * p->dx - p->x0;
            const double psi = *AT(B, i, j, 10);
            const double vx1 = *AT(Win, i, j, 0), vy1i = *AT(Win, i, j, 1), vz1 = *AT(Win, i, j, 2);
            const double sxx1 = *AT(Win, i, j, 3), syy1i = *AT(Win, i, j, 4), szz1 = *AT(Win, i, j, 5);
            const double sxy1i = *AT(Win, i, j, 6), sxz1 = *AT(Win, i, j, 7);
            const double vx2 = *AT(Win, i, j, 11), vy2i = *AT(Win, i, j, 12), vz2 = *AT(Win, i, j, 13);
            const double sxx2 = *AT(Win, i, j, 14), syy2i = *AT(Win, i, j, 15), szz2 = *AT(Win, i, j, 16);
            const double sxy2i = *AT(Win, i, j, 17), sxz2 = *AT(Win, i, j, 18);
            const double wpxy1 = -sxy1i - Zs * vx1;
            const double wpyy1 = -syy1i - Zp * vy1i;
            const double wpxy2 = sxy2i - Zs * vx2;
            const double wpyy2 = syy2i - Zp * vy2i;
            const double wpyz1 = 0.0, wpyz2 = 0.0;
            double tau_xy_0 = p->nsig * gaussianfunc(p->r0, x);
            if (p->f_tau) tau_xy_0 = tau_xy_0 + face_val(p->f_tau, p, i, 0.0);
            const double tau_yz_0 = 0.0;
            const double phi_xy = 0.5 * (wpxy2 - wpxy1) + tau_xy_0;
            const double phi_yz = 0.5 * (wpyz2 - wpyz1) + tau_yz_0;
            const double phi = sqrt(phi_xy * phi_xy + phi_yz * phi_yz);
            const double syy1 = -((wpyy1 - wpyy2) / 2.0);
            const double syy2 = -((wpyy1 - wpyy2) / 2.0);
            const double vy1 = -((wpyy1 + wpyy2) / 2.0 / Zp);
            const double vy2 = -((wpyy1 + wpyy2) / 2.0 / Zp);
            double V = sqrt((vx1 - vx2) * (vx1 - vx2) + (vz1 - vz2) * (vz1 - vz2));
            const double c_ramp = fmin(1.0, fmax(0.0, (x - p->ramp_x) / p->ramp_w));
            const double abase = face_val(p->f_a, p, i, p->rs_a);
            const double c_rs_a = abase + p->ramp_a * c_ramp;
            double sn = -syy1;
            if (p->f_syy) sn = -(syy1 + face_val(p->f_syy, p, i, 0.0));
            const double tau = tmo_rsnewton(p, phi, psi, sn, c_rs_a, &V, NULL);
            const double sxy1 = (phi_xy / phi) * tau - tau_xy_0;
            const double syz1 = (phi_yz / phi) * tau - tau_yz_0;
            const double Vx = (phi_xy / phi) * V;
            const double Vz = (phi_yz / phi) * V;
            const double vx1o = -((sxy1 + wpxy1) / Zs);
            const double vz1o = -((syz1 + wpyz1) / Zs);
            double SHF = -((vy1 - syy1) * lam / (lam + 2.0 * G));
            *AT(Wg, i, j, 0) = vx1o;
            *AT(Wg, i, j, 1) = vy1;
            *AT(Wg, i, j, 2) = vz1o;
            *AT(Wg, i, j, 3) = sxx1 + SHF;
            *AT(Wg, i, j, 4) = syy1;
            *AT(Wg, i, j, 5) = szz1 + SHF;
            *AT(Wg, i, j, 6) = sxy1;
            *AT(Wg, i, j, 7) = sxz1;
            *AT(Wg, i, j, 8) = syz1;
            *AT(Wg, i, j, 10) = 0.0;
            *AT(Wg, i, j, 9) = 0.0;
            const double sxy2 = (phi_xy / phi) * tau - tau_xy_0;
            const double syz2 = (phi_yz / phi) * tau - tau_yz_0;
            const double vx2o = vx1o - Vx;
            const double vz2o = vz1o - Vz;
            SHF = -((vy2 - syy2) * lam / (lam + 2.0 * G));
            *AT(Wg, i, j, 11) = vx2o;
            *AT(Wg, i, j, 12) = vy2;
            *AT(Wg, i, j, 13) = vz2o;
            *AT(Wg, i, j, 14) = sxx2 + SHF;
            *AT(Wg, i, j, 15) = syy2;
            *AT(Wg, i, j, 16) = szz2 + SHF;
            *AT(Wg, i, j, 17) = sxy2;
            *AT(Wg, i, j, 18) = sxz2;
            *AT(Wg, i, j, 19) = syz2;
            *AT(Wg, i, j, 21) = 0.0;
            *AT(Wg, i, j, 20) = 0.0;
        }
}

/* RSSETBND (RSIBCF.ChF:61-151), plus the aging-law extension (state_law == 1).
 * Returns through *nan_flag when the new state is NaN (the reference stops). */
static int rssetbnd_impl(tmo_fab *Bout, const tmo_fab *Bin, const tmo_fab *U, const tmo_params *p,
                         double dt, double time, tmo_box box)
{
    int nan_hit = 0;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            const double x = (i + 0.5) * p->dx - p->x0;
            double tau_xy_0 = p->nsig * gaussianfunc(p->r0, x);
            if (p->f_tau) tau_xy_0 = tau_xy_0 + face_val(p->f_tau, p, i, 0.0);
            const double tau_yz_0 = 0.0;
            const double c_ramp = fmin(1.0, fmax(0.0, (x - p->ramp_x) / p->ramp_w));
            const double abase = face_val(p->f_a, p, i, p->rs_a);
            const double c_rs_a = abase + p->ramp_a * c_ramp;
            const double c_rs_Vw = p->rs_Vw + p->ramp_Vw * c_ramp;
            const double L = face_val(p->f_L, p, i, p->rs_L);
            const double b = face_val(p->f_b, p, i, p->rs_b);
            const double Vx = *AT(U, i, j, 0) - *AT(U, i, j, 11);
            const double Vz = *AT(U, i, j, 2) - *AT(U, i, j, 13);
            const double V = sqrt(Vx * Vx + Vz * Vz);
            const double binDX = *AT(Bin, i, j, 6), binDZ = *AT(Bin, i, j, 7), binDT = *AT(Bin, i, j, 8);
            const double binRT = *AT(Bin, i, j, 9), psi = *AT(Bin, i, j, 10);
            *AT(Bout, i, j, 0) = Vx;
            *AT(Bout, i, j, 1) = Vz;
            *AT(Bout, i, j, 2) = V;
            *AT(Bout, i, j, 3) = *AT(U, i, j, 6) + tau_xy_0;
            *AT(Bout, i, j, 4) = *AT(U, i, j, 8) + tau_yz_0;
            *AT(Bout, i, j, 6) = binDX + dt * Vx;
            *AT(Bout, i, j, 7) = binDZ + dt * Vz;
            *AT(Bout, i, j, 8) = binDT + dt * V;
            if (binRT > time && V > p->rs_rupture_thr) *AT(Bout, i, j, 9) = time;
            else if (binRT > time) *AT(Bout, i, j, 9) = 1e10;
            *AT(Bout, i, j, 5) = *AT(U, i, j, 4);
            double dPsi;
            if (p->state_law == 0) {
                const double f = c_rs_a * asinh(V * exp(psi / c_rs_a) / (2.0 * p->rs_V0));
                double fss = p->rs_f0 - (b - c_rs_a) * log(V / p->rs_V0);
                if (p->rs_fExp > 0.0) {
                    const double r = V / c_rs_Vw;
                    const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
                    fss = p->rs_fw + (fss - p->rs_fw) / pow(1.0 + r8, 0.125);
                }
                dPsi = psi - dt * (V / L) * (f - fss);
            } else {
                dPsi = psi + dt * (b * p->rs_V0 / L) * (exp((p->rs_f0 - psi) / b) - V / p->rs_V0);
            }
            if (isnan(dPsi)) nan_hit = 1;
            *AT(Bout, i, j, 10) = dPsi;
        }
    return nan_hit;
}

void tmo_rssetbnd(tmo_fab *Bout, const tmo_fab *Bin, const tmo_fab *U, const tmo_params *p,
                  double dt, double time, tmo_box box)
{
    (void)rssetbnd_impl(Bout, Bin, U, p, dt, time, box);
}

/* SWFAULTBCF (SWIBCF.ChF:79-267), side 2 read and written at IX2_* = 11+n
 * (the shipped routine uses 9+n; see SWIBCF.ChF:139-147,255-263). */
void tmo_swfaultbc(tmo_fab *Wg, const tmo_fab *Win, const tmo_fab *B, const tmo_params *p, tmo_box box)
{
    const double Zp = p->Zp, Zs = p->Zs, lam = p->lam, G = p->G;
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            const double x = (i + 0.5) * p->dx - p->fric_center[0];
            const double vx1 = *AT(Win, i, j, 0), vy1i = *AT(Win, i, j, 1), vz1 = *AT(Win, i, j, 2);
            const double sxx1 = *AT(Win, i, j, 3), syy1i = *AT(Win, i, j, 4), szz1 = *AT(Win, i, j, 5);
            const double sxy1i = *AT(Win, i, j, 6), sxz1 = *AT(Win, i, j, 7), syz1i = *AT(Win, i, j, 8);
            const double vx2 = *AT(Win, i, j, 11), vy2i = *AT(Win, i, j, 12), vz2 = *AT(Win, i, j, 13);
            const double sxx2 = *AT(Win, i, j, 14), syy2i = *AT(Win, i, j, 15), szz2 = *AT(Win, i, j, 16);
            const double sxy2i = *AT(Win, i, j, 17), sxz2 = *AT(Win, i, j, 18), syz2i = *AT(Win, i, j, 19);
            (void)vy1i; (void)vy2i;
            const double wpxy1 = -sxy1i - Zs * vx1;
            const double wpyy1 = -syy1i - Zp * *AT(Win, i, j, 1);
            const double wpyz1 = -syz1i - Zs * vz1;
            const double wpxy2 = sxy2i - Zs * vx2;
            const double wpyy2 = syy2i - Zp * *AT(Win, i, j, 12);
            const double wpyz2 = syz2i - Zs * vz2;
            double tau_0 = 0.0;
            for (int ip = 0; ip < p->num_patches; ++ip) {
                const double valx = smoothfunc(p->smValue, fabs(x - p->xc[ip]) - p->xw[ip]);
                tau_0 = tau_0 + p->tau_p[ip] * valx;
            }
            if (p->f_tau) tau_0 = tau_0 + face_val(p->f_tau, p, i, 0.0);
            const double ta = 0.5 * (wpxy2 - wpxy1) + tau_0;
            const double tb = 0.5 * (wpyz2 - wpyz1);
            const double tau_lock = sqrt(ta * ta + tb * tb);
            const double syy1 = -((wpyy1 - wpyy2) / 2.0);
            const double syy2 = -((wpyy1 - wpyy2) / 2.0);
            const double vy1 = -((wpyy1 + wpyy2) / 2.0 / Zp);
            const double vy2 = -((wpyy1 + wpyy2) / 2.0 / Zp);
            const double valx = boxcarfunc(p->fric_width[0], fabs(x));
            const double fs = face_val(p->f_fs, p, i, p->fs);
            const double fd = face_val(p->f_fd, p, i, p->fd);
            const double dw = face_val(p->f_dc, p, i, p->dw);
            double fric = fs * valx + (1.0 - valx) * p->outside_f;
            fric = fric - (fric - fd) * fmin(1.0, *AT(B, i, j, 7) / dw);
            double sn_syy = syy1;
            if (p->f_syy) sn_syy = syy1 + face_val(p->f_syy, p, i, 0.0);
            const double tau = fmin(tau_lock, -sn_syy * fric) - tau_0;
            const double phi_xy = 0.5 * (wpxy2 - wpxy1);
            const double phi_yz = 0.5 * (wpyz2 - wpyz1);
            const double phi = sqrt(phi_xy * phi_xy + phi_yz * phi_yz);
            const double sxy1 = (phi_xy / phi) * tau;
            const double syz1 = (phi_yz / phi) * tau;
            const double vx1o = -((sxy1 + wpxy1) / Zs);
            const double vz1o = -((syz1 + wpyz1) / Zs);
            double SHF = -((vy1 - syy1) * lam / (lam + 2.0 * G));
            *AT(Wg, i, j, 0) = vx1o;
            *AT(Wg, i, j, 1) = vy1;
            *AT(Wg, i, j, 2) = vz1o;
            *AT(Wg, i, j, 3) = sxx1 + SHF;
            *AT(Wg, i, j, 4) = syy1;
            *AT(Wg, i, j, 5) = szz1 + SHF;
            *AT(Wg, i, j, 6) = sxy1;
            *AT(Wg, i, j, 7) = sxz1;
            *AT(Wg, i, j, 8) = syz1;
            *AT(Wg, i, j, 9) = 0.0;
            *AT(Wg, i, j, 10) = 0.0;
            const double sxy2 = (phi_xy / phi) * tau;
            const double syz2 = (phi_yz / phi) * tau;
            const double vx2o = (sxy2 - wpxy2) / Zs;
            const double vz2o = (syz2 - wpyz2) / Zs;
            SHF = -((vy2 - syy2) * lam / (lam + 2.0 * G));
            *AT(Wg, i, j, 11) = vx2o;
            *AT(Wg, i, j, 12) = vy2;
            *AT(Wg, i, j, 13) = vz2o;
            *AT(Wg, i, j, 14) = sxx2 + SHF;
            *AT(Wg, i, j, 15) = syy2;
            *AT(Wg, i, j, 16) = szz2 + SHF;
            *AT(Wg, i, j, 17) = sxy2;
            *AT(Wg, i, j, 18) = sxz2;
            *AT(Wg, i, j, 19) = syz2;
            *AT(Wg, i, j, 20) = 0.0;
            *AT(Wg, i, j, 21) = 0.0;
        }
}

/* SWSETBND (SWIBCF.ChF:32-72) with side-2 velocities read at IX2_VX=11, IX2_VZ=13
 * (the shipped routine reads 9 and 11).  The rupture test reads Bin(V) as shipped. */
void tmo_swsetbnd(tmo_fab *Bout, const tmo_fab *Bin, const tmo_fab *U, const tmo_params *p,
                  double dt, double time, tmo_box box)
{
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            const double b5 = *AT(Bin, i, j, 5), b6 = *AT(Bin, i, j, 6), b7 = *AT(Bin, i, j, 7);
            const double b8 = *AT(Bin, i, j, 8);
            const double Vx = *AT(U, i, j, 0) - *AT(U, i, j, 11);
            const double Vz = *AT(U, i, j, 2) - *AT(U, i, j, 13);
            *AT(Bout, i, j, 0) = Vx;
            *AT(Bout, i, j, 1) = Vz;
            *AT(Bout, i, j, 2) = *AT(U, i, j, 6);
            *AT(Bout, i, j, 3) = *AT(U, i, j, 8);
            const double V = sqrt(Vx * Vx + Vz * Vz);
            *AT(Bout, i, j, 4) = V;
            *AT(Bout, i, j, 5) = b5 + dt * Vx;
            *AT(Bout, i, j, 6) = b6 + dt * Vz;
            *AT(Bout, i, j, 7) = b7 + dt * V;
            const double bV = *AT(Bin, i, j, 4);
            if (b8 > time && bV > p->sw_rupture_thr) *AT(Bout, i, j, 8) = time;
        }
}

/* PLASTICFLOW + INVARIANTS + YIELD (PlasticPhysicsF.ChF:74-181) */
void tmo_plasticflow(const tmo_params *p, double dt, double *sxx, double *syy, double *szz,
                     double *sxy, double *sxz, double *syz, double *plam, double *pgam)
{
    double sigma = (*sxx + *syy + *szz) / 3.0;
    const double a = *sxx - *syy, b = *syy - *szz, c = *szz - *sxx;
    const double J2 = (a * a + b * b + c * c) / 6.0 + *sxy * *sxy + *sxz * *sxz + *syz * *syz;
    double tau = sqrt(J2);
    const double bb = p->p_b + *pgam * p->p_h;
    const double Y = tau - fmax(-p->p_mu * sigma + bb, 0.0);
    if (Y <= 0.0) { *plam = 0.0; return; }
    const double G = p->G, K = p->K;
    *plam = Y / (dt * (p->p_h + G + p->p_mu * p->p_beta * K) + p->p_eta);
    *pgam = *pgam + dt * *plam;
    const double sig_e = sigma;
    tau = tau - dt * *plam * G;
    sigma = sigma - dt * *plam * K * p->p_beta;
    const double factor = tau / (tau + dt * *plam * G);
    *sxx = sigma + factor * (*sxx - sig_e);
    *syy = sigma + factor * (*syy - sig_e);
    *szz = sigma + factor * (*szz - sig_e);
    *sxy = factor * *sxy;
    *sxz = factor * *sxz;
    *syz = factor * *syz;
}

/* PLASTICCORRECTION (PlasticPhysicsF.ChF:30-69) */
void tmo_plastic_correction(tmo_fab *U, const tmo_params *p, double dt, tmo_box box)
{
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i)
            for (int s = 0; s < 2; ++s) {
                const int o = s * OFF;
                tmo_plasticflow(p, dt, AT(U, i, j, o + 3), AT(U, i, j, o + 4), AT(U, i, j, o + 5),
                                AT(U, i, j, o + 6), AT(U, i, j, o + 7), AT(U, i, j, o + 8),
                                AT(U, i, j, o + 9), AT(U, i, j, o + 10));
            }
}

/* ---------------- GodunovUtilities, restated from the published methods ---------------- */

/* classification of a cell along dir: 0 centered, 1 at domain lo, 2 at domain hi */
static int cell_side(int i, int d, tmo_box domain)
{
    if (i == domain.lo[d]) return 1;
    if (i == domain.hi[d]) return 2;
    return 0;
}

/* centered / one-sided differences (second-order slope differences) */
static void slope_diffs(const tmo_fab *W, int c, int i, int j, int dir, tmo_box domain,
                        double *dWC, double *dWL, double *dWR)
{
    const int io = dir == 0, jo = dir == 1;
    const int ci = dir == 0 ? i : j;
    const int side = cell_side(ci, dir, domain);
    if (side == 0) {
        const double dr = *AT(W, i + io, j + jo, c) - *AT(W, i, j, c);
        const double dl = *AT(W, i, j, c) - *AT(W, i - io, j - jo, c);
        *dWR = dr; *dWL = dl; *dWC = 0.5 * (dr + dl);
    } else if (side == 1) {
        const double dr = *AT(W, i + io, j + jo, c) - *AT(W, i, j, c);
        *dWC = *dWL = *dWR = dr;
    } else {
        const double dl = *AT(W, i, j, c) - *AT(W, i - io, j - jo, c);
        *dWC = *dWL = *dWR = dl;
    }
}

/* van Leer (monotonized central) limiter */
static double vl_limit(double dWc, double dWl, double dWr)
{
    double lim = fmin(2.0 * fabs(dWl), 2.0 * fabs(dWr));
    lim = fmin(lim, fabs(dWc));
    if (dWl * dWr < 0.0) lim = 0.0;
    return lim * copysign(1.0, dWc);
}

void tmo_vanleer_slopes(tmo_fab *dW, const tmo_fab *W, int nc, int limit, int dir,
                        tmo_box box, tmo_box domain)
{
    for (int c = 0; c < nc; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                double C, L, R;
                slope_diffs(W, c, i, j, dir, domain, &C, &L, &R);
                *AT(dW, i, j, c) = limit ? vl_limit(C, L, R) : C;
            }
}

/* one-sided differences W(i)-W(i-1), W(i+1)-W(i) (copies at domain edges) */
static void one_sided_differences(tmo_fab *dWm, tmo_fab *dWp, const tmo_fab *W, int dir,
                                  tmo_box box, tmo_box domain)
{
    for (int c = 0; c < TMO_NCOMP; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                double C, L, R;
                slope_diffs(W, c, i, j, dir, domain, &C, &L, &R);
                *AT(dWm, i, j, c) = L;
                *AT(dWp, i, j, c) = R;
            }
}

static void slope_limiter(tmo_fab *dW, const tmo_fab *dWl, const tmo_fab *dWr, tmo_box box)
{
    for (int c = 0; c < TMO_NCOMP; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i)
                *AT(dW, i, j, c) = vl_limit(*AT(dW, i, j, c), *AT(dWl, i, j, c), *AT(dWr, i, j, c));
}

/* fourth-order slopes from limited second-order slopes (Colella & Sekora form) */
void tmo_fourth_order_slopes(tmo_fab *dW, const tmo_fab *W, const tmo_fab *dWvL, int dir,
                                tmo_box box, tmo_box domain)
{
    const int io = dir == 0, jo = dir == 1;
    for (int c = 0; c < TMO_NCOMP; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                const int ci = dir == 0 ? i : j;
                if (cell_side(ci, dir, domain) != 0) {
                    *AT(dW, i, j, c) = *AT(dWvL, i, j, c);
                } else {
                    const double hi = *AT(W, i + io, j + jo, c) - 0.25 * *AT(dWvL, i + io, j + jo, c);
                    const double lo = *AT(W, i - io, j - jo, c) + 0.25 * *AT(dWvL, i - io, j - jo, c);
                    *AT(dW, i, j, c) = (2.0 / 3.0) * (hi - lo);
                }
            }
}

/* PPM face values, Colella & Woodward (1984) eq. 1.6/1.9 on a uniform grid */
void tmo_ppm_face_values(tmo_fab *WFace, const tmo_fab *W, int nc, int limit, int dir,
                         tmo_box facebox, tmo_box domain)
{
    tmo_box cells = facebox;
    cells.hi[dir] -= 1;                        /* enclosedCells */
    cells = band(bgrowd(cells, dir, 1), domain);
    tmo_fab dW;
    fab_new(&dW, cells, nc, NAN_FILL);
    tmo_vanleer_slopes(&dW, W, nc, limit, dir, cells, domain);
    const int io = dir == 0, jo = dir == 1;
    const double sixth = 1.0 / 6.0;
    for (int c = 0; c < nc; ++c)
        for (int j = facebox.lo[1]; j <= facebox.hi[1]; ++j)
            for (int i = facebox.lo[0]; i <= facebox.hi[0]; ++i) {
                const int f = dir == 0 ? i : j;
                double v;
                if (f == domain.lo[dir]) {
                    v = *AT(W, i, j, c) - 0.5 * *AT(&dW, i, j, c);
                } else if (f == domain.hi[dir] + 1) {
                    v = *AT(W, i - io, j - jo, c) + 0.5 * *AT(&dW, i - io, j - jo, c);
                } else {
                    v = 0.5 * (*AT(W, i, j, c) + *AT(W, i - io, j - jo, c))
                      + sixth * (*AT(&dW, i - io, j - jo, c) - *AT(&dW, i, j, c));
                }
                *AT(WFace, i, j, c) = v;
            }
    fab_free(&dW);
}

/* PPM limiter of Colella & Woodward (1984) eq. 1.10 on the face increments */
void tmo_ppm_limiter(tmo_fab *dWm, tmo_fab *dWp, tmo_box box)
{
    for (int c = 0; c < TMO_NCOMP; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                const double m = *AT(dWm, i, j, c), pp = *AT(dWp, i, j, c);
                if (m * pp < 0.0) {
                    *AT(dWm, i, j, c) = copysign(fmin(2.0 * fabs(pp), fabs(m)), m);
                    *AT(dWp, i, j, c) = copysign(fmin(2.0 * fabs(m), fabs(pp)), pp);
                } else {
                    *AT(dWm, i, j, c) = 0.0;
                    *AT(dWp, i, j, c) = 0.0;
                }
            }
}

/* PPM characteristic tracing (Colella & Woodward 1984 eq. 1.12), per field */
void tmo_ppm_normal_pred(tmo_fab *dWm, tmo_fab *dWp, const tmo_fab *lambda, double dtbydx, tmo_box box)
{
    for (int c = 0; c < TMO_NCOMP; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                const double lk = *AT(lambda, i, j, c);
                const double sm = -fmin(lk, 0.0) * dtbydx;
                const double sp = fmax(lk, 0.0) * dtbydx;
                const double m = *AT(dWm, i, j, c), pp = *AT(dWp, i, j, c);
                *AT(dWm, i, j, c) = m + sm * ((pp - m) - (pp + m) * (3.0 - 2.0 * sm)) / 2.0;
                *AT(dWp, i, j, c) = pp + sp * ((m - pp) - (pp + m) * (3.0 - 2.0 * sp)) / 2.0;
            }
}

/* PLM characteristic predictor (Colella 1990 reference-state form) */
void tmo_plm_normal_pred(tmo_fab *dWm, tmo_fab *dWp, const tmo_fab *dW, const tmo_fab *lambda,
                            double dtbydx, tmo_box box)
{
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            const double lmin = fmin(*AT(lambda, i, j, 0), 0.0);
            const double lmax = fmax(*AT(lambda, i, j, TMO_NCOMP - 1), 0.0);
            for (int c = 0; c < TMO_NCOMP; ++c) {
                const double lk = *AT(lambda, i, j, c), d = *AT(dW, i, j, c);
                if (lk > 0.0) {
                    *AT(dWm, i, j, c) = -(0.5 * (1.0 + dtbydx * lmin) * d);
                    *AT(dWp, i, j, c) = 0.5 * (1.0 - dtbydx * lk) * d;
                } else if (lk < 0.0) {
                    *AT(dWm, i, j, c) = -(0.5 * (1.0 + dtbydx * lk) * d);
                    *AT(dWp, i, j, c) = 0.5 * (1.0 - dtbydx * lmax) * d;
                } else {
                    *AT(dWm, i, j, c) = -(0.5 * (1.0 + dtbydx * lmin) * d);
                    *AT(dWp, i, j, c) = 0.5 * (1.0 - dtbydx * lmax) * d;
                }
            }
        }
}

/* ---------------- patch integrator ---------------- */

typedef struct {
    const tmo_params *p;
    tmo_fab *B;        /* m_bdryData */
    tmo_fab tmpB;      /* m_tmpBdryData */
    int tmp_set;       /* m_tmpBdryDataSet */
    int err;
} ibc_t;

static void fab_add(tmo_fab *a, const tmo_fab *b, tmo_box box)
{
    for (int c = 0; c < a->nc; ++c)
        for (int j = box.lo[1]; j <= box.hi[1]; ++j)
            for (int i = box.lo[0]; i <= box.hi[0]; ++i) *AT(a, i, j, c) = *AT(a, i, j, c) + *AT(b, i, j, c);
}

/* RSIBC::primBC / SWIBC::primBC (RSIBC.cpp:179-273, SWIBC.cpp:161-271).
 * wgbox is the face box of the WGdnv FArrayBox. */
static void prim_bc(ibc_t *ibc, tmo_fab *Wg, tmo_box wgbox, const tmo_fab *Win, int dir, int lohisign)
{
    const tmo_params *p = ibc->p;
    tmo_box tmp = wgbox;
    /* shiftHalf(dir, lohisign) of a face box: Lo -> cells flo-1..fhi-1, Hi -> cells flo..fhi */
    if (lohisign == -1) { tmp.lo[dir] -= 1; tmp.hi[dir] -= 1; }
    if (bcontains(p->domain, tmp)) return;
    tmp = band(tmp, p->domain);
    if (bempty(tmp)) return;
    tmo_box bb = tmp;
    if (lohisign == -1) { bb.hi[dir] = bb.lo[dir]; }
    else { bb.lo[dir] = bb.hi[dir] + 1; bb.hi[dir] = bb.lo[dir]; }
    if (lohisign == -1 && dir == 1) {
        tmo_fab *Bsrc = ibc->tmp_set ? &ibc->tmpB : ibc->B;
        if (!Bsrc || !Bsrc->d) { ibc->err = -4; return; }
        if (p->fault_kind == 0) tmo_rsfaultbc(Wg, Win, Bsrc, p, bb);
        else tmo_swfaultbc(Wg, Win, Bsrc, p, bb);
        return;
    }
    const int t = p->bc[dir * 2 + (lohisign + 1) / 2];
    if (t == 0) tmo_outbc(Wg, Win, p, lohisign, dir, bb);
    else if (t == 1) tmo_freebc(Wg, Win, p, lohisign, dir, bb);
    else ibc->err = -2;
}

/* LinElastPhysics::riemann (LinElastPhysics.cpp:301-359) */
static void riemann(ibc_t *ibc, tmo_fab *WStar, tmo_box wsbox, const tmo_fab *WPlus, const tmo_fab *WMinus,
                    int dir, tmo_box facebox)
{
    const tmo_fab WL = fab_shift(WPlus, dir, 1);   /* shiftHalf(dir,+1): cell i -> face i+1 */
    const tmo_fab WR = fab_shift(WMinus, dir, 0);  /* shiftHalf(dir,-1): cell i -> face i   */
    if (!bempty(facebox)) tmo_riemann_kernel(WStar, &WL, &WR, ibc->p, dir, facebox);
    prim_bc(ibc, WStar, wsbox, &WL, dir, +1);
    prim_bc(ibc, WStar, wsbox, &WR, dir, -1);
}

/* RSIBC::updateBoundary / SWIBC::updateBoundary (RSIBC.cpp:295-336, SWIBC.cpp:293-331) */
static void update_boundary(ibc_t *ibc, const tmo_fab *WHalf, int dir, double dt, double time, int final)
{
    const tmo_params *p = ibc->p;
    if (dir != 1) return;
    tmo_box fb = WHalf->b;             /* face box of WHalf */
    tmo_box bl = fb; bl.hi[1] = bl.lo[1];  /* bdryLo(WHalf.box(),1) */
    tmo_box dlo = p->domain; dlo.hi[1] = dlo.lo[1];  /* bdryLo(domain,1) */
    if (!bcontains(dlo, bl)) return;
    if (!ibc->B || !ibc->B->d) { ibc->err = -4; return; }
    if (final) {
        if (p->fault_kind == 0) { if (rssetbnd_impl(ibc->B, ibc->B, WHalf, p, dt, time, bl)) ibc->err = -1; }
        else tmo_swsetbnd(ibc->B, ibc->B, WHalf, p, dt, time, bl);
    } else {
        if (p->fault_kind == 0) { if (rssetbnd_impl(&ibc->tmpB, ibc->B, WHalf, p, dt, time, bl)) ibc->err = -1; }
        else tmo_swsetbnd(&ibc->tmpB, ibc->B, WHalf, p, dt, time, bl);
        ibc->tmp_set = 1;
    }
}

/* LinElastPatchGodunov::PLMNormalPred / PPMNormalPred / CTUNormalPred */
static void normal_pred(const tmo_params *p, tmo_fab *WM, tmo_fab *WP, const tmo_fab *W,
                        double dt, int dir, tmo_box box)
{
    const int nc = TMO_NCOMP;
    const int limit = p->char_limiting || p->prim_limiting;
    tmo_fab lambda;
    if (p->normal_pred == 0) {
        for (int c = 0; c < nc; ++c)
            for (int j = box.lo[1]; j <= box.hi[1]; ++j)
                for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                    *AT(WM, i, j, c) = *AT(W, i, j, c);
                    *AT(WP, i, j, c) = *AT(W, i, j, c);
                }
        return;
    }
    fab_new(&lambda, box, nc, 0.0);
    if (p->normal_pred == 1) {
        tmo_fab dW;
        fab_new(&dW, box, nc, NAN_FILL);
        if (p->fourth_order) {
            tmo_box boxVL = band(bgrowd(box, dir, 1), p->domain);
            tmo_fab dWvL;
            fab_new(&dWvL, boxVL, nc, NAN_FILL);
            tmo_vanleer_slopes(&dWvL, W, nc, limit, dir, boxVL, p->domain);
            tmo_fourth_order_slopes(&dW, W, &dWvL, dir, box, p->domain);
            fab_free(&dWvL);
        } else {
            tmo_vanleer_slopes(&dW, W, nc, limit, dir, box, p->domain);
        }
        for (size_t k = 0; k < fab_npts(WM) * nc; ++k) { WM->d[k] = 0.0; WP->d[k] = 0.0; }
        if (p->char_limiting || p->fourth_order) one_sided_differences(WM, WP, W, dir, box, p->domain);
        tmo_charvalues(&lambda, p, box);
        if (p->char_limiting) {
            tmo_charanalysis(WM, p, dir, box);
            tmo_charanalysis(WP, p, dir, box);
            tmo_charanalysis(&dW, p, dir, box);
        }
        if (p->char_limiting || p->prim_limiting)
            if (p->char_limiting || p->fourth_order) slope_limiter(&dW, WM, WP, box);
        if (!p->char_limiting) tmo_charanalysis(&dW, p, dir, box);
        tmo_plm_normal_pred(WM, WP, &dW, &lambda, dt / p->dx, box);
        fab_free(&dW);
    } else {
        tmo_box facebox = bnodes(box, dir);
        tmo_fab WFace;
        fab_new(&WFace, facebox, nc, NAN_FILL);
        tmo_ppm_face_values(&WFace, W, nc, limit, dir, facebox, p->domain);
        const int io = dir == 0, jo = dir == 1;
        for (int c = 0; c < nc; ++c)
            for (int j = box.lo[1]; j <= box.hi[1]; ++j)
                for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                    const double w = *AT(W, i, j, c);
                    *AT(WM, i, j, c) = (0.0 - w) + *AT(&WFace, i, j, c);
                    *AT(WP, i, j, c) = (0.0 - w) + *AT(&WFace, i + io, j + jo, c);
                }
        fab_free(&WFace);
        tmo_charvalues(&lambda, p, box);
        if (p->char_limiting) {
            tmo_charanalysis(WM, p, dir, box);
            tmo_charanalysis(WP, p, dir, box);
        }
        if (p->char_limiting || p->prim_limiting) tmo_ppm_limiter(WM, WP, box);
        if (!p->char_limiting) {
            tmo_charanalysis(WM, p, dir, box);
            tmo_charanalysis(WP, p, dir, box);
        }
        tmo_ppm_normal_pred(WM, WP, &lambda, dt / p->dx, box);
    }
    tmo_charsynthesis(WM, p, dir, box);
    tmo_charsynthesis(WP, p, dir, box);
    fab_add(WM, W, box);
    fab_add(WP, W, box);
    fab_free(&lambda);
}

int tmo_update_state(tmo_fab *U, tmo_fab *B, tmo_fab *F0, tmo_fab *F1,
                     const tmo_params *p, tmo_box box, double time, double dt)
{
    const int nc = TMO_NCOMP;
    if (p->normal_pred < 0 || p->normal_pred > 2) return -3;
    if (p->char_limiting && p->prim_limiting && p->normal_pred != 0) return -3;
    ibc_t ibc;
    memset(&ibc, 0, sizeof ibc);
    ibc.p = p;
    ibc.B = B;
    if (B) fab_new(&ibc.tmpB, B->b, B->nc, NAN_FILL);  /* RSIBC::setBdryData */

    /* WHalf (FluxBox, zeroed: LinElastPatchGodunov.cpp:114-115) */
    tmo_fab WHalf[2];
    for (int d = 0; d < 2; ++d) fab_new(&WHalf[d], bnodes(box, d), nc, 0.0);

    /* ---- computeWHalf (LinElastPatchGodunov.cpp:174-454) ---- */
    tmo_box faceBox[2], fluxBox[2], ccBox[2];
    for (int d = 0; d < 2; ++d) {
        faceBox[d] = bnodes(bgrowd(band(bgrow(box, 1), p->domain), d, -1), d);
        ccBox[d] = band(bgrowd(bgrow(box, 1), d, -1), p->domain);
        fluxBox[d] = bnodes(ccBox[d], d);
    }
    tmo_box WBox = band(U->b, p->domain);
    tmo_fab W;
    fab_new(&W, WBox, nc, 0.0);
    for (int c = 0; c < nc; ++c)
        for (int j = WBox.lo[1]; j <= WBox.hi[1]; ++j)
            for (int i = WBox.lo[0]; i <= WBox.hi[0]; ++i) *AT(&W, i, j, c) = *AT(U, i, j, c);
    tmo_box slopeBox = band(bgrow(box, 1), p->domain);
    tmo_fab WM[2], WP[2], WHalf1[2];
    for (int d = 0; d < 2; ++d) {
        fab_new(&WM[d], slopeBox, nc, NAN_FILL);
        fab_new(&WP[d], slopeBox, nc, NAN_FILL);
        normal_pred(p, &WM[d], &WP[d], &W, dt, d, slopeBox);
        fab_new(&WHalf1[d], fluxBox[d], nc, NAN_FILL);
        riemann(&ibc, &WHalf1[d], fluxBox[d], &WP[d], &WM[d], d, faceBox[d]);
    }
    for (int d = 0; d < 2; ++d) {
        faceBox[d] = bnodes(bgrowd(band(bgrowd(box, d, 1), p->domain), d, -1), d);
        fluxBox[d] = bnodes(box, d);
    }
    const double scale = -(1.0 / 2.0) * dt / p->dx;
    for (int d1 = 0; d1 < 2; ++d1) {
        for (int d2 = 0; d2 < 2; ++d2) {
            if (d2 == d1) continue;
            tmo_fab A;
            fab_new(&A, slopeBox, nc, NAN_FILL);
            tmo_getadwdx(&A, &WHalf1[d2], p, scale, d2, ccBox[d2]);
            fab_add(&WM[d1], &A, slopeBox);
            tmo_getadwdx(&A, &WHalf1[d2], p, scale, d2, ccBox[d2]);
            fab_add(&WP[d1], &A, slopeBox);
            fab_free(&A);
            /* quasilinearBoundaryUpdate -> updateBoundary(..., time 0, final false) */
            update_boundary(&ibc, &WHalf1[d2], d2, (1.0 / 2.0) * dt, 0.0, 0);
        }
        riemann(&ibc, &WHalf[d1], fluxBox[d1], &WP[d1], &WM[d1], d1, faceBox[d1]);
    }

    /* ---- computeUpdate (LinElastPhysics.cpp:135-179) ---- */
    tmo_fab dU;
    fab_new(&dU, U->b, nc, 0.0);
    for (int d = 0; d < 2; ++d) {
        tmo_fab Fd;
        tmo_fab *F = d == 0 ? F0 : F1;
        if (F) Fd = *F; else fab_new(&Fd, bnodes(box, d), nc, 0.0);
        for (size_t k = 0; k < fab_npts(&Fd) * nc; ++k) Fd.d[k] = 0.0;   /* a_F.setVal(0) */
        tmo_getflux(&Fd, &WHalf[d], p, d, Fd.b);
        const int io = d == 0, jo = d == 1;
        for (int c = 0; c < nc; ++c)
            for (int j = box.lo[1]; j <= box.hi[1]; ++j)
                for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
                    const double diff = *AT(&Fd, i + io, j + jo, c) - *AT(&Fd, i, j, c);   /* FLUXDIFFF */
                    *AT(&dU, i, j, c) = *AT(&dU, i, j, c) + diff;
                }
        update_boundary(&ibc, &WHalf[d], d, dt, time + dt, 1);
        if (!F) fab_free(&Fd);
    }
    const double sc = -dt / p->dx;
    for (int c = 0; c < nc; ++c)
        for (int j = U->b.lo[1]; j <= U->b.hi[1]; ++j)
            for (int i = U->b.lo[0]; i <= U->b.hi[0]; ++i) {
                const double du = *AT(&dU, i, j, c) * sc;
                *AT(U, i, j, c) = *AT(U, i, j, c) + du;
            }

    fab_free(&dU);
    for (int d = 0; d < 2; ++d) {
        fab_free(&WM[d]); fab_free(&WP[d]); fab_free(&WHalf1[d]); fab_free(&WHalf[d]);
    }
    fab_free(&W);
    if (B) fab_free(&ibc.tmpB);
    return ibc.err;
}

/* GETFULLGRADF + MAXF over both directions (AMRLevelLinElast.cpp:610-662) */
void tmo_fullgrad(tmo_fab *grad, const tmo_fab *U, tmo_box box, tmo_box domain)
{
    for (int j = box.lo[1]; j <= box.hi[1]; ++j)
        for (int i = box.lo[0]; i <= box.hi[0]; ++i) {
            double g[2];
            for (int d = 0; d < 2; ++d) {
                const int io = d == 0, jo = d == 1;
                const int ci = d == 0 ? i : j;
                double diff = 0.0;
                for (int iv = 0; iv < 3; ++iv) {
                    double v;
                    if (ci == domain.lo[d]) v = fabs(*AT(U, i + io, j + jo, iv) - *AT(U, i, j, iv));
                    else if (ci == domain.hi[d]) v = fabs(*AT(U, i, j, iv) - *AT(U, i - io, j - jo, iv));
                    else v = 0.5 * fabs(*AT(U, i + io, j + jo, iv) - *AT(U, i - io, j - jo, iv));
                    diff = fmax(diff, v);
                }
                g[d] = diff;
            }
            double cur = 0.0;
            for (int d = 0; d < 2; ++d) cur = fmax(cur, fabs(g[d]));
            *AT(grad, i, j, 0) = cur;
        }
}

/* ---------------- uniform-grid level driver ---------------- */

/* LinElastLevelGodunov::step semantics on a level split into boxes of at most
 * maxbox x maxbox cells: every box reads the pre-step level (its 4-ghost copy),
 * updates its valid cells, and fault data is copied back for valid faces. */
static int level_step(double *Uold, double *Unew, double *Bold, double *Bnew, const tmo_params *p,
                      int nx, int ny, double time, double dt, int maxbox, int nthreads)
{
    const int g = 4, nc = TMO_NCOMP;
    const int nb = p->fault_kind == 0 ? TMO_NB_RS : TMO_NB_SW;
    const tmo_box lb = bx(-g, -g, nx - 1 + g, ny - 1 + g);
    const int LX = nx + 2 * g, LY = ny + 2 * g;
    const int bxn = (nx + maxbox - 1) / maxbox, byn = (ny + maxbox - 1) / maxbox;
    const int nboxes = bxn * byn;
    int err = 0;
    (void)lb;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(min : err)
#endif
    for (int k = 0; k < nboxes; ++k) {
        const int bi = k % bxn, bj = k / bxn;
        tmo_box box = bx(bi * maxbox, bj * maxbox, (bi + 1) * maxbox - 1, (bj + 1) * maxbox - 1);
        if (box.hi[0] > nx - 1) box.hi[0] = nx - 1;
        if (box.hi[1] > ny - 1) box.hi[1] = ny - 1;
        tmo_fab U;
        fab_new(&U, bgrow(box, g), nc, 0.0);
        for (int c = 0; c < nc; ++c)
            for (int j = U.b.lo[1]; j <= U.b.hi[1]; ++j)
                for (int i = U.b.lo[0]; i <= U.b.hi[0]; ++i)
                    *AT(&U, i, j, c) = Uold[(size_t)(i + g) + (size_t)(j + g) * LX + (size_t)c * LX * LY];
        tmo_fab Bf, *Bp = NULL;
        if (box.lo[1] == 0) {
            fab_new(&Bf, bx(box.lo[0] - g, 0, box.hi[0] + g, 0), nb, 0.0);
            for (int c = 0; c < nb; ++c)
                for (int i = Bf.b.lo[0]; i <= Bf.b.hi[0]; ++i)
                    *AT(&Bf, i, 0, c) = Bold[(size_t)(i + g) + (size_t)c * (nx + 2 * g)];
            Bp = &Bf;
        }
        const int e = tmo_update_state(&U, Bp, NULL, NULL, p, box, time, dt);
        if (e < err) err = e;
        if (p->use_plasticity) tmo_plastic_correction(&U, p, dt, box);
        for (int c = 0; c < nc; ++c)
            for (int j = box.lo[1]; j <= box.hi[1]; ++j)
                for (int i = box.lo[0]; i <= box.hi[0]; ++i)
                    Unew[(size_t)(i + g) + (size_t)(j + g) * LX + (size_t)c * LX * LY] = *AT(&U, i, j, c);
        if (Bp) {
            for (int c = 0; c < nb; ++c)
                for (int i = box.lo[0]; i <= box.hi[0]; ++i)
                    Bnew[(size_t)(i + g) + (size_t)c * (nx + 2 * g)] = *AT(&Bf, i, 0, c);
            fab_free(&Bf);
        }
        fab_free(&U);
    }
    return err;
}

int tmo_level_steps_boxed(double *U, double *B, const tmo_params *p, int nx, int ny,
                          double time, double dt, int nsteps, int maxbox, int nthreads)
{
    const int g = 4, nc = TMO_NCOMP;
    const int nb = p->fault_kind == 0 ? TMO_NB_RS : TMO_NB_SW;
    const size_t nU = (size_t)(nx + 2 * g) * (ny + 2 * g) * nc;
    const size_t nB = (size_t)(nx + 2 * g) * nb;
    double *U2 = (double *)malloc(nU * sizeof(double));
    double *B2 = (double *)malloc(nB * sizeof(double));
    memcpy(U2, U, nU * sizeof(double));
    memcpy(B2, B, nB * sizeof(double));
    int err = 0;
    double t = time;
    for (int s = 0; s < nsteps && err == 0; ++s) {
        err = level_step(U, U2, B, B2, p, nx, ny, t, dt, maxbox, nthreads);
        t = t + dt;
        memcpy(U, U2, nU * sizeof(double));
        memcpy(B, B2, nB * sizeof(double));
    }
    free(U2);
    free(B2);
    return err;
}

int tmo_level_steps(double *U, double *B, const tmo_params *p, int nx, int ny,
                    double time, double dt, int nsteps, int nthreads)
{
    const int mb = nx > ny ? nx : ny;
    return tmo_level_steps_boxed(U, B, p, nx, ny, time, dt, nsteps, mb, nthreads);
}

/* ---------------- refinement tagging ---------------- */

int tmo_tag_level(const double *U, const double *B, const tmo_params *p, const tmo_tagparams *tp,
                  int nx, int ny, unsigned char *tags)
{
    const int g = 4;
    const int LX = nx + 2 * g, LY = ny + 2 * g;
    tmo_fab Uf;
    Uf.d = (double *)U;
    Uf.b = bx(-g, -g, nx - 1 + g, ny - 1 + g);
    Uf.nc = TMO_NCOMP;
    (void)LY;
    const tmo_box dom = bx(0, 0, nx - 1, ny - 1);
    tmo_fab grad;
    fab_new(&grad, dom, 1, 0.0);
    tmo_fullgrad(&grad, &Uf, dom, dom);
    /* IBC override of the gradient magnitude on the nucleation cells (RSIBC.cpp:351-368,
     * SWIBC.cpp:341-372) */
    if (p->fault_kind == 0) {
        const int i0 = (int)floor(p->x0 / p->dx), i1 = (int)ceil(p->x0 / p->dx);
        for (int i = i0; i <= i1; ++i)
            if (i >= 0 && i < nx) *AT(&grad, i, 0, 0) = 1.0;
    } else {
        for (int k = 0; k < p->num_patches; ++k) {
            const int i0 = (int)floor((p->fric_center[0] + p->xc[k] - p->xw[k]) / p->dx);
            const int i1 = (int)ceil((p->fric_center[0] + p->xc[k] + p->xw[k]) / p->dx);
            for (int i = i0; i <= i1; ++i)
                if (i >= 0 && i < nx) *AT(&grad, i, 0, 0) = 1.0;
        }
    }
    unsigned char *raw = (unsigned char *)calloc((size_t)nx * ny, 1);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double r = 0.0;
            const double x = i * p->dx;
            r = fmax(r, fmax(tp->x_coarsen[0] - x, x - tp->x_coarsen[1]));
            const double y = j * p->dx;
            r = fmax(r, fmax(tp->y_coarsen[0] - y, y - tp->y_coarsen[1]));
            const double thr = tp->refine_thresh * (1 + r * tp->slope_coarsen / tp->width_coarsen);
            if (!(tp->level < tp->coarsen2_max_level ||
                  (x >= tp->x_coarsen2[0] && x <= tp->x_coarsen2[1] && y >= tp->y_coarsen2[0] &&
                   y <= tp->y_coarsen2[1])))
                continue;
            const double *u = U + (size_t)(i + g) + (size_t)(j + g) * LX;
            const size_t pl = (size_t)LX * LY;
            int t = *AT(&grad, i, j, 0) >= thr || u[10 * pl] >= tp->plastic_thresh || u[21 * pl] >= tp->plastic_thresh;
            if (!t && tp->stress_thresh > 0.0) {
                double gs = 0.0;
                for (int d = 0; d < 2; ++d) {
                    const int io = d == 0, jo = d == 1;
                    const int ci = d == 0 ? i : j;
                    for (int c = 3; c <= 6; ++c) {
                        if (c == 5) continue;   /* sxx, syy, sxy */
                        double v;
                        if (ci == dom.lo[d]) v = fabs(*AT(&Uf, i + io, j + jo, c) - *AT(&Uf, i, j, c));
                        else if (ci == dom.hi[d]) v = fabs(*AT(&Uf, i, j, c) - *AT(&Uf, i - io, j - jo, c));
                        else v = 0.5 * fabs(*AT(&Uf, i + io, j + jo, c) - *AT(&Uf, i - io, j - jo, c));
                        gs = fmax(gs, v);
                    }
                }
                t = gs >= tp->stress_thresh;
            }
            if (!t && tp->sliprate_thresh > 0.0 && j == 0 && B) {
                const int vc = p->fault_kind == 0 ? 2 : 4;
                t = B[(size_t)(i + g) + (size_t)vc * LX] >= tp->sliprate_thresh;
            }
            raw[(size_t)i + (size_t)j * nx] = (unsigned char)t;
        }
    /* IntVectSet::grow(tag_buffer), then clip to the domain */
    const int b = tp->tag_buffer;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            unsigned char t = 0;
            for (int jj = j - b; jj <= j + b && !t; ++jj)
                for (int ii = i - b; ii <= i + b && !t; ++ii)
                    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) t = raw[(size_t)ii + (size_t)jj * nx];
            tags[(size_t)i + (size_t)j * nx] = t;
        }
    free(raw);
    fab_free(&grad);
    return 0;
}
