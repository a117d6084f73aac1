// tmk_kernels.hpp -- fused per-tile Godunov update for gfx950 (CDNA4, wave64).
//
// One workgroup updates one tile of TX x TY valid cells of one box
// (LinElastPatchGodunov::updateState, LinElastPatchGodunov.cpp:101-129, run on
// a sub-rectangle; every value the reference computes is a function of its
// location only, so tiling a box reproduces the box result exactly).
// Thread (lx, ly) owns cell (x0-1+lx, y0-1+ly) of the tile grown by one ring
// (the reference's slopeBox, LinElastPatchGodunov.cpp:244-246).
//
// Stages (each ends with a workgroup barrier):
//   1  x-sweep stencil staged in LDS, normal predictor in x (per thread)
//   2  y-sweep stencil staged in LDS, normal predictor in y
//   3  predictor Riemann states WHalf1 on the lo x- and y-face of each thread,
//      domain boundary states and the fault solve (RIEMANNF / RSIBC::primBC)
//   4  transverse correction W(d1) += A(d2) dWHalf1(d2)  (GETADWDXF)
//   5  final Riemann states on valid faces, fault solve and fault-data update
//   6  flux difference, U += dU, optional Drucker-Prager return
//
// Only the components a sweep can move are carried (tmk_physics.hpp): the
// state planes vz/sxz/syz are skipped when the anti-plane fields are zero
// (NS == 2), and lam/gam are only touched by the plastic return.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

#include "tetemoko_hip.h"
#include "tmk_physics.hpp"

namespace tmk {

// one box of the level as seen by the kernel
template <typename T>
struct Patch {
    const T* in[22];   // state planes read this step
    T* out[22];        // state planes written this step (== in for pointwise planes)
    int ulo0, ulo1, unx; // U array box (ghosted) lo and x extent
    unsigned pbytes;   // bytes of one state plane (the bound of buffer descriptors over a plane)
    const double* bin; // fault data read (NULL: box has none)
    double* bout;      // fault data written
    int blo0, bnx;     // fault-data x lo and extent
    int bnb;           // fault-data components (11 rate-and-state, 9 slip weakening)
    double* F[2];      // optional face fluxes (22 comps, on surroundingNodes(box,d)) or NULL
    int vlo0, vlo1, vhi0, vhi1; // valid box
    int scen;          // index of the scenario (parameters, time, step) this box belongs to
    // with F: per-cell flags over the valid box grown by 1 (bit 0: a cell of the level, bit 1:
    // covered by the next finer level); NULL keeps every face, else only the faces between a
    // cell of the level and one outside it, or between covered and uncovered cells of the level
    const unsigned char* fmask;
};

struct StateVec {
    double c[22];
};

// one scenario's parameters, clock and error word; device resident so that
// boxes of several scenarios can share one launch
struct Scen {
    DevParams P;
    double time;  // time at the start of the current step
    double dt;
    // step constants, evaluated on the host with the expressions the update uses:
    // dt/dx, -(1/2) dt/dx (transverse correction), -dt/dx (flux difference), and in FP32
    double dtdx, hdtdx, mdtdx;
    float fdtdx, fhdtdx, fmdtdx;
    int err;      // first error reported by any launch (0 = none)
    int pad;
};

// host: the clock and step constants of a scenario
__host__ inline void scen_set_clock(Scen& s, double time, double dt)
{
    s.time = time;
    s.dt = dt;
    s.dtdx = dt / s.P.dx;
    s.hdtdx = -(1.0 / 2.0) * dt / s.P.dx;
    s.mdtdx = -dt / s.P.dx;
    s.fdtdx = (float)s.dtdx;
    s.fhdtdx = (float)s.hdtdx;
    s.fmdtdx = (float)s.mdtdx;
}

// step constants in the bulk arithmetic type
template <typename T>
struct StepConst {
    T dtbydx, scale, sc;  // dt/dx, -(1/2) dt/dx, -dt/dx
};
template <typename T>
__device__ __forceinline__ StepConst<T> step_const(const Scen& S)
{
    if constexpr (sizeof(T) == sizeof(double))
        return {(T)S.dtdx, (T)S.hdtdx, (T)S.mdtdx};
    else
        return {(T)S.fdtdx, (T)S.fhdtdx, (T)S.fmdtdx};
}

struct Tile {
    int patch;
    int x0, y0, x1, y1;  // valid cells of this tile (inclusive)
};

template <typename T>
struct KArgs {
    Scen* scen;
    const Patch<T>* patches;
    const Tile* tiles;
    int ntiles;
    int plastic;   // 0: skip the Drucker-Prager return even if the scenario enables it
};

__device__ __forceinline__ void report(int* err, int code)
{
    // vector atomic from the lanes that hit the error
    atomicCAS(err, 0, code);
}

template <typename T>
__device__ __forceinline__ size_t uidx(const Patch<T>& p, int i, int j, size_t plane)
{
    return (size_t)(i - p.ulo0) + (size_t)(j - p.ulo1) * (size_t)p.unx;
}

// stage-1/2 staging: component list per direction, [side][slot][v|s]
template <int NS>
__device__ __forceinline__ int stage_comp(int d, int c)
{
    const int s = c / (2 * NS);
    const int r = c - s * 2 * NS;
    return slot_comp(d, r >> 1, r & 1) + 11 * s;
}

// GETFLUXF (LinElastPhysicsF.ChF:303-350) from the face state of one side, in slot form:
// f[0]=F(vn) f[1]=F(snn) f[2]=F(vt) f[3]=F(snt) f[4]=F(s_transverse normal) f[5]=F(va) f[6]=F(sna);
// r is the side's velocity sign in the face direction (a product with r = 1 is exact)
template <typename T, int NS>
__device__ __forceinline__ void face_flux_side(const Mat<T>& m, const T (&v)[NS], const T (&s)[NS], T r, T (&f)[7])
{
    f[0] = -m.divrho(r * s[0]);
    f[2] = -m.divrho(r * s[1]);
    f[1] = -(r * v[0] * m.lam2G);
    f[4] = -(r * v[0] * m.lam);
    f[3] = -(r * v[1] * m.G);
    if (NS == 3) {
        f[5] = -m.divrho(r * s[NS - 1]);
        f[6] = -(r * v[NS - 1] * m.G);
    }
}

template <typename T, int NS>
__device__ __forceinline__ void face_flux(const Mat<T>& m, const Slots<T, NS>& S, T rv, T (&f)[2][7])
{
    face_flux_side<T, NS>(m, S.v[0], S.s[0], (T)1, f[0]);
    face_flux_side<T, NS>(m, S.v[1], S.s[1], rv, f[1]);
}

// the fault solve's input from the lo-y face traces W (FP64, absolute: stores_delta adds the
// background) and the face states H from its output (the background removed in FP64)
template <typename T, int NS>
__device__ __forceinline__ void fault_input(const DevParams& P, const Slots<T, NS>& W, FaultIn& fin)
{
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        fin.vy[s] = (double)W.v[s][0]; fin.syy[s] = (double)W.s[s][0];
        fin.vx[s] = (double)W.v[s][1]; fin.sxy[s] = (double)W.s[s][1];
        fin.vz[s] = NS == 3 ? (double)W.v[s][NS - 1] : 0.0;
        fin.syz[s] = NS == 3 ? (double)W.s[s][NS - 1] : 0.0;
        if constexpr (stores_delta<T>()) {
            const double* b = P.bg + 11 * s;
            fin.vx[s] += b[0]; fin.vy[s] += b[1]; fin.syy[s] += b[4]; fin.sxy[s] += b[6];
            if (NS == 3) { fin.vz[s] += b[2]; fin.syz[s] += b[8]; }
        }
    }
}
template <typename T, int NS>
__device__ __forceinline__ void fault_output(const DevParams& P, const FaultOut& fo, Slots<T, NS>& H)
{
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        double vy = fo.vy[s], syy = fo.syy[s], vx = fo.vx[s], sxy = fo.sxy[s], vz = fo.vz[s], syz = fo.syz[s];
        if constexpr (stores_delta<T>()) {
            const double* b = P.bg + 11 * s;
            vx -= b[0]; vy -= b[1]; syy -= b[4]; sxy -= b[6]; vz -= b[2]; syz -= b[8];
        }
        H.v[s][0] = (T)vy; H.s[s][0] = (T)syy;
        H.v[s][1] = (T)vx; H.s[s][1] = (T)sxy;
        if (NS == 3) { H.v[s][NS - 1] = (T)vz; H.s[s][NS - 1] = (T)syz; }
    }
}

// the 11 flux components of side sd from face state (v, s) on the lo d-face of cell (i, j) into
// the flux register array of direction d (surroundingNodes(valid box, d)).  With stores_delta
// this is the flux of U - bg: flux registers only ever difference fluxes of one precision, and
// the drop-in path adds the background's flux (bg_flux) on the host.
// whether a face between cells with face-mask bits a and b carries a register flux: a face of
// the level's boundary (exactly one side a cell of the level), or, between two cells of the
// level, a face around the finer level (exactly one side covered by it)
__device__ __forceinline__ bool register_face(unsigned a, unsigned b)
{
    const bool edge = ((a ^ b) & 1) != 0;           // level boundary (fine side)
    const bool cf = (a & b & 1) && ((a ^ b) & 2);  // around the finer level (coarse side)
    return edge || cf;
}

// the 11 components of side sd of the face flux in direction d at index fi of the flux array
// Fo (fp faces per component plane)
template <typename T, int NS>
__device__ __forceinline__ void put_face_flux(__attribute__((address_space(1))) double* Fo, size_t fi, size_t fp,
                                              const Mat<T>& m, int d, const T (&v)[NS], const T (&sv)[NS], T r,
                                              int sd)
{
    T f[7];
    face_flux_side<T, NS>(m, v, sv, r, f);
    const int o = 11 * sd;
    if (d == 0) {
        Fo[fi + (o + 0) * fp] = (double)f[0];
        Fo[fi + (o + 1) * fp] = (double)f[2];
        Fo[fi + (o + 2) * fp] = NS == 3 ? (double)f[5] : -(sd ? 1.0 : 1.0) * 0.0 * 0.0;
        Fo[fi + (o + 3) * fp] = (double)f[1];
        Fo[fi + (o + 4) * fp] = (double)f[4];
        Fo[fi + (o + 5) * fp] = (double)f[4];
        Fo[fi + (o + 6) * fp] = (double)f[3];
        Fo[fi + (o + 7) * fp] = NS == 3 ? (double)f[6] : 0.0;
        Fo[fi + (o + 8) * fp] = 0.0;
    } else {
        Fo[fi + (o + 1) * fp] = (double)f[0];
        Fo[fi + (o + 0) * fp] = (double)f[2];
        Fo[fi + (o + 2) * fp] = NS == 3 ? (double)f[5] : 0.0;
        Fo[fi + (o + 4) * fp] = (double)f[1];
        Fo[fi + (o + 5) * fp] = (double)f[4];
        Fo[fi + (o + 3) * fp] = (double)f[4];
        Fo[fi + (o + 6) * fp] = (double)f[3];
        Fo[fi + (o + 8) * fp] = NS == 3 ? (double)f[6] : 0.0;
        Fo[fi + (o + 7) * fp] = 0.0;
    }
    Fo[fi + (o + 9) * fp] = 0.0;
    Fo[fi + (o + 10) * fp] = 0.0;
}

template <typename T, int NS>
__device__ __forceinline__ void store_flux_side(const Patch<T>& pt, const Mat<T>& m, int d, const T (&v)[NS],
                                                const T (&sv)[NS], T r, int i, int j, int sd)
{
    const int fnx = pt.vhi0 - pt.vlo0 + 1 + (d == 0), fny = pt.vhi1 - pt.vlo1 + 1 + (d == 1);
    if (pt.fmask) {
        const int mw = pt.vhi0 - pt.vlo0 + 3;
        const __attribute__((address_space(1))) unsigned char* mk =
            gmem(pt.fmask) + (size_t)(i - pt.vlo0 + 1) + (size_t)(j - pt.vlo1 + 1) * mw;
        if (!register_face(d == 0 ? mk[-1] : mk[-mw], *mk)) return;
    }
    const size_t fp = (size_t)fnx * fny;
    const size_t fi = (size_t)(i - pt.vlo0) + (size_t)(j - pt.vlo1) * fnx;
    put_face_flux<T, NS>(gmem(pt.F[d]), fi, fp, m, d, v, sv, r, sd);
}

template <typename T, int NS>
__device__ __forceinline__ void store_flux(const Patch<T>& pt, const Mat<T>& m, int d, const Slots<T, NS>& Fs, T rv,
                                           int i, int j)
{
    store_flux_side<T, NS>(pt, m, d, Fs.v[0], Fs.s[0], (T)1, i, j, 0);
    store_flux_side<T, NS>(pt, m, d, Fs.v[1], Fs.s[1], rv, i, j, 1);
}

// face states of one side on the four faces of a cell: lo/hi x face, lo/hi y face
template <typename T, int NS>
struct Faces1 {
    T xlv[NS], xls[NS], xhv[NS], xhs[NS], ylv[NS], yls[NS], yhv[NS], yhs[NS];
};

// New values of one side's state of a valid cell from its old values u[0..10] (side-local
// component index) and the final states on its four faces: flux difference and
// U += dU (-dt/dx) (LinElastPatchGodunov.cpp:117-126, LinElastPhysics.cpp:135-179), then, with
// plastic, the Drucker-Prager return (LinElastLevelGodunov::plasticUpdate after the level step,
// AMRLevelLinElast.cpp:476).  Changes u[0 1 3 4 5 6] (+ 2 7 8 with NS 3, + 7 8 9 10 with plastic).
// PP: DevParams, or PlasticConstL (FP32) / PlasticConstL64 (FP64): the plastic return's constants in LDS
template <typename T, int NS, typename PP>
__device__ __forceinline__ void cell_update_side(const PP& P, const Mat<T>& m, double dt, T sc, bool plastic,
                                                 T ry, const Faces1<T, NS>& F, T (&u)[11])
{
    if constexpr (regroup<T>()) {
        // the flux is linear in the face state: difference the face states first, then one
        // uniform coefficient per component (-sc / rho, -sc lam2G, -sc lam, -sc G)
        const T xs0 = F.xhs[0] - F.xls[0], xs1 = F.xhs[1] - F.xls[1];
        const T xv0 = F.xhv[0] - F.xlv[0], xv1 = F.xhv[1] - F.xlv[1];
        const T ys0 = F.yhs[0] - F.yls[0], ys1 = F.yhs[1] - F.yls[1];
        const T yv0 = F.yhv[0] - F.ylv[0], yv1 = F.yhv[1] - F.ylv[1];
        const T kv = -(sc * m.irho), kl2 = -(sc * m.lam2G), kl = -(sc * m.lam), kg = -(sc * m.G);
        const T ryv0 = ry * yv0;
        u[0] = fma(kv, fma(ry, ys1, xs0), u[0]);
        u[1] = fma(kv, fma(ry, ys0, xs1), u[1]);
        u[3] = fma(kl2, xv0, fma(kl, ryv0, u[3]));
        u[4] = fma(kl, xv0, fma(kl2, ryv0, u[4]));
        u[5] = fma(kl, xv0 + ryv0, u[5]);
        u[6] = fma(kg, fma(ry, yv1, xv1), u[6]);
        if (NS == 3) {
            const T xs2 = F.xhs[NS - 1] - F.xls[NS - 1], ys2 = F.yhs[NS - 1] - F.yls[NS - 1];
            const T xv2 = F.xhv[NS - 1] - F.xlv[NS - 1], yv2 = F.yhv[NS - 1] - F.ylv[NS - 1];
            u[2] = fma(kv, fma(ry, ys2, xs2), u[2]);
            u[7] = fma(kg, xv2, u[7]);
            u[8] = fma(kg, ry * yv2, u[8]);
        }
        if (plastic) plastic_flow_delta<T, PP>(P, (T)dt, u[3], u[4], u[5], u[6], u[7], u[8], u[9], u[10]);
        return;
    }
    T fxl[7], fxh[7], fyl[7], fyh[7];
    face_flux_side<T, NS>(m, F.xlv, F.xls, (T)1, fxl);
    face_flux_side<T, NS>(m, F.xhv, F.xhs, (T)1, fxh);
    face_flux_side<T, NS>(m, F.ylv, F.yls, ry, fyl);
    face_flux_side<T, NS>(m, F.yhv, F.yhs, ry, fyh);
    // dU = (0 + (Fx_hi - Fx_lo)) + (Fy_hi - Fy_lo); U += dU * (-dt/dx)
    // x flux: vx<-F(vn) vy<-F(vt) sxx<-F(snn) syy,szz<-F(stt) sxy<-F(snt)
    // y flux: vy<-F(vn) vx<-F(vt) syy<-F(snn) sxx,szz<-F(stt) sxy<-F(snt)
    const T dvx_x = fxh[0] - fxl[0], dvx_y = fyh[2] - fyl[2];
    const T dvy_x = fxh[2] - fxl[2], dvy_y = fyh[0] - fyl[0];
    const T dsxx_x = fxh[1] - fxl[1], dsxx_y = fyh[4] - fyl[4];
    const T dsyy_x = fxh[4] - fxl[4], dsyy_y = fyh[1] - fyl[1];
    const T dszz_x = fxh[4] - fxl[4], dszz_y = fyh[4] - fyl[4];
    const T dsxy_x = fxh[3] - fxl[3], dsxy_y = fyh[3] - fyl[3];
    u[0] = u[0] + (((T)0 + dvx_x) + dvx_y) * sc;
    u[1] = u[1] + (((T)0 + dvy_x) + dvy_y) * sc;
    u[3] = u[3] + (((T)0 + dsxx_x) + dsxx_y) * sc;
    u[4] = u[4] + (((T)0 + dsyy_x) + dsyy_y) * sc;
    u[5] = u[5] + (((T)0 + dszz_x) + dszz_y) * sc;
    u[6] = u[6] + (((T)0 + dsxy_x) + dsxy_y) * sc;
    if (NS == 3) {
        const T dvz_x = fxh[5] - fxl[5], dvz_y = fyh[5] - fyl[5];
        const T dsxz_x = fxh[6] - fxl[6], dsxz_y = (T)0 - (T)0;
        const T dsyz_x = (T)0 - (T)0, dsyz_y = fyh[6] - fyl[6];
        u[2] = u[2] + (((T)0 + dvz_x) + dvz_y) * sc;
        u[7] = u[7] + (((T)0 + dsxz_x) + dsxz_y) * sc;
        u[8] = u[8] + (((T)0 + dsyz_x) + dsyz_y) * sc;
    }
    if constexpr (!regroup<T>())
        if (plastic) plastic_flow<T, PP>(P, (T)dt, u[3], u[4], u[5], u[6], u[7], u[8], u[9], u[10]);
}

// U += dU of side sd of valid cell (i, j) from the final states on its four faces, then the
// optional Drucker-Prager return (cell_update_side)
template <typename T, int NS>
__device__ __forceinline__ void update_cell_side(const KArgs<T>& a, const Patch<T>& pt, const DevParams& P,
                                                 const Mat<T>& m, double dt, T sc, int i, int j, int sd,
                                                 const Faces1<T, NS>& F)
{
    const T ry = sd ? (T)-1 : (T)1;
    const unsigned u = (unsigned)uidx(pt, i, j, 0);
    const int o = 11 * sd;
    const bool plastic = P.plast && a.plastic;
    T w[11];
    w[0] = gld(pt.in[o + 0], u);
    w[1] = gld(pt.in[o + 1], u);
    w[3] = gld(pt.in[o + 3], u);
    w[4] = gld(pt.in[o + 4], u);
    w[5] = gld(pt.in[o + 5], u);
    w[6] = gld(pt.in[o + 6], u);
    // mode II keeps sxz, syz zero (the plastic return scales them); lam is only written
    w[2] = NS == 3 ? gld(pt.in[o + 2], u) : (T)0;
    w[7] = NS == 3 ? gld(pt.in[o + 7], u) : (T)0;
    w[8] = NS == 3 ? gld(pt.in[o + 8], u) : (T)0;
    w[9] = (T)0;
    w[10] = plastic ? gld(pt.in[o + 10], u) : (T)0;
    cell_update_side<T, NS>(P, m, dt, sc, plastic, ry, F, w);
    gst(pt.out[o + 0], u, w[0]);
    gst(pt.out[o + 1], u, w[1]);
    if (NS == 3) gst(pt.out[o + 2], u, w[2]);
    gst(pt.out[o + 3], u, w[3]);
    gst(pt.out[o + 4], u, w[4]);
    gst(pt.out[o + 5], u, w[5]);
    gst(pt.out[o + 6], u, w[6]);
    if (NS == 3) {
        gst(pt.out[o + 7], u, w[7]);
        gst(pt.out[o + 8], u, w[8]);
    }
    if (plastic) {
        gst(pt.out[o + 9], u, w[9]);
        gst(pt.out[o + 10], u, w[10]);
    }
}

// both sides
template <typename T, int NS>
__device__ __forceinline__ void update_cell(const KArgs<T>& a, const Patch<T>& pt, const DevParams& P,
                                            const Mat<T>& m, double dt, T sc, int i, int j, const Slots<T, NS>& F0,
                                            const Slots<T, NS>& F0h, const Slots<T, NS>& F1,
                                            const Slots<T, NS>& F1h)
{
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        Faces1<T, NS> F;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            F.xlv[k] = F0.v[sd][k]; F.xls[k] = F0.s[sd][k];
            F.xhv[k] = F0h.v[sd][k]; F.xhs[k] = F0h.s[sd][k];
            F.ylv[k] = F1.v[sd][k]; F.yls[k] = F1.s[sd][k];
            F.yhv[k] = F1h.v[sd][k]; F.yhs[k] = F1h.s[sd][k];
        }
        update_cell_side<T, NS>(a, pt, P, m, dt, sc, i, j, sd, F);
    }
}

// INT: the tile is full size and its stencil (+-3 cells) lies inside the domain, so no cell of
// it is a domain-boundary or fault cell; the body then carries no boundary code at all.
template <typename T, int NS, int BX, int BY, int PK, bool INT>
__device__ __forceinline__ void godunov_tile(const KArgs<T>& a, const Tile& tl, T* sm)
{
    constexpr int NT = BX * BY;
    constexpr int NC = 4 * NS;  // staged components per direction
    const Patch<T>& pt = a.patches[tl.patch];
    Scen& S = a.scen[pt.scen];
    const DevParams& P = S.P;
    const Mat<T> m(P);
    int* const errp = &S.err;
    const int tid = threadIdx.x;
    const int lx = tid % BX, ly = tid / BX;
    const int i = tl.x0 - 1 + lx, j = tl.y0 - 1 + ly;
    const int tnx = INT ? BX - 2 : tl.x1 - tl.x0 + 1, tny = INT ? BY - 2 : tl.y1 - tl.y0 + 1;
    const int dlo0 = P.dlo[0], dlo1 = P.dlo[1], dhi0 = P.dhi[0], dhi1 = P.dhi[1];
    const bool colD = INT || (i >= dlo0 && i <= dhi0);
    const bool rowD = INT || (j >= dlo1 && j <= dhi1);
    const bool r1 = colD && rowD && lx <= tnx + 1 && ly <= tny + 1;
    const bool valid = lx >= 1 && lx <= tnx && ly >= 1 && ly <= tny;
    const double time = S.time;
    const double dt = S.dt;
    const StepConst<T> K = step_const<T>(S);
    const T dtbydx = K.dtbydx;
    const int pred = P.pred;
    const T rv1 = (T)-1;

    Slots<T, NS> WmX, WpX, WmY, WpY;

    // Staging of a sweep's stencil planes into LDS: the planes in an unrolled outer loop (each
    // plane's address uniform), every load of this thread issued before the first LDS store, so
    // that the loads' latencies overlap (one plane address per element, read through a varying
    // index, had put two dependent global loads and waits on every element).
    const int ulo0 = pt.ulo0, ulo1 = pt.ulo1, unx = pt.unx;
    auto stage = [&](int d, int RW, int RH, int gx0, int gy0, int hx, int hy) {
        // plane c of sweep d over the RW x RH region from (gx0, gy0); only the tile's stencil
        // region (cells up to tl.x1 + hx, tl.y1 + hy), inside the domain, is read
        const int n = RW * RH;
        constexpr int IT = (((BY + 4) * (BX + 4) - 16) + NT - 1) / NT;  // bound on per-plane trips
        T v[NC][IT];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const T* base = pt.in[stage_comp<NS>(d, c)];
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int r = tid + k * NT;
                const int yy = r / RW, xx = r - yy * RW;
                const int gi = gx0 + xx, gj = gy0 + yy;
                v[c][k] = (T)0;
                if (r < n && (INT || (gi >= dlo0 && gi <= dhi0 && gj >= dlo1 && gj <= dhi1 && gi <= tl.x1 + hx &&
                                      gj <= tl.y1 + hy)))
                    v[c][k] = gld(base, (unsigned)((gi - ulo0) + (gj - ulo1) * unx));
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int r = tid + k * NT;
                if (r < n) sm[c * n + r] = v[c][k];
            }
    };

    // ---------------- stage 1: x predictor ----------------
    {
        constexpr int SW = BX + 4;
        stage(0, SW, BY, tl.x0 - 3, tl.y0 - 1, 3, 1);
        __syncthreads();
        if (r1) {
            T w[2][NS][2][5];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < NS; ++k)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const int c = s * 2 * NS + k * 2 + q;
#pragma unroll
                        for (int o = 0; o < 5; ++o) w[s][k][q][o] = sm[(c * BY + ly) * SW + lx + o];
                    }
            if (INT) normal_pred<T, NS, PK>(P, m, pred, (T)1, w, 0, -(1 << 20), 1 << 20, dtbydx, WmX, WpX);
            else normal_pred<T, NS, PK>(P, m, pred, (T)1, w, i, dlo0, dhi0, dtbydx, WmX, WpX);
        }
        __syncthreads();
    }
    // ---------------- stage 2: y predictor ----------------
    {
        constexpr int SH = BY + 4;
        stage(1, BX, SH, tl.x0 - 1, tl.y0 - 3, 1, 3);
        __syncthreads();
        if (r1) {
            T w[2][NS][2][5];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < NS; ++k)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const int c = s * 2 * NS + k * 2 + q;
#pragma unroll
                        for (int o = 0; o < 5; ++o) w[s][k][q][o] = sm[(c * SH + ly + o) * BX + lx];
                    }
            if (INT) normal_pred<T, NS, PK>(P, m, pred, rv1, w, 0, -(1 << 20), 1 << 20, dtbydx, WmY, WpY);
            else normal_pred<T, NS, PK>(P, m, pred, rv1, w, j, dlo1, dhi1, dtbydx, WmY, WpY);
        }
        __syncthreads();
    }

    // exchange helpers: region r of the LDS holds NV values per thread
    auto put = [&](int slot, T v) { sm[slot * NT + tid] = v; };
    auto get = [&](int slot, int t) -> T { return sm[slot * NT + t]; };

    // ---------------- stage 3: predictor face states ----------------
    Slots<T, NS> H0, H1;
    double tmp_state = 0.0;  // RS: psi of the half step; SW: slip-total of the half step
    {
        T cpx[2][NS], cpy[2][NS];
        if (r1) {
            right_chars<T, NS>(m, (T)1, WpX, cpx);
            right_chars<T, NS>(m, rv1, WpY, cpy);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    put(s * NS + k, cpx[s][k]);
                    put(2 * NS + s * NS + k, cpy[s][k]);
                }
        }
        __syncthreads();
        // x faces: lo face of cell i, rows of the ring-grown tile
        if (lx >= 1 && lx <= tnx + 1 && rowD && ly <= tny + 1 && (INT || (i >= dlo0 && i <= dhi0 + 1))) {
            T wpL[2][NS], wmR[2][NS];
            if (!INT && i == dlo0) {
                left_chars<T, NS>(m, (T)1, WmX, wmR);
                bc_state<T, NS>(P, m, 0, (T)1, P.bc[0], -1, wmR, H0);
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(s * NS + k, tid - 1);
                if (!INT && i == dhi0 + 1) {
                    bc_state<T, NS>(P, m, 0, (T)1, P.bc[1], +1, wpL, H0);
                } else {
                    left_chars<T, NS>(m, (T)1, WmX, wmR);
                    riemann_state<T, NS>(m, (T)1, wpL, wmR, H0);
                }
            }
        }
        // y faces: lo face of cell j
        if (ly >= 1 && ly <= tny + 1 && colD && lx <= tnx + 1 && (INT || (j >= dlo1 && j <= dhi1 + 1))) {
            T wpL[2][NS], wmR[2][NS];
            if (!INT && j == dlo1) {
                // fault face (RSIBC::primBC lo-y branch, RSIBC.cpp:220-246)
                if (!pt.bin) {
                    report(errp, TMK_E_NO_BDRY);
                } else {
                    FaultIn fin;
                    fault_input<T, NS>(P, WmY, fin);
                    FaultOut fo;
                    const size_t bi = (size_t)(i - pt.blo0);
                    const size_t bp = (size_t)pt.bnx;
                    if (P.fault_kind == 0) {
                        rs_fault(P, i, gmem(pt.bin)[bi + 10 * bp], fin, fo);
                        // RSSETBND into the half-step copy (RSIBC.cpp:325-333): only psi is consumed
                        const double Vx = fo.vx[0] - fo.vx[1];
                        const double Vz = NS == 3 ? fo.vz[0] - fo.vz[1] : 0.0;
                        const double V = sqrt(Vx * Vx + Vz * Vz);
                        tmp_state = rs_psi_update(P, i, gmem(pt.bin)[bi + 10 * bp], V, 0.5 * dt);
                        if (isnan(tmp_state)) report(errp, TMK_E_PSI_NAN);
                    } else {
                        sw_fault(P, i, gmem(pt.bin)[bi + 7 * bp], fin, fo);
                        const double Vx = fo.vx[0] - fo.vx[1];
                        const double Vz = NS == 3 ? fo.vz[0] - fo.vz[1] : 0.0;
                        const double V = sqrt(Vx * Vx + Vz * Vz);
                        tmp_state = gmem(pt.bin)[bi + 7 * bp] + (0.5 * dt) * V;
                    }
                    fault_output<T, NS>(P, fo, H1);
                }
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(2 * NS + s * NS + k, tid - BX);
                if (!INT && j == dhi1 + 1) {
                    bc_state<T, NS>(P, m, 1, rv1, P.bc[3], +1, wpL, H1);
                } else {
                    left_chars<T, NS>(m, rv1, WmY, wmR);
                    riemann_state<T, NS>(m, rv1, wpL, wmR, H1);
                }
            }
        }
        __syncthreads();
    }
    // ---------------- stage 4: transverse corrections ----------------
    {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                put(s * NS + k, H0.v[s][k]);
                put(2 * NS + s * NS + k, H0.s[s][k]);
                put(4 * NS + s * NS + k, H1.v[s][k]);
                put(6 * NS + s * NS + k, H1.s[s][k]);
            }
        __syncthreads();
        const T scale = K.scale;
        // x predictor (cells of ccBox[1]: ring columns, valid rows) corrected with WHalf1[1]
        if (colD && lx <= tnx + 1 && ly >= 1 && ly <= tny) {
            Slots<T, NS> Hhi;
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    Hhi.v[s][k] = get(4 * NS + s * NS + k, tid + BX);
                    Hhi.s[s][k] = get(6 * NS + s * NS + k, tid + BX);
                }
            transverse_add<T, NS>(m, scale, rv1 * scale, H1, Hhi, WmX);
            transverse_add<T, NS>(m, scale, rv1 * scale, H1, Hhi, WpX);
        }
        // y predictor (cells of ccBox[0]: valid columns, ring rows) corrected with WHalf1[0]
        if (rowD && ly <= tny + 1 && lx >= 1 && lx <= tnx) {
            Slots<T, NS> Hhi;
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    Hhi.v[s][k] = get(s * NS + k, tid + 1);
                    Hhi.s[s][k] = get(2 * NS + s * NS + k, tid + 1);
                }
            transverse_add<T, NS>(m, scale, scale, H0, Hhi, WmY);
            transverse_add<T, NS>(m, scale, scale, H0, Hhi, WpY);
        }
        __syncthreads();
    }
    // ---------------- stage 5: final face states ----------------
    Slots<T, NS> F0, F1;
    {
        T cpx[2][NS], cpy[2][NS];
        right_chars<T, NS>(m, (T)1, WpX, cpx);
        right_chars<T, NS>(m, rv1, WpY, cpy);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                put(s * NS + k, cpx[s][k]);
                put(2 * NS + s * NS + k, cpy[s][k]);
            }
        __syncthreads();
        if (lx >= 1 && lx <= tnx + 1 && ly >= 1 && ly <= tny && (INT || (i >= dlo0 && i <= dhi0 + 1))) {
            T wpL[2][NS], wmR[2][NS];
            if (!INT && i == dlo0) {
                left_chars<T, NS>(m, (T)1, WmX, wmR);
                bc_state<T, NS>(P, m, 0, (T)1, P.bc[0], -1, wmR, F0);
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(s * NS + k, tid - 1);
                if (!INT && i == dhi0 + 1) {
                    bc_state<T, NS>(P, m, 0, (T)1, P.bc[1], +1, wpL, F0);
                } else {
                    left_chars<T, NS>(m, (T)1, WmX, wmR);
                    riemann_state<T, NS>(m, (T)1, wpL, wmR, F0);
                }
            }
        }
        if (ly >= 1 && ly <= tny + 1 && lx >= 1 && lx <= tnx && (INT || (j >= dlo1 && j <= dhi1 + 1))) {
            T wpL[2][NS], wmR[2][NS];
            if (!INT && j == dlo1) {
                if (pt.bin) {
                    FaultIn fin;
                    fault_input<T, NS>(P, WmY, fin);
                    FaultOut fo;
                    const size_t bi = (size_t)(i - pt.blo0);
                    const size_t bp = (size_t)pt.bnx;
                    const __attribute__((address_space(1))) double* B = gmem(pt.bin);
                    __attribute__((address_space(1))) double* Bo = gmem(pt.bout);
                    const double tnew = time + dt;
                    if (P.fault_kind == 0) {
                        rs_fault(P, i, tmp_state, fin, fo);
                        // RSSETBND final (RSIBCF.ChF:61-151) on m_bdryData
                        const double x = (i + 0.5) * P.dx - P.x0;
                        double tau_xy_0 = P.nsig * gaussianfunc(P.r0, x);
                        if (P.f_tau) tau_xy_0 = tau_xy_0 + face_val(P.f_tau, P, i, 0.0);
                        const double Vx = fo.vx[0] - fo.vx[1];
                        const double Vz = NS == 3 ? fo.vz[0] - fo.vz[1] : 0.0;
                        const double V = sqrt(Vx * Vx + Vz * Vz);
                        const double dX = B[bi + 6 * bp], dZ = B[bi + 7 * bp], dT = B[bi + 8 * bp];
                        const double RT = B[bi + 9 * bp], psi = B[bi + 10 * bp];
                        Bo[bi + 0 * bp] = Vx;
                        Bo[bi + 1 * bp] = Vz;
                        Bo[bi + 2 * bp] = V;
                        Bo[bi + 3 * bp] = fo.sxy[0] + tau_xy_0;
                        Bo[bi + 4 * bp] = (NS == 3 ? fo.syz[0] : 0.0) + 0.0;
                        Bo[bi + 6 * bp] = dX + dt * Vx;
                        Bo[bi + 7 * bp] = dZ + dt * Vz;
                        Bo[bi + 8 * bp] = dT + dt * V;
                        if (RT > tnew && V > P.rs_thr) Bo[bi + 9 * bp] = tnew;
                        else if (RT > tnew) Bo[bi + 9 * bp] = 1e10;
                        else Bo[bi + 9 * bp] = RT;
                        Bo[bi + 5 * bp] = fo.syy[0];
                        const double np = rs_psi_update(P, i, psi, V, dt);
                        if (isnan(np)) report(errp, TMK_E_PSI_NAN);
                        Bo[bi + 10 * bp] = np;
                    } else {
                        sw_fault(P, i, tmp_state, fin, fo);
                        // SWSETBND final (SWIBCF.ChF:32-72) on m_bdryData (Bout aliases Bin)
                        const double Vx = fo.vx[0] - fo.vx[1];
                        const double Vz = NS == 3 ? fo.vz[0] - fo.vz[1] : 0.0;
                        const double b5 = B[bi + 5 * bp], b6 = B[bi + 6 * bp], b7 = B[bi + 7 * bp];
                        const double b8 = B[bi + 8 * bp];
                        const double V = sqrt(Vx * Vx + Vz * Vz);
                        Bo[bi + 0 * bp] = Vx;
                        Bo[bi + 1 * bp] = Vz;
                        Bo[bi + 2 * bp] = fo.sxy[0];
                        Bo[bi + 3 * bp] = NS == 3 ? fo.syz[0] : 0.0;
                        Bo[bi + 4 * bp] = V;
                        Bo[bi + 5 * bp] = b5 + dt * Vx;
                        Bo[bi + 6 * bp] = b6 + dt * Vz;
                        Bo[bi + 7 * bp] = b7 + dt * V;
                        Bo[bi + 8 * bp] = (b8 > tnew && V > P.sw_thr) ? tnew : b8;
                    }
                    fault_output<T, NS>(P, fo, F1);
                }
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(2 * NS + s * NS + k, tid - BX);
                if (!INT && j == dhi1 + 1) {
                    bc_state<T, NS>(P, m, 1, rv1, P.bc[3], +1, wpL, F1);
                } else {
                    left_chars<T, NS>(m, rv1, WmY, wmR);
                    riemann_state<T, NS>(m, rv1, wpL, wmR, F1);
                }
            }
        }
        __syncthreads();
    }
    // ---------------- stage 6: flux difference and update ----------------
    {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                put(s * NS + k, F0.v[s][k]);
                put(2 * NS + s * NS + k, F0.s[s][k]);
                put(4 * NS + s * NS + k, F1.v[s][k]);
                put(6 * NS + s * NS + k, F1.s[s][k]);
            }
        __syncthreads();
        // optional flux output for flux registers (LinElastPatchGodunov.cpp:117-118, LinElastPhysics.cpp:153)
        if (pt.F[0] && lx >= 1 && lx <= tnx + 1 && ly >= 1 && ly <= tny) store_flux<T, NS>(pt, m, 0, F0, (T)1, i, j);
        if (pt.F[1] && ly >= 1 && ly <= tny + 1 && lx >= 1 && lx <= tnx) store_flux<T, NS>(pt, m, 1, F1, rv1, i, j);
        if (valid) {
            Slots<T, NS> F0h, F1h;
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    F0h.v[s][k] = get(s * NS + k, tid + 1);
                    F0h.s[s][k] = get(2 * NS + s * NS + k, tid + 1);
                    F1h.v[s][k] = get(4 * NS + s * NS + k, tid + BX);
                    F1h.s[s][k] = get(6 * NS + s * NS + k, tid + BX);
                }
            update_cell<T, NS>(a, pt, P, m, dt, K.sc, i, j, F0, F0h, F1, F1h);
        }
    }
}

// a tile is interior when its stencil (+-3 cells) lies inside the domain (and it fits the
// workgroup): then no cell of it is a domain-boundary or fault cell
template <int BX, int BY>
__device__ __forceinline__ bool tile_interior(const Tile& t, const DevParams& P)
{
    return t.x1 - t.x0 + 1 <= BX - 2 && t.y1 - t.y0 + 1 <= BY - 2 && t.x0 - 3 >= P.dlo[0] && t.x1 + 3 <= P.dhi[0] &&
           t.y0 - 3 >= P.dlo[1] && t.y1 + 3 <= P.dhi[1];
}

// Boundary pieces: the bands of cells within 3 of the domain's edges, whose stencils need the
// boundary conditions; every band is at most 3 cells thick.  A piece along a row band (at most
// BND_BX - 2 = 49 wide, 3 high) runs on a BND_BX x BND_BY workgroup, a piece along a column band
// (taller than 3) on the transposed BND_BY x BND_BX one: either way the 255 lanes cover 3 rows or
// columns of cells and the ring around them, with no idle row.
constexpr int BND_BX = 51, BND_BY = 5;

template <typename T, int NS>
constexpr size_t bnd_lds_elems()
{
    constexpr size_t NC = 4 * NS;
    constexpr size_t s1 = NC * BND_BY * (BND_BX + 4), s2 = NC * (BND_BY + 4) * BND_BX;
    constexpr size_t s1t = NC * BND_BX * (BND_BY + 4), s2t = NC * (BND_BX + 4) * BND_BY;
    constexpr size_t s3 = (size_t)8 * NS * BND_BX * BND_BY;  // the exchange slots (either shape)
    constexpr size_t a = s1 > s2 ? s1 : s2, b = s1t > s2t ? s1t : s2t;
    return (a > b ? a : b) > s3 ? (a > b ? a : b) : s3;
}

// FP32: held to 128 VGPRs, 4 wavefronts per SIMD (a spill of about 20 VGPRs per lane); measured
// 3-4 % faster in batches of serving scenarios than 145 VGPRs at 3 per SIMD
template <typename T, int NS, int PK>
__global__ void __launch_bounds__(BND_BX * BND_BY) __attribute__((amdgpu_waves_per_eu(sizeof(T) == sizeof(float) ? 4 : 1)))
k_godunov_bnd(const KArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    const Tile tl = a.tiles[blockIdx.x];
    if (tl.y1 - tl.y0 + 1 > BND_BY - 2) godunov_tile<T, NS, BND_BY, BND_BX, PK, false>(a, tl, sm);
    else godunov_tile<T, NS, BND_BX, BND_BY, PK, false>(a, tl, sm);
}

// LinElastPatchGodunov::plasticUpdate on its own (PLASTICCORRECTION over the valid box)
template <typename T>
__global__ void k_plastic(const DevParams P, T* U, int ulo0, int ulo1, int unx, size_t plane, int vlo0, int vlo1,
                          int vnx, int vny, double dt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= vnx * vny) return;
    const int i = vlo0 + t % vnx, j = vlo1 + t / vnx;
    const size_t u = (size_t)(i - ulo0) + (size_t)(j - ulo1) * unx;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int o = 11 * s;
        T sxx = U[u + (o + 3) * plane], syy = U[u + (o + 4) * plane], szz = U[u + (o + 5) * plane];
        T sxy = U[u + (o + 6) * plane], sxz = U[u + (o + 7) * plane], syz = U[u + (o + 8) * plane];
        T lam = U[u + (o + 9) * plane], gam = U[u + (o + 10) * plane];
        if constexpr (stores_delta<T>())
            plastic_flow_delta<T, DevParams>(P, (T)dt, sxx, syy, szz, sxy, sxz, syz, lam, gam);
        else
            plastic_flow<T>(P, (T)dt, sxx, syy, szz, sxy, sxz, syz, lam, gam);
        U[u + (o + 3) * plane] = sxx; U[u + (o + 4) * plane] = syy; U[u + (o + 5) * plane] = szz;
        U[u + (o + 6) * plane] = sxy; U[u + (o + 7) * plane] = sxz; U[u + (o + 8) * plane] = syz;
        U[u + (o + 9) * plane] = lam; U[u + (o + 10) * plane] = gam;
    }
}

// same-level ghost exchange (LevelData::exchange, LinElastLevelGodunov.cpp:222-223):
// copy rectangles of valid cells of a source box into the ghost cells of a destination box
struct CopyRect {
    int src, dst;      // patch indices
    int lo0, lo1, hi0, hi1;
};

// one workgroup per rectangle (most are a few ghost rows of a small box)
template <typename T>
__global__ void k_exchange(const Patch<T>* patches, const CopyRect* rects, int nrects, int ncomp_mask)
{
    const CopyRect r = rects[blockIdx.x];
    const Patch<T>& S = patches[r.src];
    const Patch<T>& D = patches[r.dst];
    const int nx = r.hi0 - r.lo0 + 1, ny = r.hi1 - r.lo1 + 1;
    for (int t = threadIdx.x; t < nx * ny; t += blockDim.x) {
        const int i = r.lo0 + t % nx, j = r.lo1 + t / nx;
        const size_t si = (size_t)(i - S.ulo0) + (size_t)(j - S.ulo1) * S.unx;
        const size_t di = (size_t)(i - D.ulo0) + (size_t)(j - D.ulo1) * D.unx;
        for (int c = 0; c < 22; ++c)
            if ((ncomp_mask >> c) & 1) gmem(D.out[c])[di] = gmem(S.out[c])[si];
    }
}

// fault-data ghost exchange along x (m_B.exchange, LinElastLevelGodunov.cpp:223)
template <typename T>
__global__ void k_bexchange(const Patch<T>* patches, const CopyRect* rects, int nrects)
{
    const CopyRect r = rects[blockIdx.x];
    const Patch<T>& S = patches[r.src];
    const Patch<T>& D = patches[r.dst];
    const int nx = r.hi0 - r.lo0 + 1;
    const int nb = D.bnb;
    for (int t = threadIdx.x; t < nx * nb; t += blockDim.x) {
        const int i = r.lo0 + t % nx, c = t / nx;
        gmem(D.bout)[(size_t)(i - D.blo0) + (size_t)c * D.bnx] = gmem(S.bout)[(size_t)(i - S.blo0) + (size_t)c * S.bnx];
    }
}

// fault stations of a level (AMRLevelLinElast::writeStationLevel, AMRLevelLinElast.cpp:1758-1810)
struct StationDev {
    int n, nb, capacity;
    const int* patch;   // [n] patch whose fault row holds face i
    const int* face;    // [n] face index i = floor(x/dx - 0.5)
    const double* d;    // [n] x - (i + 0.5) dx, the reference's weight (metres)
    double* rec;        // [capacity][n][1 + nb]: time, then the interpolated fault data
    int* count;         // [0] records written since the last read, [1] set when one was dropped
};

// one record per station after a level step: the new time and B[i] (1 - d) + B[i+1] d per
// fault-data component, from the step's output fault data (ghosts exchanged)
template <typename T>
__global__ void k_station_record(const Patch<T>* patches, const Scen* scen, StationDev s)
{
    __shared__ int base;
    if (threadIdx.x == 0) base = s.count[0];
    __syncthreads();
    const int b = base;
    if (b >= s.capacity) {
        if (threadIdx.x == 0) s.count[1] = 1;
        return;
    }
    for (int k = threadIdx.x; k < s.n; k += blockDim.x) {
        const Patch<T>& p = patches[s.patch[k]];
        const size_t i = (size_t)(s.face[k] - p.blo0), bp = (size_t)p.bnx;
        const double d = s.d[k];
        double* r = s.rec + ((size_t)b * s.n + k) * (size_t)(1 + s.nb);
        r[0] = scen[p.scen].time;
        for (int c = 0; c < s.nb; ++c) r[1 + c] = gmem(p.bout)[i + c * bp] * (1.0 - d) + gmem(p.bout)[i + 1 + c * bp] * d;
    }
    __syncthreads();
    if (threadIdx.x == 0) s.count[0] = b + 1;
}

// refinement tags (AMRLevelLinElast::tagCells, AMRLevelLinElast.cpp:577-717; RSIBC::tagCellsInit
// RSIBC.cpp:341-377; SWIBC::tagCellsInit SWIBC.cpp:336-396) on the current state of a level
struct TagArgs {
    double refine_thresh, plastic_thresh;
    double xc0, xc1, yc0, yc1, slope, width;
    double x2c0, x2c1, y2c0, y2c1;
    int level, c2max;
    double stress_thresh, sliprate_thresh;
    int nx, ny;  // domain extent (lo = P.dlo)
};

template <typename T>
__device__ __forceinline__ double undiv(const Patch<T>& p, int c, int i, int j, int d, int dlo, int dhi)
{
    const int io = d == 0, jo = d == 1;
    const int ci = d == 0 ? i : j;
    const __attribute__((address_space(1))) T* q = gmem(p.in[c]);
    const size_t u = (size_t)(i - p.ulo0) + (size_t)(j - p.ulo1) * p.unx;
    const size_t s = d == 0 ? 1 : (size_t)p.unx;
    (void)io; (void)jo;
    if (ci == dlo) return fabs((double)q[u + s] - (double)q[u]);
    if (ci == dhi) return fabs((double)q[u] - (double)q[u - s]);
    return 0.5 * fabs((double)q[u + s] - (double)q[u - s]);
}

template <typename T>
__global__ void k_tag(const Patch<T>* patches, const Scen* scen, TagArgs ta, unsigned char* raw)
{
    const Patch<T>& p = patches[blockIdx.x];  // box index on x (no 65535 limit), cell chunks on y
    const DevParams& P = scen[p.scen].P;
    const int vnx = p.vhi0 - p.vlo0 + 1, vny = p.vhi1 - p.vlo1 + 1;
    for (int t = blockIdx.y * blockDim.x + threadIdx.x; t < vnx * vny; t += gridDim.y * blockDim.x) {
        const int i = p.vlo0 + t % vnx, j = p.vlo1 + t / vnx;
        // GETFULLGRADF over the side-1 velocities, then MAXF over directions
        double gm = 0.0;
        for (int d = 0; d < 2; ++d) {
            double g = 0.0;
            for (int c = 0; c < 3; ++c) g = fmax(g, undiv(p, c, i, j, d, P.dlo[d], P.dhi[d]));
            gm = fmax(gm, fabs(g));
        }
        // IBC override on the nucleation cells of the fault row
        if (j == P.dlo[1]) {
            if (P.fault_kind == 0) {
                const int i0 = (int)floor(P.x0 / P.dx), i1 = (int)ceil(P.x0 / P.dx);
                if (i >= i0 && i <= i1) gm = 1.0;
            } else {
                for (int k = 0; k < P.npatch; ++k) {
                    const int i0 = (int)floor((P.fric_c0 + P.xc[k] - P.xw[k]) / P.dx);
                    const int i1 = (int)ceil((P.fric_c0 + P.xc[k] + P.xw[k]) / P.dx);
                    if (i >= i0 && i <= i1) gm = 1.0;
                }
            }
        }
        double r = 0.0;
        const double x = i * P.dx;
        r = fmax(r, fmax(ta.xc0 - x, x - ta.xc1));
        const double y = j * P.dx;
        r = fmax(r, fmax(ta.yc0 - y, y - ta.yc1));
        const double thr = ta.refine_thresh * (1 + r * ta.slope / ta.width);
        unsigned char tag = 0;
        if (ta.level < ta.c2max || (x >= ta.x2c0 && x <= ta.x2c1 && y >= ta.y2c0 && y <= ta.y2c1)) {
            const size_t u = (size_t)(i - p.ulo0) + (size_t)(j - p.ulo1) * p.unx;
            tag = gm >= thr || (double)gmem(p.in[10])[u] >= ta.plastic_thresh ||
                  (double)gmem(p.in[21])[u] >= ta.plastic_thresh;
            if (!tag && ta.stress_thresh > 0.0) {
                double gs = 0.0;
                for (int d = 0; d < 2; ++d)
                    for (int c : {3, 4, 6}) gs = fmax(gs, undiv(p, c, i, j, d, P.dlo[d], P.dhi[d]));
                tag = gs >= ta.stress_thresh;
            }
            if (!tag && ta.sliprate_thresh > 0.0 && j == P.dlo[1] && p.bin) {
                const int vc = P.fault_kind == 0 ? 2 : 4;
                tag = gmem(p.bin)[(size_t)(i - p.blo0) + (size_t)vc * p.bnx] >= ta.sliprate_thresh;
            }
        }
        raw[(size_t)(i - P.dlo[0]) + (size_t)(j - P.dlo[1]) * ta.nx] = tag;
    }
}

// IntVectSet::grow(b) followed by the clip to the domain
static __global__ void k_dilate(const unsigned char* raw, unsigned char* out, int nx, int ny, int b)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int i = t % nx, j = t / nx;
    unsigned char v = 0;
    for (int jj = max(0, j - b); jj <= min(ny - 1, j + b) && !v; ++jj)
        for (int ii = max(0, i - b); ii <= min(nx - 1, i + b) && !v; ++ii) v = raw[(size_t)ii + (size_t)jj * nx];
    out[t] = v;
}

// every scenario's clock advances by its own step (AMRLevelLinElast.cpp:478, m_time += m_dt)
static __global__ void k_advance_time(Scen* s, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) s[k].time = s[k].time + s[k].dt;
}

}  // namespace tmk

