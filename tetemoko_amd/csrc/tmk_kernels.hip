// tmk_kernels.hip -- fused per-tile Godunov update for gfx950 (CDNA4, wave64).
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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "tetemoko_hip.h"
#include "tmk_physics.hpp"

namespace tmk {

// one box of the level as seen by the kernel
template <typename T>
struct Patch {
    const T* in[22];   // state planes read this step
    T* out[22];        // state planes written this step (== in for pointwise planes)
    int ulo0, ulo1, unx; // U array box (ghosted) lo and x extent
    const double* bin; // fault data read (NULL: box has none)
    double* bout;      // fault data written
    int blo0, bnx;     // fault-data x lo and extent
    double* F[2];      // optional face fluxes (22 comps, on surroundingNodes(box,d)) or NULL
    int vlo0, vlo1, vhi0, vhi1; // valid box
};

struct Tile {
    int patch;
    int x0, y0, x1, y1;  // valid cells of this tile (inclusive)
};

template <typename T>
struct KArgs {
    DevParams P;
    const Patch<T>* patches;
    const Tile* tiles;
    int ntiles;
    double time, dt;   // current time and step
    const double* time_dev;  // if non-null, time is read from device memory
    int* err;          // error word (first error wins)
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

template <typename T, int NS, int BX, int BY>
__global__ void __launch_bounds__(BX * BY)
k_godunov(const KArgs<T> a)
{
    constexpr int NT = BX * BY;
    constexpr int NC = 4 * NS;  // staged components per direction
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);

    const DevParams& P = a.P;
    const Mat<T> m(P);
    const Tile tl = a.tiles[blockIdx.x];
    const Patch<T>& pt = a.patches[tl.patch];
    const int tid = threadIdx.x;
    const int lx = tid % BX, ly = tid / BX;
    const int i = tl.x0 - 1 + lx, j = tl.y0 - 1 + ly;
    const int tnx = tl.x1 - tl.x0 + 1, tny = tl.y1 - tl.y0 + 1;
    const int dlo0 = P.dlo[0], dlo1 = P.dlo[1], dhi0 = P.dhi[0], dhi1 = P.dhi[1];
    const bool colD = i >= dlo0 && i <= dhi0;
    const bool rowD = j >= dlo1 && j <= dhi1;
    const bool r1 = colD && rowD && lx <= tnx + 1 && ly <= tny + 1;
    const bool valid = lx >= 1 && lx <= tnx && ly >= 1 && ly <= tny;
    const double time = a.time_dev ? *a.time_dev : a.time;
    const double dt = a.dt;
    const T dtbydx = (T)(dt / P.dx);
    const int pred = P.pred;
    const T rv1 = (T)-1;

    Slots<T, NS> WmX, WpX, WmY, WpY;

    // ---------------- stage 1: x predictor ----------------
    {
        constexpr int SW = BX + 4;
        for (int e = tid; e < NC * BY * SW; e += NT) {
            const int c = e / (BY * SW);
            const int r = e - c * (BY * SW);
            const int yy = r / SW, xx = r - yy * SW;
            const int gi = tl.x0 - 3 + xx, gj = tl.y0 - 1 + yy;
            T v = (T)0;
            // only the tile's stencil region (x: tile +-3, y: tile +-1), inside the domain
            if (gi >= dlo0 && gi <= dhi0 && gj >= dlo1 && gj <= dhi1 && gi <= tl.x1 + 3 && gj <= tl.y1 + 1)
                v = pt.in[stage_comp<NS>(0, c)][uidx(pt, gi, gj, 0)];
            sm[e] = v;
        }
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
            normal_pred<T, NS>(P, m, pred, (T)1, w, i, dlo0, dhi0, dtbydx, WmX, WpX);
        }
        __syncthreads();
    }
    // ---------------- stage 2: y predictor ----------------
    {
        constexpr int SH = BY + 4;
        for (int e = tid; e < NC * SH * BX; e += NT) {
            const int c = e / (SH * BX);
            const int r = e - c * (SH * BX);
            const int yy = r / BX, xx = r - yy * BX;
            const int gi = tl.x0 - 1 + xx, gj = tl.y0 - 3 + yy;
            T v = (T)0;
            // only the tile's stencil region (x: tile +-1, y: tile +-3), inside the domain
            if (gi >= dlo0 && gi <= dhi0 && gj >= dlo1 && gj <= dhi1 && gi <= tl.x1 + 1 && gj <= tl.y1 + 3)
                v = pt.in[stage_comp<NS>(1, c)][uidx(pt, gi, gj, 0)];
            sm[e] = v;
        }
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
            normal_pred<T, NS>(P, m, pred, rv1, w, j, dlo1, dhi1, dtbydx, WmY, WpY);
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
        if (lx >= 1 && lx <= tnx + 1 && rowD && ly <= tny + 1 && i >= dlo0 && i <= dhi0 + 1) {
            T wpL[2][NS], wmR[2][NS];
            if (i == dlo0) {
                left_chars<T, NS>(m, (T)1, WmX, wmR);
                bc_state<T, NS>(P, m, 0, (T)1, P.bc[0], -1, wmR, H0);
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(s * NS + k, tid - 1);
                if (i == dhi0 + 1) {
                    bc_state<T, NS>(P, m, 0, (T)1, P.bc[1], +1, wpL, H0);
                } else {
                    left_chars<T, NS>(m, (T)1, WmX, wmR);
                    riemann_state<T, NS>(m, (T)1, wpL, wmR, H0);
                }
            }
        }
        // y faces: lo face of cell j
        if (ly >= 1 && ly <= tny + 1 && colD && lx <= tnx + 1 && j >= dlo1 && j <= dhi1 + 1) {
            T wpL[2][NS], wmR[2][NS];
            if (j == dlo1) {
                // fault face (RSIBC::primBC lo-y branch, RSIBC.cpp:220-246)
                if (!pt.bin) {
                    report(a.err, TMK_E_NO_BDRY);
                } else {
                    FaultIn fin;
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        fin.vy[s] = (double)WmY.v[s][0]; fin.syy[s] = (double)WmY.s[s][0];
                        fin.vx[s] = (double)WmY.v[s][1]; fin.sxy[s] = (double)WmY.s[s][1];
                        fin.vz[s] = NS == 3 ? (double)WmY.v[s][NS - 1] : 0.0;
                        fin.syz[s] = NS == 3 ? (double)WmY.s[s][NS - 1] : 0.0;
                    }
                    FaultOut fo;
                    const size_t bi = (size_t)(i - pt.blo0);
                    const size_t bp = (size_t)pt.bnx;
                    if (P.fault_kind == 0) {
                        rs_fault(P, i, pt.bin[bi + 10 * bp], fin, fo);
                        // RSSETBND into the half-step copy (RSIBC.cpp:325-333): only psi is consumed
                        const double Vx = fo.vx[0] - fo.vx[1];
                        const double Vz = NS == 3 ? fo.vz[0] - fo.vz[1] : 0.0;
                        const double V = sqrt(Vx * Vx + Vz * Vz);
                        tmp_state = rs_psi_update(P, i, pt.bin[bi + 10 * bp], V, 0.5 * dt);
                        if (isnan(tmp_state)) report(a.err, TMK_E_PSI_NAN);
                    } else {
                        sw_fault(P, i, pt.bin[bi + 7 * bp], fin, fo);
                        const double Vx = fo.vx[0] - fo.vx[1];
                        const double Vz = NS == 3 ? fo.vz[0] - fo.vz[1] : 0.0;
                        const double V = sqrt(Vx * Vx + Vz * Vz);
                        tmp_state = pt.bin[bi + 7 * bp] + (0.5 * dt) * V;
                    }
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        H1.v[s][0] = (T)fo.vy[s]; H1.s[s][0] = (T)fo.syy[s];
                        H1.v[s][1] = (T)fo.vx[s]; H1.s[s][1] = (T)fo.sxy[s];
                        if (NS == 3) { H1.v[s][NS - 1] = (T)fo.vz[s]; H1.s[s][NS - 1] = (T)fo.syz[s]; }
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(2 * NS + s * NS + k, tid - BX);
                if (j == dhi1 + 1) {
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
        const T scale = (T)(-(1.0 / 2.0) * dt / P.dx);
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
        if (lx >= 1 && lx <= tnx + 1 && ly >= 1 && ly <= tny && i >= dlo0 && i <= dhi0 + 1) {
            T wpL[2][NS], wmR[2][NS];
            if (i == dlo0) {
                left_chars<T, NS>(m, (T)1, WmX, wmR);
                bc_state<T, NS>(P, m, 0, (T)1, P.bc[0], -1, wmR, F0);
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(s * NS + k, tid - 1);
                if (i == dhi0 + 1) {
                    bc_state<T, NS>(P, m, 0, (T)1, P.bc[1], +1, wpL, F0);
                } else {
                    left_chars<T, NS>(m, (T)1, WmX, wmR);
                    riemann_state<T, NS>(m, (T)1, wpL, wmR, F0);
                }
            }
        }
        if (ly >= 1 && ly <= tny + 1 && lx >= 1 && lx <= tnx && j >= dlo1 && j <= dhi1 + 1) {
            T wpL[2][NS], wmR[2][NS];
            if (j == dlo1) {
                if (pt.bin) {
                    FaultIn fin;
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        fin.vy[s] = (double)WmY.v[s][0]; fin.syy[s] = (double)WmY.s[s][0];
                        fin.vx[s] = (double)WmY.v[s][1]; fin.sxy[s] = (double)WmY.s[s][1];
                        fin.vz[s] = NS == 3 ? (double)WmY.v[s][NS - 1] : 0.0;
                        fin.syz[s] = NS == 3 ? (double)WmY.s[s][NS - 1] : 0.0;
                    }
                    FaultOut fo;
                    const size_t bi = (size_t)(i - pt.blo0);
                    const size_t bp = (size_t)pt.bnx;
                    const double* B = pt.bin;
                    double* Bo = pt.bout;
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
                        if (isnan(np)) report(a.err, TMK_E_PSI_NAN);
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
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        F1.v[s][0] = (T)fo.vy[s]; F1.s[s][0] = (T)fo.syy[s];
                        F1.v[s][1] = (T)fo.vx[s]; F1.s[s][1] = (T)fo.sxy[s];
                        if (NS == 3) { F1.v[s][NS - 1] = (T)fo.vz[s]; F1.s[s][NS - 1] = (T)fo.syz[s]; }
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int k = 0; k < NS; ++k) wpL[s][k] = get(2 * NS + s * NS + k, tid - BX);
                if (j == dhi1 + 1) {
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
        // GETFLUXF (LinElastPhysicsF.ChF:303-350) from a face state, in slot form:
        // f[s][0]=F(vn) f[s][1]=F(snn) f[s][2]=F(vt) f[s][3]=F(snt) f[s][4]=F(s_transverse normal)
        // f[s][5]=F(va) f[s][6]=F(sna)
        auto flux = [&](const Slots<T, NS>& S, T rv, T (&f)[2][7]) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const T r = s ? rv : (T)1;
                if (s == 0) {
                    f[s][0] = -(S.s[s][0] / m.rho);
                    f[s][2] = -(S.s[s][1] / m.rho);
                    f[s][1] = -(S.v[s][0] * m.lam2G);
                    f[s][4] = -(S.v[s][0] * m.lam);
                    f[s][3] = -(S.v[s][1] * m.G);
                    if (NS == 3) { f[s][5] = -(S.s[s][NS - 1] / m.rho); f[s][6] = -(S.v[s][NS - 1] * m.G); }
                } else {
                    f[s][0] = -(r * S.s[s][0] / m.rho);
                    f[s][2] = -(r * S.s[s][1] / m.rho);
                    f[s][1] = -(r * S.v[s][0] * m.lam2G);
                    f[s][4] = -(r * S.v[s][0] * m.lam);
                    f[s][3] = -(r * S.v[s][1] * m.G);
                    if (NS == 3) { f[s][5] = -(r * S.s[s][NS - 1] / m.rho); f[s][6] = -(r * S.v[s][NS - 1] * m.G); }
                }
            }
        };
        // optional flux output for flux registers (LinElastPatchGodunov.cpp:117-118, LinElastPhysics.cpp:153)
        if (pt.F[0] && lx >= 1 && lx <= tnx + 1 && ly >= 1 && ly <= tny) {
            T f[2][7];
            flux(F0, (T)1, f);
            double* Fo = pt.F[0];
            const int fnx = pt.vhi0 - pt.vlo0 + 2, fny = pt.vhi1 - pt.vlo1 + 1;
            const size_t fp = (size_t)fnx * fny;
            const size_t fi = (size_t)(i - pt.vlo0) + (size_t)(j - pt.vlo1) * fnx;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int o = 11 * s;
                Fo[fi + (o + 0) * fp] = (double)f[s][0];
                Fo[fi + (o + 1) * fp] = (double)f[s][2];
                Fo[fi + (o + 2) * fp] = NS == 3 ? (double)f[s][5] : -(s ? 1.0 : 1.0) * 0.0 * 0.0;
                Fo[fi + (o + 3) * fp] = (double)f[s][1];
                Fo[fi + (o + 4) * fp] = (double)f[s][4];
                Fo[fi + (o + 5) * fp] = (double)f[s][4];
                Fo[fi + (o + 6) * fp] = (double)f[s][3];
                Fo[fi + (o + 7) * fp] = NS == 3 ? (double)f[s][6] : 0.0;
                Fo[fi + (o + 8) * fp] = 0.0;
                Fo[fi + (o + 9) * fp] = 0.0;
                Fo[fi + (o + 10) * fp] = 0.0;
            }
        }
        if (pt.F[1] && ly >= 1 && ly <= tny + 1 && lx >= 1 && lx <= tnx) {
            T f[2][7];
            flux(F1, rv1, f);
            double* Fo = pt.F[1];
            const int fnx = pt.vhi0 - pt.vlo0 + 1, fny = pt.vhi1 - pt.vlo1 + 2;
            const size_t fp = (size_t)fnx * fny;
            const size_t fi = (size_t)(i - pt.vlo0) + (size_t)(j - pt.vlo1) * fnx;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int o = 11 * s;
                Fo[fi + (o + 1) * fp] = (double)f[s][0];
                Fo[fi + (o + 0) * fp] = (double)f[s][2];
                Fo[fi + (o + 2) * fp] = NS == 3 ? (double)f[s][5] : 0.0;
                Fo[fi + (o + 4) * fp] = (double)f[s][1];
                Fo[fi + (o + 5) * fp] = (double)f[s][4];
                Fo[fi + (o + 3) * fp] = (double)f[s][4];
                Fo[fi + (o + 6) * fp] = (double)f[s][3];
                Fo[fi + (o + 8) * fp] = NS == 3 ? (double)f[s][6] : 0.0;
                Fo[fi + (o + 7) * fp] = 0.0;
                Fo[fi + (o + 9) * fp] = 0.0;
                Fo[fi + (o + 10) * fp] = 0.0;
            }
        }
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
            T fxl[2][7], fxh[2][7], fyl[2][7], fyh[2][7];
            flux(F0, (T)1, fxl);
            flux(F0h, (T)1, fxh);
            flux(F1, rv1, fyl);
            flux(F1h, rv1, fyh);
            const T sc = (T)(-dt / P.dx);
            const size_t u = uidx(pt, i, j, 0);
            // dU = (0 + (Fx_hi - Fx_lo)) + (Fy_hi - Fy_lo); U += dU * (-dt/dx)
            auto upd = [&](int c, T dx_, T dy_) {
                const T dU = ((T)0 + dx_) + dy_;
                pt.out[c][u] = pt.in[c][u] + dU * sc;
            };
            T st[2][6];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int o = 11 * s;
                // x flux: vx<-F(vn) vy<-F(vt) sxx<-F(snn) syy,szz<-F(stt) sxy<-F(snt)
                // y flux: vy<-F(vn) vx<-F(vt) syy<-F(snn) sxx,szz<-F(stt) sxy<-F(snt)
                const T dvx_x = fxh[s][0] - fxl[s][0], dvx_y = fyh[s][2] - fyl[s][2];
                const T dvy_x = fxh[s][2] - fxl[s][2], dvy_y = fyh[s][0] - fyl[s][0];
                const T dsxx_x = fxh[s][1] - fxl[s][1], dsxx_y = fyh[s][4] - fyl[s][4];
                const T dsyy_x = fxh[s][4] - fxl[s][4], dsyy_y = fyh[s][1] - fyl[s][1];
                const T dszz_x = fxh[s][4] - fxl[s][4], dszz_y = fyh[s][4] - fyl[s][4];
                const T dsxy_x = fxh[s][3] - fxl[s][3], dsxy_y = fyh[s][3] - fyl[s][3];
                upd(o + 0, dvx_x, dvx_y);
                upd(o + 1, dvy_x, dvy_y);
                const T sxx = pt.in[o + 3][u] + (((T)0 + dsxx_x) + dsxx_y) * sc;
                const T syy = pt.in[o + 4][u] + (((T)0 + dsyy_x) + dsyy_y) * sc;
                const T szz = pt.in[o + 5][u] + (((T)0 + dszz_x) + dszz_y) * sc;
                const T sxy = pt.in[o + 6][u] + (((T)0 + dsxy_x) + dsxy_y) * sc;
                st[s][0] = sxx; st[s][1] = syy; st[s][2] = szz; st[s][3] = sxy;
                if (NS == 3) {
                    const T dvz_x = fxh[s][5] - fxl[s][5], dvz_y = fyh[s][5] - fyl[s][5];
                    const T dsxz_x = fxh[s][6] - fxl[s][6], dsxz_y = (T)0 - (T)0;
                    const T dsyz_x = (T)0 - (T)0, dsyz_y = fyh[s][6] - fyl[s][6];
                    upd(o + 2, dvz_x, dvz_y);
                    st[s][4] = pt.in[o + 7][u] + (((T)0 + dsxz_x) + dsxz_y) * sc;
                    st[s][5] = pt.in[o + 8][u] + (((T)0 + dsyz_x) + dsyz_y) * sc;
                } else {
                    st[s][4] = (T)0;
                    st[s][5] = (T)0;
                }
            }
            if (P.plast) {
                // LinElastLevelGodunov::plasticUpdate after the level step (AMRLevelLinElast.cpp:476)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int o = 11 * s;
                    T lam = pt.in[o + 9][u], gam = pt.in[o + 10][u];
                    T sxz = NS == 3 ? st[s][4] : pt.in[o + 7][u];
                    T syz = NS == 3 ? st[s][5] : pt.in[o + 8][u];
                    plastic_flow<T>(P, (T)dt, st[s][0], st[s][1], st[s][2], st[s][3], sxz, syz, lam, gam);
                    st[s][4] = sxz; st[s][5] = syz;
                    pt.out[o + 9][u] = lam;
                    pt.out[o + 10][u] = gam;
                    if (NS == 2) { pt.out[o + 7][u] = sxz; pt.out[o + 8][u] = syz; }
                }
            }
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int o = 11 * s;
                pt.out[o + 3][u] = st[s][0];
                pt.out[o + 4][u] = st[s][1];
                pt.out[o + 5][u] = st[s][2];
                pt.out[o + 6][u] = st[s][3];
                if (NS == 3) { pt.out[o + 7][u] = st[s][4]; pt.out[o + 8][u] = st[s][5]; }
            }
        }
    }
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

template <typename T>
__global__ void k_exchange(const Patch<T>* patches, const CopyRect* rects, int nrects, int ncomp_mask)
{
    const CopyRect r = rects[blockIdx.y];
    const Patch<T>& S = patches[r.src];
    const Patch<T>& D = patches[r.dst];
    const int nx = r.hi0 - r.lo0 + 1, ny = r.hi1 - r.lo1 + 1;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nx * ny; t += gridDim.x * blockDim.x) {
        const int i = r.lo0 + t % nx, j = r.lo1 + t / nx;
        const size_t si = (size_t)(i - S.ulo0) + (size_t)(j - S.ulo1) * S.unx;
        const size_t di = (size_t)(i - D.ulo0) + (size_t)(j - D.ulo1) * D.unx;
        for (int c = 0; c < 22; ++c)
            if ((ncomp_mask >> c) & 1) D.out[c][di] = S.out[c][si];
    }
}

// fault-data ghost exchange along x (m_B.exchange, LinElastLevelGodunov.cpp:223)
template <typename T>
__global__ void k_bexchange(const Patch<T>* patches, const CopyRect* rects, int nrects, int nb)
{
    const CopyRect r = rects[blockIdx.y];
    const Patch<T>& S = patches[r.src];
    const Patch<T>& D = patches[r.dst];
    const int nx = r.hi0 - r.lo0 + 1;
    for (int t = threadIdx.x; t < nx * nb; t += blockDim.x) {
        const int i = r.lo0 + t % nx, c = t / nx;
        D.bout[(size_t)(i - D.blo0) + (size_t)c * D.bnx] = S.bout[(size_t)(i - S.blo0) + (size_t)c * S.bnx];
    }
}

__global__ void k_advance_time(double* t, double dt) { if (threadIdx.x == 0 && blockIdx.x == 0) t[0] = t[0] + dt; }

}  // namespace tmk

// ============================================================================
// host side
// ============================================================================
using namespace tmk;

#define HIPCHK(x)                                                      \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) {                                        \
            fprintf(stderr, "tetemoko_hip: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return e_ == hipErrorOutOfMemory ? TMK_E_NOMEM : TMK_E_HIP; \
        }                                                              \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= n) return TMK_OK;
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return TMK_E_NOMEM;
        n = bytes;
        return TMK_OK;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

struct tmk_ctx {
    tmk_params hp;
    DevParams P;
    int precision;
    hipStream_t stream = nullptr;
    std::vector<double*> face_dev;
    DevBuf scratch[8];
    int* err_dev = nullptr;
    int bx = 32, by = 16;
    bool antiplane = true;  // per-patch path keeps the anti-plane fields unless they are provably zero
};

struct tmk_level {
    tmk_ctx* ctx;
    std::vector<tmk_box> boxes;
    std::vector<int> has_b;
    int ns;                      // 2: anti-plane planes skipped, 3: carried
    size_t tsize;                // sizeof bulk type
    std::vector<void*> ubuf[2];  // per box: stencil planes A/B
    std::vector<void*> upt;      // per box: pointwise planes
    std::vector<double*> bbuf[2];
    int cur = 0;
    void* patches_dev[2] = {nullptr, nullptr};  // Patch arrays for cur=0 and cur=1
    Tile* tiles_dev = nullptr;
    int ntiles = 0;
    CopyRect* rects_dev = nullptr;
    int nrects = 0;
    CopyRect* brects_dev = nullptr;
    int nbrects = 0;
    double* time_dev = nullptr;
    hipGraphExec_t gexec = nullptr;
    double gdt = 0.0;
    bool dirty = true;  // uploaded data not yet exchanged
};

static const char* k_errs[] = {"ok",
                               "NaN state variable psi (dPsi is nan)",
                               "invalid boundary type",
                               "invalid normal predictor or limiter combination",
                               "fault face without fault data",
                               "array box does not cover the update stencil",
                               "invalid argument",
                               "HIP runtime failure",
                               "unsupported option",
                               "device memory exhausted"};

extern "C" const char* tmk_strerror(int code)
{
    const int k = -code;
    if (k < 0 || k >= (int)(sizeof(k_errs) / sizeof(k_errs[0]))) return "unknown error";
    return k_errs[k];
}

extern "C" int tmk_device_ok(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return strstr(prop.gcnArchName, "gfx950") != nullptr ? 1 : 0;
}

static void fill_dev_params(const tmk_params& h, DevParams& P)
{
    memset(&P, 0, sizeof P);
    // LINELASTSETF (LinElastPhysicsF.ChF:1165-1207)
    P.cs = h.cs; P.cp = h.cp; P.G = h.G;
    P.rho = h.G / (h.cs * h.cs);
    P.lam = (h.cp * h.cp) * P.rho - 2.0 * h.G;
    P.Zs = P.rho * h.cs;
    P.Zp = P.rho * h.cp;
    P.K = P.lam + 2.0 * h.G / 3.0;
    P.den = 2.0 * h.G + P.lam;
    P.lam2G = P.lam + 2.0 * h.G;
    const double s9[9] = {0.0, 0.0, 0.0, h.sxx0, h.syy0, h.szz0, h.sxy0, h.sxz0, h.syz0};
    for (int s = 0; s < 2; ++s) {
        for (int c = 0; c < 9; ++c) P.bg[11 * s + c] = s9[c];
        P.bg[11 * s + 9] = 0.0;
        P.bg[11 * s + 10] = 0.0;
    }
    P.plast = h.use_plasticity;
    P.p_mu = h.p_mu; P.p_beta = h.p_beta; P.p_eta = h.p_eta; P.p_b = h.p_b; P.p_h = h.p_h;
    for (int k = 0; k < 4; ++k) P.bc[k] = h.bc[k];
    P.fault_kind = h.fault_kind;
    P.r0 = h.r0; P.nsig = h.nsig; P.rs_a = h.rs_a; P.rs_b = h.rs_b; P.rs_V0 = h.rs_V0; P.rs_f0 = h.rs_f0;
    P.rs_L = h.rs_L; P.rs_fw = h.rs_fw; P.rs_Vw = h.rs_Vw; P.rs_fexp = h.rs_fexp;
    P.x0 = h.x0; P.rs_thr = h.rs_rupture_thr; P.ramp_x = h.ramp_x; P.ramp_w = h.ramp_w; P.ramp_a = h.ramp_a;
    P.ramp_Vw = h.ramp_Vw;
    P.state_law = h.state_law;
    P.fs = h.fs; P.fd = h.fd; P.dw = h.dw; P.sm = h.sm_value; P.sw_thr = h.sw_rupture_thr;
    P.npatch = h.num_patches;
    for (int k = 0; k < TMK_MAXPATCH; ++k) { P.xc[k] = h.xc[k]; P.xw[k] = h.xw[k]; P.taup[k] = h.tau_p[k]; }
    P.fric_c0 = h.fric_center[0]; P.fric_w0 = h.fric_width[0]; P.outside_f = h.outside_f;
    P.face_lo = h.face_lo; P.nface = h.nface;
    P.pred = h.normal_pred; P.fourth = h.fourth_order; P.prim_lim = h.prim_limiting; P.char_lim = h.char_limiting;
    P.dx = h.dx;
    P.dlo[0] = h.domain.lo[0]; P.dlo[1] = h.domain.lo[1];
    P.dhi[0] = h.domain.hi[0]; P.dhi[1] = h.domain.hi[1];
}

extern "C" int tmk_ctx_create(tmk_ctx** out, const tmk_params* p, int precision)
{
    if (!out || !p) return TMK_E_ARG;
    *out = nullptr;
    if (precision != TMK_PREC_FP64 && precision != TMK_PREC_FP32_BULK) return TMK_E_ARG;
    if (p->normal_pred < 0 || p->normal_pred > 2) return TMK_E_PRED_ORDER;
    if (p->normal_pred != 0 && p->char_limiting && p->prim_limiting) return TMK_E_PRED_ORDER;
    for (int k = 0; k < 4; ++k)
        if (p->bc[k] != 0 && p->bc[k] != 1 && k != 2) return TMK_E_BAD_BC;
    if (p->fault_kind != TMK_FAULT_RS && p->fault_kind != TMK_FAULT_SW) return TMK_E_ARG;
    if (p->num_patches < 0 || p->num_patches > TMK_MAXPATCH) return TMK_E_ARG;
    if (p->dx <= 0.0 || p->domain.hi[0] - p->domain.lo[0] < 3 || p->domain.hi[1] - p->domain.lo[1] < 3)
        return TMK_E_ARG;
    tmk_ctx* c = new (std::nothrow) tmk_ctx;
    if (!c) return TMK_E_NOMEM;
    c->hp = *p;
    c->precision = precision;
    fill_dev_params(*p, c->P);
    c->antiplane = !(p->sxz0 == 0.0 && p->syz0 == 0.0);
    if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return TMK_E_HIP; }
    if (hipMalloc(&c->err_dev, sizeof(int)) != hipSuccess) { tmk_ctx_destroy(c); return TMK_E_NOMEM; }
    (void)hipMemset(c->err_dev, 0, sizeof(int));
    const double* fa[8] = {p->f_a, p->f_b, p->f_L, p->f_fs, p->f_fd, p->f_dc, p->f_syy, p->f_tau};
    const double** fd[8] = {&c->P.f_a, &c->P.f_b, &c->P.f_L, &c->P.f_fs, &c->P.f_fd, &c->P.f_dc, &c->P.f_syy,
                            &c->P.f_tau};
    for (int k = 0; k < 8; ++k) {
        *fd[k] = nullptr;
        if (fa[k]) {
            if (p->nface <= 0) { tmk_ctx_destroy(c); return TMK_E_ARG; }
            double* d = nullptr;
            if (hipMalloc(&d, sizeof(double) * p->nface) != hipSuccess) { tmk_ctx_destroy(c); return TMK_E_NOMEM; }
            c->face_dev.push_back(d);
            if (hipMemcpy(d, fa[k], sizeof(double) * p->nface, hipMemcpyHostToDevice) != hipSuccess) {
                tmk_ctx_destroy(c);
                return TMK_E_HIP;
            }
            *fd[k] = d;
        }
    }
    c->hp.f_a = c->hp.f_b = c->hp.f_L = c->hp.f_fs = c->hp.f_fd = c->hp.f_dc = c->hp.f_syy = c->hp.f_tau = nullptr;
    *out = c;
    return TMK_OK;
}

extern "C" int tmk_ctx_destroy(tmk_ctx* c)
{
    if (!c) return TMK_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (double* d : c->face_dev) (void)hipFree(d);
    for (auto& b : c->scratch) b.release();
    if (c->err_dev) (void)hipFree(c->err_dev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TMK_OK;
}

extern "C" int tmk_set_tile(tmk_ctx* c, int bx, int by)
{
    if (!c) return TMK_E_ARG;
    if (bx == 0) bx = 32;
    if (by == 0) by = 16;
    const bool ok = (bx == 32 && (by == 8 || by == 16 || by == 32)) || (bx == 64 && (by == 4 || by == 8 || by == 16));
    if (!ok) return TMK_E_ARG;
    c->bx = bx; c->by = by;
    return TMK_OK;
}

static inline bool box_contains(const tmk_box& o, const tmk_box& b)
{
    return b.lo[0] >= o.lo[0] && b.lo[1] >= o.lo[1] && b.hi[0] <= o.hi[0] && b.hi[1] <= o.hi[1];
}
static inline tmk_box box_and(tmk_box a, const tmk_box& b)
{
    for (int d = 0; d < 2; ++d) { a.lo[d] = std::max(a.lo[d], b.lo[d]); a.hi[d] = std::min(a.hi[d], b.hi[d]); }
    return a;
}
static inline tmk_box box_grow(tmk_box a, int n)
{
    a.lo[0] -= n; a.lo[1] -= n; a.hi[0] += n; a.hi[1] += n;
    return a;
}

static void make_tiles(const tmk_box& v, int patch, int tx, int ty, std::vector<Tile>& out)
{
    for (int y0 = v.lo[1]; y0 <= v.hi[1]; y0 += ty)
        for (int x0 = v.lo[0]; x0 <= v.hi[0]; x0 += tx) {
            Tile t;
            t.patch = patch;
            t.x0 = x0; t.y0 = y0;
            t.x1 = std::min(x0 + tx - 1, v.hi[0]);
            t.y1 = std::min(y0 + ty - 1, v.hi[1]);
            out.push_back(t);
        }
}

// Order tiles so that each XCD (round-robin over workgroup ids) walks a
// contiguous band of tiles and keeps the shared halos in its own L2.
static void xcd_order(std::vector<Tile>& t)
{
    const size_t nx = 8;
    const size_t n = t.size();
    if (n < 64 || n % nx != 0) return;
    const size_t per = n / nx;
    std::vector<Tile> o(n);
    for (size_t g = 0; g < n; ++g) o[g] = t[(g % nx) * per + g / nx];
    t.swap(o);
}

template <typename T, int NS, int BX, int BY>
static int launch_tiles_t(hipStream_t st, const KArgs<T>& a)
{
    constexpr int NT = BX * BY;
    constexpr int NC = 4 * NS;
    const size_t s1 = (size_t)NC * BY * (BX + 4), s2 = (size_t)NC * (BY + 4) * BX, s3 = (size_t)8 * NS * NT;
    const size_t shm = std::max(s1, std::max(s2, s3)) * sizeof(T);
    if (a.ntiles == 0) return TMK_OK;
    if (shm > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_godunov<T, NS, BX, BY>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm);
    hipLaunchKernelGGL((k_godunov<T, NS, BX, BY>), dim3(a.ntiles), dim3(NT), shm, st, a);
    return hipGetLastError() == hipSuccess ? TMK_OK : TMK_E_HIP;
}

template <typename T, int NS>
static int launch_tiles_ns(hipStream_t st, int bx, int by, const KArgs<T>& a)
{
    if (bx == 32 && by == 16) return launch_tiles_t<T, NS, 32, 16>(st, a);
    if (bx == 32 && by == 8) return launch_tiles_t<T, NS, 32, 8>(st, a);
    if (bx == 32 && by == 32) return launch_tiles_t<T, NS, 32, 32>(st, a);
    if (bx == 64 && by == 8) return launch_tiles_t<T, NS, 64, 8>(st, a);
    if (bx == 64 && by == 4) return launch_tiles_t<T, NS, 64, 4>(st, a);
    if (bx == 64 && by == 16) return launch_tiles_t<T, NS, 64, 16>(st, a);
    return TMK_E_ARG;
}

template <typename T>
static int launch_tiles(hipStream_t st, int ns, int bx, int by, const KArgs<T>& a)
{
    return ns == 3 ? launch_tiles_ns<T, 3>(st, bx, by, a) : launch_tiles_ns<T, 2>(st, bx, by, a);
}

// ---------------------------------------------------------------------------
// per-patch drop-in (host FArrayBox data)
// ---------------------------------------------------------------------------
template <typename T>
static int patch_update_t(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, double* F0, double* F1, double* B,
                          tmk_box bb, int nb, double time, double dt)
{
    const size_t unx = (size_t)(ub.hi[0] - ub.lo[0] + 1), uny = (size_t)(ub.hi[1] - ub.lo[1] + 1);
    const size_t plane = unx * uny;
    const size_t nU = plane * TMK_NCOMP;
    const size_t bnx = B ? (size_t)(bb.hi[0] - bb.lo[0] + 1) : 0;
    const size_t nB = bnx * (size_t)nb;
    const size_t nF0 = F0 ? (size_t)(valid.hi[0] - valid.lo[0] + 2) * (valid.hi[1] - valid.lo[1] + 1) * TMK_NCOMP : 0;
    const size_t nF1 = F1 ? (size_t)(valid.hi[0] - valid.lo[0] + 1) * (valid.hi[1] - valid.lo[1] + 2) * TMK_NCOMP : 0;
    int e;
    if ((e = c->scratch[0].ensure(nU * sizeof(double)))) return e;
    if ((e = c->scratch[1].ensure(nU * sizeof(T)))) return e;
    if ((e = c->scratch[2].ensure(nU * sizeof(T)))) return e;
    if ((e = c->scratch[3].ensure(std::max<size_t>(2 * nB, 1) * sizeof(double)))) return e;
    if ((e = c->scratch[4].ensure(std::max<size_t>(nF0 + nF1, 1) * sizeof(double)))) return e;
    if ((e = c->scratch[5].ensure(sizeof(Patch<T>) + 2 * sizeof(Tile) * 4096))) return e;
    hipStream_t st = c->stream;
    double* Ud = (double*)c->scratch[0].p;
    T* Uin = (T*)c->scratch[1].p;
    T* Uout = (T*)c->scratch[2].p;
    double* Bd = (double*)c->scratch[3].p;
    double* Fd = (double*)c->scratch[4].p;
    HIPCHK(hipMemcpyAsync(Ud, U, nU * sizeof(double), hipMemcpyHostToDevice, st));
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(hipMemcpyAsync(Uin, Ud, nU * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
        // round the host FP64 state to the FP32 bulk type
        std::vector<float> tmp(nU);
        for (size_t k = 0; k < nU; ++k) tmp[k] = (float)U[k];
        HIPCHK(hipMemcpyAsync(Uin, tmp.data(), nU * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    HIPCHK(hipMemcpyAsync(Uout, Uin, nU * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (B) {
        HIPCHK(hipMemcpyAsync(Bd, B, nB * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(Bd + nB, Bd, nB * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    Patch<T> pt;
    memset(&pt, 0, sizeof pt);
    for (int k = 0; k < TMK_NCOMP; ++k) { pt.in[k] = Uin + k * plane; pt.out[k] = Uout + k * plane; }
    pt.ulo0 = ub.lo[0]; pt.ulo1 = ub.lo[1]; pt.unx = (int)unx;
    pt.bin = B ? Bd : nullptr;
    pt.bout = B ? Bd + nB : nullptr;
    pt.blo0 = bb.lo[0]; pt.bnx = (int)bnx;
    pt.F[0] = F0 ? Fd : nullptr;
    pt.F[1] = F1 ? Fd + nF0 : nullptr;
    pt.vlo0 = valid.lo[0]; pt.vlo1 = valid.lo[1]; pt.vhi0 = valid.hi[0]; pt.vhi1 = valid.hi[1];
    std::vector<Tile> tiles;
    make_tiles(valid, 0, c->bx - 2, c->by - 2, tiles);
    if ((e = c->scratch[5].ensure(sizeof(Patch<T>) + sizeof(Tile) * tiles.size() + 64))) return e;
    char* meta = (char*)c->scratch[5].p;
    Patch<T>* pd = (Patch<T>*)meta;
    Tile* td = (Tile*)(meta + ((sizeof(Patch<T>) + 63) / 64) * 64);
    HIPCHK(hipMemcpyAsync(pd, &pt, sizeof pt, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(td, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(c->err_dev, 0, sizeof(int), st));
    KArgs<T> a;
    a.P = c->P;
    a.patches = pd;
    a.tiles = td;
    a.ntiles = (int)tiles.size();
    a.time = time;
    a.dt = dt;
    a.time_dev = nullptr;
    a.err = c->err_dev;
    // the per-patch path cannot prove the anti-plane fields zero, unless background and data say so
    int ns = 3;
    if (!c->antiplane) {
        bool zero = true;
        for (size_t jj = 0; jj < uny && zero; ++jj)
            for (size_t ii = 0; ii < unx && zero; ++ii)
                for (int s = 0; s < 2 && zero; ++s)
                    for (int cc : {2, 7, 8})
                        if (U[ii + jj * unx + (size_t)(11 * s + cc) * plane] != 0.0) { zero = false; break; }
        ns = zero ? 2 : 3;
    }
    const int plast = a.P.plast;
    a.P.plast = 0;  // updateState does not include the plastic return
    if ((e = launch_tiles<T>(st, ns, c->bx, c->by, a))) return e;
    a.P.plast = plast;
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, c->err_dev, sizeof(int), hipMemcpyDeviceToHost, st));
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(hipMemcpyAsync(U, Uout, nU * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    } else {
        std::vector<float> tmp(nU);
        HIPCHK(hipMemcpyAsync(tmp.data(), Uout, nU * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (size_t k = 0; k < nU; ++k) U[k] = (double)tmp[k];
    }
    if (B) HIPCHK(hipMemcpy(B, Bd + nB, nB * sizeof(double), hipMemcpyDeviceToHost));
    if (F0) HIPCHK(hipMemcpy(F0, Fd, nF0 * sizeof(double), hipMemcpyDeviceToHost));
    if (F1) HIPCHK(hipMemcpy(F1, Fd + nF0, nF1 * sizeof(double), hipMemcpyDeviceToHost));
    return herr;
}

extern "C" int tmk_patch_update_state(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, int ncomp, double* F0,
                                      double* F1, double* B, tmk_box bb, int nbdry, double time, double dt,
                                      double* max_wave_speed)
{
    if (!c || !U || ncomp != TMK_NCOMP) return TMK_E_ARG;
    if (valid.hi[0] < valid.lo[0] || valid.hi[1] < valid.lo[1]) return TMK_E_ARG;
    const tmk_box dom = c->hp.domain;
    if (!box_contains(dom, valid)) return TMK_E_BOX;
    // the update reads W on grow(valid,3) & domain and writes valid cells
    if (!box_contains(ub, box_and(box_grow(valid, 3), dom))) return TMK_E_BOX;
    const bool fault = valid.lo[1] == dom.lo[1];
    if (fault) {
        if (!B) return TMK_E_NO_BDRY;
        const int need = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        if (nbdry != need) return TMK_E_ARG;
        tmk_box need_b = valid;
        need_b.lo[0] = std::max(valid.lo[0] - 1, dom.lo[0]);
        need_b.hi[0] = std::min(valid.hi[0] + 1, dom.hi[0]);
        if (bb.lo[0] > need_b.lo[0] || bb.hi[0] < need_b.hi[0]) return TMK_E_BOX;
    } else {
        B = nullptr;
    }
    if (max_wave_speed) *max_wave_speed = std::max(c->hp.cs, c->hp.cp);
    if (c->precision == TMK_PREC_FP64) return patch_update_t<double>(c, valid, U, ub, F0, F1, B, bb, nbdry, time, dt);
    return patch_update_t<float>(c, valid, U, ub, F0, F1, B, bb, nbdry, time, dt);
}

extern "C" int tmk_patch_plastic_update(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, int ncomp, double dt)
{
    if (!c || !U || ncomp != TMK_NCOMP) return TMK_E_ARG;
    if (!box_contains(ub, valid)) return TMK_E_BOX;
    const size_t unx = (size_t)(ub.hi[0] - ub.lo[0] + 1), uny = (size_t)(ub.hi[1] - ub.lo[1] + 1);
    const size_t plane = unx * uny, nU = plane * TMK_NCOMP;
    int e;
    if ((e = c->scratch[0].ensure(nU * sizeof(double)))) return e;
    double* Ud = (double*)c->scratch[0].p;
    HIPCHK(hipMemcpyAsync(Ud, U, nU * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const int vnx = valid.hi[0] - valid.lo[0] + 1, vny = valid.hi[1] - valid.lo[1] + 1;
    const int n = vnx * vny;
    hipLaunchKernelGGL(k_plastic<double>, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->P, Ud, ub.lo[0],
                       ub.lo[1], (int)unx, plane, valid.lo[0], valid.lo[1], vnx, vny, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(U, Ud, nU * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TMK_OK;
}

// ---------------------------------------------------------------------------
// device-resident level
// ---------------------------------------------------------------------------
static const int kStencil2[] = {0, 1, 3, 4, 6};            // vx vy sxx syy sxy
static const int kStencil3[] = {0, 1, 2, 3, 4, 6, 7, 8};   // + vz sxz syz

static bool is_stencil(int c, int ns)
{
    const int k = c % 11;
    if (ns == 2) { for (int x : kStencil2) if (x == k) return true; return false; }
    for (int x : kStencil3) if (x == k) return true;
    return false;
}

template <typename T>
static int level_build_patches(tmk_level* L)
{
    const int nbox = (int)L->boxes.size();
    const int nb = L->ctx->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    for (int cur = 0; cur < 2; ++cur) {
        std::vector<Patch<T>> pts(nbox);
        for (int b = 0; b < nbox; ++b) {
            const tmk_box v = L->boxes[b];
            const tmk_box ub = box_grow(v, TMK_GHOST);
            const size_t unx = ub.hi[0] - ub.lo[0] + 1, uny = ub.hi[1] - ub.lo[1] + 1, plane = unx * uny;
            Patch<T>& p = pts[b];
            memset(&p, 0, sizeof p);
            T* A = (T*)L->ubuf[cur][b];
            T* Bn = (T*)L->ubuf[1 - cur][b];
            T* W = (T*)L->upt[b];
            for (int c = 0; c < TMK_NCOMP; ++c) {
                if (is_stencil(c, L->ns)) { p.in[c] = A + c * plane; p.out[c] = Bn + c * plane; }
                else { p.in[c] = W + c * plane; p.out[c] = W + c * plane; }
            }
            p.ulo0 = ub.lo[0]; p.ulo1 = ub.lo[1]; p.unx = (int)unx;
            if (L->has_b[b]) {
                p.bin = L->bbuf[cur][b];
                p.bout = L->bbuf[1 - cur][b];
                p.blo0 = v.lo[0] - TMK_GHOST;
                p.bnx = v.hi[0] - v.lo[0] + 1 + 2 * TMK_GHOST;
            }
            p.vlo0 = v.lo[0]; p.vlo1 = v.lo[1]; p.vhi0 = v.hi[0]; p.vhi1 = v.hi[1];
            (void)nb;
        }
        if (!L->patches_dev[cur]) HIPCHK(hipMalloc(&L->patches_dev[cur], sizeof(Patch<T>) * nbox));
        HIPCHK(hipMemcpy(L->patches_dev[cur], pts.data(), sizeof(Patch<T>) * nbox, hipMemcpyHostToDevice));
    }
    return TMK_OK;
}

extern "C" int tmk_level_create(tmk_ctx* c, int nboxes, const tmk_box* boxes, tmk_level** out)
{
    if (!c || !out || nboxes <= 0 || !boxes) return TMK_E_ARG;
    *out = nullptr;
    tmk_level* L = new (std::nothrow) tmk_level;
    if (!L) return TMK_E_NOMEM;
    L->ctx = c;
    L->boxes.assign(boxes, boxes + nboxes);
    L->ns = c->antiplane ? 3 : 2;
    L->tsize = c->precision == TMK_PREC_FP64 ? sizeof(double) : sizeof(float);
    const tmk_box dom = c->hp.domain;
    const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    for (int b = 0; b < nboxes; ++b) {
        const tmk_box v = boxes[b];
        if (!box_contains(dom, v) || v.hi[0] < v.lo[0] || v.hi[1] < v.lo[1]) { delete L; return TMK_E_BOX; }
        const tmk_box ub = box_grow(v, TMK_GHOST);
        const size_t plane = (size_t)(ub.hi[0] - ub.lo[0] + 1) * (ub.hi[1] - ub.lo[1] + 1);
        const size_t bytes = plane * TMK_NCOMP * L->tsize;
        void *a = nullptr, *bb = nullptr, *w = nullptr;
        if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&bb, bytes) != hipSuccess ||
            hipMalloc(&w, bytes) != hipSuccess) {
            if (a) (void)hipFree(a);
            if (bb) (void)hipFree(bb);
            if (w) (void)hipFree(w);
            tmk_level_destroy(L);
            return TMK_E_NOMEM;
        }
        L->ubuf[0].push_back(a);
        L->ubuf[1].push_back(bb);
        L->upt.push_back(w);
        const int hb = v.lo[1] == dom.lo[1];
        L->has_b.push_back(hb);
        double *b0 = nullptr, *b1 = nullptr;
        if (hb) {
            const size_t nB = (size_t)(v.hi[0] - v.lo[0] + 1 + 2 * TMK_GHOST) * nb;
            if (hipMalloc(&b0, nB * sizeof(double)) != hipSuccess || hipMalloc(&b1, nB * sizeof(double)) != hipSuccess) {
                L->bbuf[0].push_back(b0);
                L->bbuf[1].push_back(b1);
                tmk_level_destroy(L);
                return TMK_E_NOMEM;
            }
        }
        L->bbuf[0].push_back(b0);
        L->bbuf[1].push_back(b1);
    }
    int e = c->precision == TMK_PREC_FP64 ? level_build_patches<double>(L) : level_build_patches<float>(L);
    if (e) { tmk_level_destroy(L); return e; }
    // tiles
    std::vector<Tile> tiles;
    for (int b = 0; b < nboxes; ++b) make_tiles(boxes[b], b, c->bx - 2, c->by - 2, tiles);
    xcd_order(tiles);
    L->ntiles = (int)tiles.size();
    if (hipMalloc(&L->tiles_dev, sizeof(Tile) * tiles.size()) != hipSuccess) { tmk_level_destroy(L); return TMK_E_NOMEM; }
    (void)hipMemcpy(L->tiles_dev, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice);
    // exchange plan: ghost cells (within the domain) of each box covered by another box's valid cells
    std::vector<CopyRect> rects, brects;
    for (int d = 0; d < nboxes; ++d) {
        const tmk_box g = box_and(box_grow(boxes[d], TMK_GHOST), dom);
        for (int s = 0; s < nboxes; ++s) {
            if (s == d) continue;
            const tmk_box r = box_and(g, boxes[s]);
            if (r.hi[0] < r.lo[0] || r.hi[1] < r.lo[1]) continue;
            rects.push_back(CopyRect{s, d, r.lo[0], r.lo[1], r.hi[0], r.hi[1]});
            if (L->has_b[d] && L->has_b[s]) {
                const int lo = std::max(boxes[d].lo[0] - TMK_GHOST, boxes[s].lo[0]);
                const int hi = std::min(boxes[d].hi[0] + TMK_GHOST, boxes[s].hi[0]);
                if (lo <= hi) brects.push_back(CopyRect{s, d, lo, 0, hi, 0});
            }
        }
    }
    L->nrects = (int)rects.size();
    L->nbrects = (int)brects.size();
    if (!rects.empty()) {
        (void)hipMalloc(&L->rects_dev, sizeof(CopyRect) * rects.size());
        (void)hipMemcpy(L->rects_dev, rects.data(), sizeof(CopyRect) * rects.size(), hipMemcpyHostToDevice);
    }
    if (!brects.empty()) {
        (void)hipMalloc(&L->brects_dev, sizeof(CopyRect) * brects.size());
        (void)hipMemcpy(L->brects_dev, brects.data(), sizeof(CopyRect) * brects.size(), hipMemcpyHostToDevice);
    }
    if (hipMalloc(&L->time_dev, sizeof(double)) != hipSuccess) { tmk_level_destroy(L); return TMK_E_NOMEM; }
    *out = L;
    return TMK_OK;
}

extern "C" int tmk_level_destroy(tmk_level* L)
{
    if (!L) return TMK_OK;
    if (L->ctx && L->ctx->stream) (void)hipStreamSynchronize(L->ctx->stream);
    if (L->gexec) (void)hipGraphExecDestroy(L->gexec);
    for (int k = 0; k < 2; ++k) {
        for (void* p : L->ubuf[k]) if (p) (void)hipFree(p);
        for (double* p : L->bbuf[k]) if (p) (void)hipFree(p);
        if (L->patches_dev[k]) (void)hipFree(L->patches_dev[k]);
    }
    for (void* p : L->upt) if (p) (void)hipFree(p);
    if (L->tiles_dev) (void)hipFree(L->tiles_dev);
    if (L->rects_dev) (void)hipFree(L->rects_dev);
    if (L->brects_dev) (void)hipFree(L->brects_dev);
    if (L->time_dev) (void)hipFree(L->time_dev);
    delete L;
    return TMK_OK;
}

// host FP64 <-> device bulk type, one plane set per box
extern "C" int tmk_level_upload(tmk_level* L, int b, const double* U, const double* B)
{
    if (!L || b < 0 || b >= (int)L->boxes.size() || !U) return TMK_E_ARG;
    tmk_ctx* c = L->ctx;
    const tmk_box ub = box_grow(L->boxes[b], TMK_GHOST);
    const size_t plane = (size_t)(ub.hi[0] - ub.lo[0] + 1) * (ub.hi[1] - ub.lo[1] + 1);
    const size_t nU = plane * TMK_NCOMP;
    std::vector<float> tmp;
    const void* src = U;
    if (L->tsize == sizeof(float)) {
        tmp.resize(nU);
        for (size_t k = 0; k < nU; ++k) tmp[k] = (float)U[k];
        src = tmp.data();
    }
    // the level skips the anti-plane planes only when they are zero
    if (L->ns == 2) {
        for (int s = 0; s < 2; ++s)
            for (int cc : {2, 7, 8})
                for (size_t k = 0; k < plane; ++k)
                    if (U[k + (size_t)(11 * s + cc) * plane] != 0.0) return TMK_E_ARG;
    }
    const size_t bytes = nU * L->tsize;
    L->dirty = true;
    for (int k = 0; k < 2; ++k) HIPCHK(hipMemcpy(L->ubuf[k][b], src, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L->upt[b], src, bytes, hipMemcpyHostToDevice));
    if (L->has_b[b]) {
        if (!B) return TMK_E_NO_BDRY;
        const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        const size_t nB = (size_t)(L->boxes[b].hi[0] - L->boxes[b].lo[0] + 1 + 2 * TMK_GHOST) * nb;
        for (int k = 0; k < 2; ++k) HIPCHK(hipMemcpy(L->bbuf[k][b], B, nB * sizeof(double), hipMemcpyHostToDevice));
    }
    return TMK_OK;
}

extern "C" int tmk_level_download(tmk_level* L, int b, double* U, double* B)
{
    if (!L || b < 0 || b >= (int)L->boxes.size() || !U) return TMK_E_ARG;
    tmk_ctx* c = L->ctx;
    HIPCHK(hipStreamSynchronize(c->stream));
    const tmk_box ub = box_grow(L->boxes[b], TMK_GHOST);
    const size_t plane = (size_t)(ub.hi[0] - ub.lo[0] + 1) * (ub.hi[1] - ub.lo[1] + 1);
    const size_t nU = plane * TMK_NCOMP;
    std::vector<char> a(nU * L->tsize), w(nU * L->tsize);
    HIPCHK(hipMemcpy(a.data(), L->ubuf[L->cur][b], nU * L->tsize, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(w.data(), L->upt[b], nU * L->tsize, hipMemcpyDeviceToHost));
    for (int cc = 0; cc < TMK_NCOMP; ++cc) {
        const char* src = is_stencil(cc, L->ns) ? a.data() : w.data();
        for (size_t k = 0; k < plane; ++k) {
            const size_t q = k + (size_t)cc * plane;
            U[q] = L->tsize == sizeof(double) ? ((const double*)src)[q] : (double)((const float*)src)[q];
        }
    }
    if (L->has_b[b] && B) {
        const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        const size_t nB = (size_t)(L->boxes[b].hi[0] - L->boxes[b].lo[0] + 1 + 2 * TMK_GHOST) * nb;
        HIPCHK(hipMemcpy(B, L->bbuf[L->cur][b], nB * sizeof(double), hipMemcpyDeviceToHost));
    }
    return TMK_OK;
}

template <typename T>
static int level_enqueue_step(tmk_level* L, double time, double dt, bool device_time)
{
    tmk_ctx* c = L->ctx;
    hipStream_t st = c->stream;
    const Patch<T>* pd = (const Patch<T>*)L->patches_dev[L->cur];
    KArgs<T> a;
    a.P = c->P;
    a.patches = pd;
    a.tiles = L->tiles_dev;
    a.ntiles = L->ntiles;
    a.time = time;
    a.dt = dt;
    a.time_dev = device_time ? L->time_dev : nullptr;
    a.err = c->err_dev;
    int e = launch_tiles<T>(st, L->ns, c->bx, c->by, a);
    if (e) return e;
    L->cur = 1 - L->cur;
    // ghost exchange of the new state into the ghost cells of the other boxes
    if (L->nrects > 0) {
        const Patch<T>* pn = (const Patch<T>*)L->patches_dev[1 - L->cur];  // 'out' of the step just launched
        int mask = 0;
        for (int cc = 0; cc < TMK_NCOMP; ++cc)
            if (is_stencil(cc, L->ns)) mask |= 1 << cc;
        hipLaunchKernelGGL(k_exchange<T>, dim3(4, L->nrects), dim3(256), 0, st, pn, L->rects_dev, L->nrects, mask);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    if (L->nbrects > 0) {
        const Patch<T>* pn = (const Patch<T>*)L->patches_dev[1 - L->cur];
        const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        hipLaunchKernelGGL(k_bexchange<T>, dim3(1, L->nbrects), dim3(256), 0, st, pn, L->brects_dev, L->nbrects, nb);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    if (device_time) {
        hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, st, L->time_dev, dt);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    return TMK_OK;
}

template <typename T>
static int level_exchange_now(tmk_level* L)
{
    tmk_ctx* c = L->ctx;
    // patches_dev[1-cur].out == the current state buffers
    const Patch<T>* pn = (const Patch<T>*)L->patches_dev[1 - L->cur];
    if (L->nrects > 0) {
        int mask = 0;
        for (int cc = 0; cc < TMK_NCOMP; ++cc)
            if (is_stencil(cc, L->ns)) mask |= 1 << cc;
        hipLaunchKernelGGL(k_exchange<T>, dim3(4, L->nrects), dim3(256), 0, c->stream, pn, L->rects_dev, L->nrects, mask);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    if (L->nbrects > 0) {
        const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        hipLaunchKernelGGL(k_bexchange<T>, dim3(1, L->nbrects), dim3(256), 0, c->stream, pn, L->brects_dev, L->nbrects, nb);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    return TMK_OK;
}

static int level_prepare(tmk_level* L)
{
    if (!L->dirty) return TMK_OK;
    L->dirty = false;
    return L->tsize == sizeof(double) ? level_exchange_now<double>(L) : level_exchange_now<float>(L);
}

static int level_step_any(tmk_level* L, double time, double dt, bool device_time)
{
    return L->tsize == sizeof(double) ? level_enqueue_step<double>(L, time, dt, device_time)
                                      : level_enqueue_step<float>(L, time, dt, device_time);
}

extern "C" int tmk_level_step(tmk_level* L, double time, double dt, double* dt_new)
{
    if (!L) return TMK_E_ARG;
    int e = level_prepare(L);
    if (e) return e;
    e = level_step_any(L, time, dt, false);
    if (e) return e;
    if (dt_new) *dt_new = L->ctx->hp.dx / std::max(L->ctx->hp.cs, L->ctx->hp.cp);
    return TMK_OK;
}

extern "C" int tmk_level_run(tmk_level* L, int nsteps, double time, double dt)
{
    if (!L || nsteps < 0) return TMK_E_ARG;
    if (nsteps == 0) return TMK_OK;
    tmk_ctx* c = L->ctx;
    hipStream_t st = c->stream;
    int e = level_prepare(L);
    if (e) return e;
    HIPCHK(hipMemcpyAsync(L->time_dev, &time, sizeof(double), hipMemcpyHostToDevice, st));
    // the graph holds two steps (buffers A->B, B->A) so that it always starts from cur == 0
    if (L->cur != 0) {
        if ((e = level_step_any(L, 0.0, dt, true))) return e;
        if (--nsteps == 0) return TMK_OK;
    }
    if (nsteps >= 2 && (!L->gexec || L->gdt != dt)) {
        if (L->gexec) { (void)hipGraphExecDestroy(L->gexec); L->gexec = nullptr; }
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        e = level_step_any(L, 0.0, dt, true);
        if (!e) e = level_step_any(L, 0.0, dt, true);
        hipError_t ce = hipStreamEndCapture(st, &g);
        if (e) return e;
        if (ce != hipSuccess) return TMK_E_HIP;
        HIPCHK(hipGraphInstantiate(&L->gexec, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        L->gdt = dt;
    }
    for (int k = 0; k + 1 < nsteps; k += 2) HIPCHK(hipGraphLaunch(L->gexec, st));
    if (nsteps % 2) {
        if ((e = level_step_any(L, 0.0, dt, true))) return e;
    }
    return TMK_OK;
}

extern "C" int tmk_level_run_timed(tmk_level* L, int nsteps, double time, double dt, double* ms)
{
    if (!L || !ms) return TMK_E_ARG;
    hipStream_t st = L->ctx->stream;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, st));
    int e = tmk_level_run(L, nsteps, time, dt);
    if (e) { (void)hipEventDestroy(a); (void)hipEventDestroy(b); return e; }
    HIPCHK(hipEventRecord(b, st));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    *ms = f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return TMK_OK;
}

extern "C" int tmk_level_profile(tmk_level* L, int nsteps, double time, double dt, double* ms)
{
    if (!L || !ms || nsteps <= 0) return TMK_E_ARG;
    int e = level_prepare(L);
    if (e) return e;
    hipStream_t st = L->ctx->stream;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, st));
    double t = time;
    for (int k = 0; k < nsteps && !e; ++k) {
        e = level_step_any(L, t, dt, false);
        t = t + dt;
    }
    if (e) { (void)hipEventDestroy(a); (void)hipEventDestroy(b); return e; }
    HIPCHK(hipEventRecord(b, st));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    *ms = (double)f / nsteps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return TMK_OK;
}

extern "C" int tmk_level_sync(tmk_level* L)
{
    if (!L) return TMK_E_ARG;
    int herr = 0;
    HIPCHK(hipStreamSynchronize(L->ctx->stream));
    HIPCHK(hipMemcpy(&herr, L->ctx->err_dev, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) (void)hipMemset(L->ctx->err_dev, 0, sizeof(int));
    return herr;
}
