// tmk_interior.hpp -- the update of an interior tile (stencil inside the domain).
//
// Same arithmetic as godunov_tile (tmk_kernels.hpp) on the same cells, in the same order, so
// results are bit-identical; only the data movement differs:
//   * one staging pass loads the union of both sweeps' stencils, a (BX+4) x (BY+4) box of the
//     ten mode-II stencil planes (vx vy sxx syy sxy, both sides; sixteen with the anti-plane
//     planes) into LDS; the x and y normal predictors then run back to back;
//   * neighbour exchanges along x use DPP wavefront shifts (lane i <- lane i -+ 1).  A tile row
//     is BX lanes with 64 % BX == 0 or BX % 64 == 0, so a shift that crosses a row or wave edge
//     only lands in lane lx == 0 (left neighbour) or lx == BX - 1 (right neighbour), and those
//     lanes never use the shifted value;
//   * neighbour exchanges along y go through LDS, alternating between two regions so that each
//     exchange needs one barrier: region A lies past the staging box, region B reuses the
//     staging box once both predictors are done;
//   * every lane runs every stage.  Values of lanes on the tile edge that the general kernel
//     leaves uncomputed are computed from neighbours that do not exist (shifted-in or
//     out-of-region LDS values) and are never read by a cell that is updated; every LDS
//     index stays inside the workgroup's allocation.
// Stage numbering follows godunov_tile (LinElastPatchGodunov.cpp:101-129, :190-441).
#pragma once

#include "tmk_kernels.hpp"

namespace tmk {

// staged planes per side: vx vy sxx syy sxy (+ vz sxz syz)
template <int NS>
__host__ __device__ constexpr int stage_nq()
{
    return NS == 2 ? 5 : 8;
}

// state component of staged plane q (side = q / stage_nq)
template <int NS>
__device__ __forceinline__ int stage_plane_comp(int q)
{
    constexpr int nq = stage_nq<NS>();
    const int s = q / nq, r = q - s * nq;
    const int c = r == 0 ? 0 : r == 1 ? 1 : r == 2 ? 3 : r == 3 ? 4 : r == 4 ? 6 : r == 5 ? 2 : r == 6 ? 7 : 8;
    return c + 11 * s;
}

// staged plane of (direction d, side s, slot k, velocity/stress)
template <int NS>
__host__ __device__ constexpr int stage_plane(int d, int s, int k, int is_s)
{
    // d 0: N (vx, sxx) S (vy, sxy) A (vz, sxz);  d 1: N (vy, syy) S (vx, sxy) A (vz, syz)
    return s * stage_nq<NS>() +
           (d == 0 ? (k == 0 ? (is_s ? 2 : 0) : k == 1 ? (is_s ? 4 : 1) : (is_s ? 6 : 5))
                   : (k == 0 ? (is_s ? 3 : 1) : k == 1 ? (is_s ? 4 : 0) : (is_s ? 7 : 5)));
}

// LDS elements (of T) the interior kernel uses: staging box plus exchange region A
template <int NS, int BX, int BY>
__host__ __device__ constexpr size_t interior_lds_elems()
{
    return (size_t)2 * stage_nq<NS>() * (BX + 4) * (BY + 4) + (size_t)2 * NS * BX * BY;
}

constexpr int DPP_FROM_PREV = 0x138;  // wave_shr:1, lane i reads lane i - 1
constexpr int DPP_FROM_NEXT = 0x130;  // wave_shl:1, lane i reads lane i + 1

template <int CTRL>
__device__ __forceinline__ double lane_shift(double v)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float lane_shift(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <typename T, int NS, int BX, int BY, int PK>
__device__ __forceinline__ void interior_tile(const KArgs<T>& a, const Tile& tl, T* sm)
{
    static_assert(64 % BX == 0 || BX % 64 == 0, "x lane shifts need whole tile rows per wave");
    constexpr int NT = BX * BY;
    constexpr int SX = BX + 4, SY = BY + 4, SP = SX * SY;
    constexpr int NQ = 2 * stage_nq<NS>();
    T* const stg = sm;              // [NQ][SY][SX], box origin (x0 - 3, y0 - 3)
    T* const exA = sm + NQ * SP;    // [2 NS][NT]
    T* const exB = sm;              // [4 NS][NT], over the staging box

    const Patch<T>& pt = a.patches[tl.patch];
    Scen& S = a.scen[pt.scen];
    const DevParams& P = S.P;
    const Mat<T> m(P);
    const int tid = threadIdx.x;
    const int lx = tid % BX, ly = tid / BX;
    const int i = tl.x0 - 1 + lx, j = tl.y0 - 1 + ly;
    const double dt = S.dt;
    const StepConst<T> K = step_const<T>(S);
    const T dtbydx = K.dtbydx;
    const int pred = P.pred;
    const T rv1 = (T)-1;

    // ---------------- staging: both sweeps' stencils ----------------
    // a partial tile (at a box edge) stages only its own stencil region, tile +-3, which lies in
    // the box's ghosted array; the rest of the box is zero and feeds only lanes past the tile
    const int tnx = tl.x1 - tl.x0 + 1, tny = tl.y1 - tl.y0 + 1;
    const bool full = tnx == BX - 2 && tny == BY - 2;
    {
        const T* src[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) src[q] = pt.in[stage_plane_comp<NS>(q)];
#pragma unroll
        for (int r = 0; r < (SP + NT - 1) / NT; ++r) {
            const int p = tid + r * NT;
            if (p < SP) {
                const int yy = p / SX, xx = p - yy * SX;
                const unsigned g = (unsigned)(tl.x0 - 3 + xx - pt.ulo0) + (unsigned)(tl.y0 - 3 + yy - pt.ulo1) * (unsigned)pt.unx;
                if (full || (xx <= tnx + 5 && yy <= tny + 5)) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) stg[q * SP + p] = gld(src[q], g);
                } else {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) stg[q * SP + p] = (T)0;
                }
            }
        }
    }
    __syncthreads();

    // ---------------- stages 1-2: normal predictors ----------------
    Slots<T, NS> WmX, WpX, WmY, WpY;
    {
        T w[2][NS][2][5];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k)
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int o = 0; o < 5; ++o)
                        w[s][k][q][o] = stg[stage_plane<NS>(0, s, k, q) * SP + (ly + 2) * SX + lx + o];
        normal_pred<T, NS, PK>(P, m, pred, (T)1, w, 0, -(1 << 20), 1 << 20, dtbydx, WmX, WpX);
    }
    {
        T w[2][NS][2][5];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k)
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int o = 0; o < 5; ++o)
                        w[s][k][q][o] = stg[stage_plane<NS>(1, s, k, q) * SP + (ly + o) * SX + lx + 2];
        normal_pred<T, NS, PK>(P, m, pred, rv1, w, 0, -(1 << 20), 1 << 20, dtbydx, WmY, WpY);
    }

    // ---------------- stage 3: predictor face states ----------------
    Slots<T, NS> H0, H1;
    {
        T cpx[2][NS], cpy[2][NS], wpL[2][NS];
        right_chars<T, NS>(m, (T)1, WpX, cpx);
        right_chars<T, NS>(m, rv1, WpY, cpy);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                exA[(s * NS + k) * NT + tid] = cpy[s][k];
                wpL[s][k] = lane_shift<DPP_FROM_PREV>(cpx[s][k]);
            }
        __syncthreads();
        T wmR[2][NS];
        // lo x face of cell i (used for lx >= 1)
        left_chars<T, NS>(m, (T)1, WmX, wmR);
        riemann_state<T, NS>(m, (T)1, wpL, wmR, H0);
        // lo y face of cell j (used for ly >= 1)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) wpL[s][k] = exA[(s * NS + k) * NT + tid - BX];
        left_chars<T, NS>(m, rv1, WmY, wmR);
        riemann_state<T, NS>(m, rv1, wpL, wmR, H1);
    }
    // ---------------- stage 4: transverse corrections ----------------
    {
        Slots<T, NS> Hhi;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                exB[(s * NS + k) * NT + tid] = H1.v[s][k];
                exB[(2 * NS + s * NS + k) * NT + tid] = H1.s[s][k];
                Hhi.v[s][k] = lane_shift<DPP_FROM_NEXT>(H0.v[s][k]);
                Hhi.s[s][k] = lane_shift<DPP_FROM_NEXT>(H0.s[s][k]);
            }
        __syncthreads();
        const T scale = K.scale;
        // y predictor (used on valid columns, ring rows) corrected with the x-face states
        transverse_add<T, NS>(m, scale, scale, H0, Hhi, WmY);
        transverse_add<T, NS>(m, scale, scale, H0, Hhi, WpY);
        // x predictor (used on ring columns, valid rows) corrected with the y-face states
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                Hhi.v[s][k] = exB[(s * NS + k) * NT + tid + BX];
                Hhi.s[s][k] = exB[(2 * NS + s * NS + k) * NT + tid + BX];
            }
        transverse_add<T, NS>(m, scale, rv1 * scale, H1, Hhi, WmX);
        transverse_add<T, NS>(m, scale, rv1 * scale, H1, Hhi, WpX);
    }
    // ---------------- stage 5: final face states ----------------
    Slots<T, NS> F0, F1;
    {
        T cpx[2][NS], cpy[2][NS], wpL[2][NS];
        right_chars<T, NS>(m, (T)1, WpX, cpx);
        right_chars<T, NS>(m, rv1, WpY, cpy);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                exA[(s * NS + k) * NT + tid] = cpy[s][k];
                wpL[s][k] = lane_shift<DPP_FROM_PREV>(cpx[s][k]);
            }
        __syncthreads();
        T wmR[2][NS];
        // lo x face (used for lx >= 1 on valid rows), lo y face (used for ly >= 1 on valid columns)
        left_chars<T, NS>(m, (T)1, WmX, wmR);
        riemann_state<T, NS>(m, (T)1, wpL, wmR, F0);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) wpL[s][k] = exA[(s * NS + k) * NT + tid - BX];
        left_chars<T, NS>(m, rv1, WmY, wmR);
        riemann_state<T, NS>(m, rv1, wpL, wmR, F1);
    }
    // ---------------- stage 6: flux difference and update ----------------
    {
        Slots<T, NS> F0h, F1h;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                exB[(s * NS + k) * NT + tid] = F1.v[s][k];
                exB[(2 * NS + s * NS + k) * NT + tid] = F1.s[s][k];
                F0h.v[s][k] = lane_shift<DPP_FROM_NEXT>(F0.v[s][k]);
                F0h.s[s][k] = lane_shift<DPP_FROM_NEXT>(F0.s[s][k]);
            }
        __syncthreads();
        if (pt.F[0] && lx >= 1 && lx <= tnx + 1 && ly >= 1 && ly <= tny) store_flux<T, NS>(pt, m, 0, F0, (T)1, i, j);
        if (pt.F[1] && ly >= 1 && ly <= tny + 1 && lx >= 1 && lx <= tnx) store_flux<T, NS>(pt, m, 1, F1, rv1, i, j);
        if (lx >= 1 && lx <= tnx && ly >= 1 && ly <= tny) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    F1h.v[s][k] = exB[(s * NS + k) * NT + tid + BX];
                    F1h.s[s][k] = exB[(2 * NS + s * NS + k) * NT + tid + BX];
                }
            update_cell<T, NS>(a, pt, P, m, dt, K.sc, i, j, F0, F0h, F1, F1h);
        }
    }
}

// The same update with one thread per (cell, side): NT = 2 BX BY threads, side = tid / (BX BY).
// Inside an interior tile the two sides of the fault never meet (they are coupled only by the
// fault solve), so each side's thread carries half the state: half the registers per thread and
// twice the waves per workgroup.  Side 2's y-sweep sign rvy = -1 enters every expression as a
// product, which is exact, so both sides round exactly as in interior_tile.  LDS layout and
// barriers are those of interior_tile with a side index in front of every plane.
template <typename T, int NS, int BX, int BY, int PK>
__device__ __forceinline__ void interior_tile_split(const KArgs<T>& a, const Tile& tl, T* sm)
{
    static_assert(64 % BX == 0 || BX % 64 == 0, "x lane shifts need whole tile rows per wave");
    constexpr int NC = BX * BY;  // cells per tile, a multiple of 64: each wave holds one side
    static_assert(NC % 64 == 0, "a wave must not hold both sides");
    constexpr int SX = BX + 4, SY = BY + 4, SP = SX * SY;
    constexpr int NQ1 = stage_nq<NS>();
    T* const stg = sm;                  // [2][NQ1][SY][SX]
    T* const exA = sm + 2 * NQ1 * SP;   // [2][NS][NC]
    T* const exB = sm;                  // [2][2 NS][NC], over the staging box

    const Patch<T>& pt = a.patches[tl.patch];
    Scen& S = a.scen[pt.scen];
    const DevParams& P = S.P;
    const Mat<T> m(P);
    const int tid = threadIdx.x;
    const int sd = __builtin_amdgcn_readfirstlane(tid / NC);  // uniform per wave
    const int c = tid - sd * NC;
    const int lx = c % BX, ly = c / BX;
    const int i = tl.x0 - 1 + lx, j = tl.y0 - 1 + ly;
    const double dt = S.dt;
    const StepConst<T> K = step_const<T>(S);
    const T dtbydx = K.dtbydx;
    const int pred = P.pred;
    const T rv1 = (T)-1;
    const T rvy = sd ? rv1 : (T)1;  // velocity sign of this side in y
    const int q0 = sd * NQ1;        // first staged plane of this side

    // ---------------- staging: this side's stencil planes ----------------
    const int tnx = tl.x1 - tl.x0 + 1, tny = tl.y1 - tl.y0 + 1;
    const bool full = tnx == BX - 2 && tny == BY - 2;
    {
        const T* src[NQ1];
#pragma unroll
        for (int q = 0; q < NQ1; ++q) src[q] = pt.in[stage_plane_comp<NS>(q0 + q)];
#pragma unroll
        for (int r = 0; r < (SP + NC - 1) / NC; ++r) {
            const int p = c + r * NC;
            if (p < SP) {
                const int yy = p / SX, xx = p - yy * SX;
                const unsigned g = (unsigned)(tl.x0 - 3 + xx - pt.ulo0) + (unsigned)(tl.y0 - 3 + yy - pt.ulo1) * (unsigned)pt.unx;
                if (full || (xx <= tnx + 5 && yy <= tny + 5)) {
#pragma unroll
                    for (int q = 0; q < NQ1; ++q) stg[(q0 + q) * SP + p] = gld(src[q], g);
                } else {
#pragma unroll
                    for (int q = 0; q < NQ1; ++q) stg[(q0 + q) * SP + p] = (T)0;
                }
            }
        }
    }
    __syncthreads();

    // ---------------- stages 1-2: normal predictors ----------------
    T WmXv[NS], WmXs[NS], WpXv[NS], WpXs[NS], WmYv[NS], WmYs[NS], WpYv[NS], WpYs[NS];
    {
        T w[NS][2][5];
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int o = 0; o < 5; ++o)
                    w[k][q][o] = stg[(q0 + stage_plane<NS>(0, 0, k, q)) * SP + (ly + 2) * SX + lx + o];
        normal_pred_side<T, NS, PK>(P, m, pred, (T)1, w, 0, -(1 << 20), 1 << 20, dtbydx, WmXv, WmXs, WpXv, WpXs);
    }
    {
        T w[NS][2][5];
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int o = 0; o < 5; ++o)
                    w[k][q][o] = stg[(q0 + stage_plane<NS>(1, 0, k, q)) * SP + (ly + o) * SX + lx + 2];
        normal_pred_side<T, NS, PK>(P, m, pred, rvy, w, 0, -(1 << 20), 1 << 20, dtbydx, WmYv, WmYs, WpYv, WpYs);
    }

    // ---------------- stage 3: predictor face states ----------------
    T H0v[NS], H0s[NS], H1v[NS], H1s[NS];
    {
        T cpx[NS], cpy[NS], wpL[NS], wmR[NS];
        right_chars_side<T, NS>(m, (T)1, WpXv, WpXs, cpx);
        right_chars_side<T, NS>(m, rvy, WpYv, WpYs, cpy);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            exA[(sd * NS + k) * NC + c] = cpy[k];
            wpL[k] = lane_shift<DPP_FROM_PREV>(cpx[k]);
        }
        __syncthreads();
        left_chars_side<T, NS>(m, (T)1, WmXv, WmXs, wmR);
        riemann_side<T, NS>(m, (T)1, wpL, wmR, H0v, H0s);
#pragma unroll
        for (int k = 0; k < NS; ++k) wpL[k] = exA[(sd * NS + k) * NC + c - BX];
        left_chars_side<T, NS>(m, rvy, WmYv, WmYs, wmR);
        riemann_side<T, NS>(m, rvy, wpL, wmR, H1v, H1s);
    }
    // ---------------- stage 4: transverse corrections ----------------
    {
        T Hv[NS], Hs[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            exB[(sd * 2 * NS + k) * NC + c] = H1v[k];
            exB[(sd * 2 * NS + NS + k) * NC + c] = H1s[k];
            Hv[k] = lane_shift<DPP_FROM_NEXT>(H0v[k]);
            Hs[k] = lane_shift<DPP_FROM_NEXT>(H0s[k]);
        }
        __syncthreads();
        const T scale = K.scale;
        transverse_side<T, NS>(m, scale, H0v, H0s, Hv, Hs, WmYv, WmYs);
        transverse_side<T, NS>(m, scale, H0v, H0s, Hv, Hs, WpYv, WpYs);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            Hv[k] = exB[(sd * 2 * NS + k) * NC + c + BX];
            Hs[k] = exB[(sd * 2 * NS + NS + k) * NC + c + BX];
        }
        const T scx = sd ? rv1 * scale : scale;
        transverse_side<T, NS>(m, scx, H1v, H1s, Hv, Hs, WmXv, WmXs);
        transverse_side<T, NS>(m, scx, H1v, H1s, Hv, Hs, WpXv, WpXs);
    }
    // ---------------- stage 5: final face states ----------------
    Faces1<T, NS> F;
    {
        T cpx[NS], cpy[NS], wpL[NS], wmR[NS];
        right_chars_side<T, NS>(m, (T)1, WpXv, WpXs, cpx);
        right_chars_side<T, NS>(m, rvy, WpYv, WpYs, cpy);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            exA[(sd * NS + k) * NC + c] = cpy[k];
            wpL[k] = lane_shift<DPP_FROM_PREV>(cpx[k]);
        }
        __syncthreads();
        left_chars_side<T, NS>(m, (T)1, WmXv, WmXs, wmR);
        riemann_side<T, NS>(m, (T)1, wpL, wmR, F.xlv, F.xls);
#pragma unroll
        for (int k = 0; k < NS; ++k) wpL[k] = exA[(sd * NS + k) * NC + c - BX];
        left_chars_side<T, NS>(m, rvy, WmYv, WmYs, wmR);
        riemann_side<T, NS>(m, rvy, wpL, wmR, F.ylv, F.yls);
    }
    // ---------------- stage 6: flux difference and update ----------------
    {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            exB[(sd * 2 * NS + k) * NC + c] = F.ylv[k];
            exB[(sd * 2 * NS + NS + k) * NC + c] = F.yls[k];
            F.xhv[k] = lane_shift<DPP_FROM_NEXT>(F.xlv[k]);
            F.xhs[k] = lane_shift<DPP_FROM_NEXT>(F.xls[k]);
        }
        __syncthreads();
        if (pt.F[0] && lx >= 1 && lx <= tnx + 1 && ly >= 1 && ly <= tny)
            store_flux_side<T, NS>(pt, m, 0, F.xlv, F.xls, (T)1, i, j, sd);
        if (pt.F[1] && ly >= 1 && ly <= tny + 1 && lx >= 1 && lx <= tnx)
            store_flux_side<T, NS>(pt, m, 1, F.ylv, F.yls, rvy, i, j, sd);
        if (lx >= 1 && lx <= tnx && ly >= 1 && ly <= tny) {
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                F.yhv[k] = exB[(sd * 2 * NS + k) * NC + c + BX];
                F.yhs[k] = exB[(sd * 2 * NS + NS + k) * NC + c + BX];
            }
            update_cell_side<T, NS>(a, pt, P, m, dt, K.sc, i, j, sd, F);
        }
    }
}

// Occupancy target (waves per SIMD) of the interior kernel.  FP64 mode II fits 128 VGPRs,
// so two 512-thread workgroups share a CU and one's barriers overlap the other's work.
#ifndef TMK_INTERIOR_WAVES_F64
#define TMK_INTERIOR_WAVES_F64 4
#endif
#ifndef TMK_SPLIT_WAVES_F64
#define TMK_SPLIT_WAVES_F64 0
#endif
// one thread per (cell, side) for FP64 mode II on tiles of at most 384 cells (32x8, 32x12,
// 64x4): three or two workgroups of 8 or 12 waves share a CU, six waves per SIMD at 80 VGPRs.
// Larger tiles keep one thread per cell (a 1024-thread split workgroup would run alone on its CU).
// FP32 bulk keeps one thread per cell: it already holds six waves per SIMD at 74 VGPRs, and the
// split's smaller tiles cost more in ring overlap than the extra waves return (measured on the
// 4096 x 2048 level: 0.60 ms at 32x8 split against 0.57 ms at 32x16).
template <typename T, int NS, int BX, int BY>
__host__ __device__ constexpr bool interior_split()
{
    return sizeof(T) == 8 && NS == 2 && BX * BY <= 384;
}
template <typename T, int NS, int BX, int BY>
__host__ __device__ constexpr int interior_threads()
{
    return interior_split<T, NS, BX, BY>() ? 2 * BX * BY : BX * BY;
}
#ifndef TMK_INTERIOR_WAVES_F32
#define TMK_INTERIOR_WAVES_F32 0
#endif
template <typename T, int NS, int BX, int BY>
constexpr int interior_waves()
{
    return interior_split<T, NS, BX, BY>() ? TMK_SPLIT_WAVES_F64
           : sizeof(T) == 8                ? (NS == 2 ? TMK_INTERIOR_WAVES_F64 : 0)
                                           : TMK_INTERIOR_WAVES_F32;
}

// interior tiles only (the host sorts them out with the same rule); a tile that is not
// interior reports TMK_E_BOX and is left untouched.  Tiles smaller than (BX-2) x (BY-2) (at box
// edges) are interior too when their stencil lies inside the domain.
template <typename T, int NS, int BX, int BY, int PK>
__global__ void __launch_bounds__((interior_threads<T, NS, BX, BY>()), (interior_waves<T, NS, BX, BY>()))
k_godunov_interior(const KArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    const Tile tl = a.tiles[blockIdx.x];
    Scen& S = a.scen[a.patches[tl.patch].scen];
    if (!tile_interior<BX, BY>(tl, S.P)) {
        if (threadIdx.x == 0) report(&S.err, TMK_E_BOX);
        return;
    }
    if constexpr (interior_split<T, NS, BX, BY>())
        interior_tile_split<T, NS, BX, BY, PK>(a, tl, sm);
    else
        interior_tile<T, NS, BX, BY, PK>(a, tl, sm);
}

// Every update kernel the library launches: tile shapes (tmk_set_tile) x stencil-slot count x
// predictor kind.  Used to declare and to instantiate the FP32-bulk kernels, which are compiled
// in their own translation unit (tmk_f32.hip).
#define TMK_TILE_SHAPES(M, T, NS, PK)                                                           \
    M(T, NS, 16, 16, PK) M(T, NS, 32, 8, PK) M(T, NS, 32, 12, PK) M(T, NS, 32, 16, PK)        \
    M(T, NS, 32, 32, PK) M(T, NS, 64, 4, PK) M(T, NS, 64, 8, PK) M(T, NS, 64, 16, PK)
#define TMK_ALL_UPDATE_KERNELS(M, T)                                                            \
    TMK_TILE_SHAPES(M, T, 2, PK_RUNTIME) TMK_TILE_SHAPES(M, T, 3, PK_RUNTIME)                   \
    TMK_TILE_SHAPES(M, T, 2, PK_PPM_PRIM) TMK_TILE_SHAPES(M, T, 3, PK_PPM_PRIM)

}  // namespace tmk
