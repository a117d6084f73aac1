// tmk_march.hpp -- the interior update by row marching: one wavefront per strip and side.
//
// A strip is up to MARCH_COLS columns wide and any number of rows high, with its whole stencil
// (+-3 cells) inside the domain.  Lane l of the wavefront owns column x0 - 3 + l; the strip's
// columns are lanes 3 .. 3 + MARCH_COLS - 1, the other six lanes carry the x stencil.  The
// wavefront marches up the rows: after it loads row r it has what finishes row r - 3:
//
//   y slopes and face values   computed once per row and carried in registers to the next rows
//                              (the y stencil of a cell is its own column);
//   x slopes and face values   computed once per cell, the neighbours' by DPP lane shifts;
//   predictor face states      x faces from the left lane (DPP), y faces from the row below
//                              (registers);
//   transverse corrections, final face states, flux difference and update as in the tile
//   kernels (LinElastPatchGodunov.cpp:101-129, :174-454; LinElastPhysics.cpp:135-179).
//
// Every value is the one the tile kernels and the oracle compute for that cell, by the same
// expressions in the same order: a slope, face value or face state shared by two cells is
// computed once, not twice (the face value of x + 1/2 is the same expression whether it is seen
// as the hi face of cell x or the lo face of cell x + 1, FP additions being commutative).  The
// FP64 kernel is therefore bit-identical to the oracle like the tile kernels.  There is no
// barrier (each wavefront uses LDS of its own); the two sides of a strip are independent
// wavefronts (inside the domain they meet only at the fault, which no strip touches).  Mode II (NS = 2) with PPM and primitive
// limiting only: other numerics use the tile kernels.
#pragma once

#include "tmk_interior.hpp"

namespace tmk {

constexpr int MARCH_COLS = 58;  // 64 lanes minus the 3-cell x stencil on each side
constexpr unsigned kDropOffset = 0xFFFFFF00u;  // buffer offset of a lane whose access must be dropped

// Rows requested ahead of their use into an LDS ring (one slot per row: the 5 stencil planes of
// row q and the pointwise planes szz, gam of row q - 3), filled by buffer loads that write LDS
// directly, so the requests in flight hold no VGPRs.  FP32: 4 slots x 7 planes x 64 lanes x 4 B =
// 7 KiB per wavefront, 16 wavefronts per CU use 112 of its 160 KiB.  FP64: each plane row is two
// loads (low and high words), 3 slots (10.5 KiB with plasticity), 8 wavefronts per CU.  The
// ring's operations that can be in flight at a wait fit the 6-bit counter.
constexpr int MARCH_RING = 4;  // slots at most
constexpr int MARCH_RING_LOADS = 7;   // loads per slot
constexpr int MARCH_ROW_STORES = 8;   // stores per row: 5 stencil planes, szz, lam, gam

// a strip whose stencil lies in the domain (the same rule as tile_interior)
__device__ __forceinline__ bool march_interior(const Tile& t, const DevParams& P)
{
    return t.x1 - t.x0 + 1 <= MARCH_COLS && t.x0 - 3 >= P.dlo[0] && t.x1 + 3 <= P.dhi[0] && t.y0 - 3 >= P.dlo[1] &&
           t.y1 + 3 <= P.dhi[1];
}

// the monotonized central slope of a cell from its two one-sided differences (vl_slope, side 0)
template <typename T>
__device__ __forceinline__ T mc_slope(T dl, T dr)
{
    return vl_limit_centred((T)0.5 * (dr + dl), dl, dr);
}

// PPM face value between cells a (lo) and b (hi) with slopes da, db (ppm_incr's fl / fh)
template <typename T>
__device__ __forceinline__ T ppm_face(T wa, T wb, T da, T db)
{
    const T sixth = (T)(1.0 / 6.0);
    return (T)0.5 * (wb + wa) + sixth * (da - db);
}

// Buffer descriptors over one state plane: the kernel addresses a plane with a 32-bit byte
// offset per lane from a descriptor in scalar registers (no 64-bit address arithmetic per
// access), and the hardware bound check makes a load past the plane read 0.  Dword 3 holds the
// raw-access format of the gfx9 family (32-bit data format).
// The plane pointer and size must be the same in every lane (they come from the workgroup's
// patch descriptor); readfirstlane states that, so the descriptor lives in scalar registers.
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t plane_rsrc(const void* p, unsigned bytes)
{
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    void* q = (void*)(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(q, 0, (int)__builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(rsrc_t r, unsigned off)
{
    if constexpr (sizeof(T) == sizeof(double)) {
        using u2 = __attribute__((__vector_size__(2 * sizeof(unsigned)))) unsigned;
        return __builtin_bit_cast(double, (u2)__builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    } else {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    }
}
template <typename T>
__device__ __forceinline__ void bst(rsrc_t r, unsigned off, T v)
{
    if constexpr (sizeof(T) == sizeof(double)) {
        using u2 = __attribute__((__vector_size__(2 * sizeof(unsigned)))) unsigned;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)off, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
    }
}

// one plane row into LDS: lane l's dword at dst[l] (the LDS address is the wave-uniform dst plus
// 4 l); an offset past the plane reads 0
__device__ __forceinline__ void lds_row(rsrc_t r, float* dst, unsigned off)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 4, off, 0, 0, 0);
}

// wait until at most N vector memory operations of this wavefront are outstanding, and return
// addr (an LDS address) through the wait: reads at the returned address depend on it and so
// stay after it, while nothing else is tied to it (no memory clobber, which would make the
// compiler reload kernel arguments from memory inside the loop and wait for them).  Loads into
// LDS are not tracked by the compiler's own waits, so the kernel states them.  The counter
// retires loads and stores in issue order, so N no larger than the number of operations issued
// after a slot's loads makes the slot complete.
template <int N>
__device__ __forceinline__ unsigned wait_vmem(unsigned addr)
{
    static_assert(N >= 0 && N <= 63, "vmcnt is a 6-bit count");
    asm volatile("s_waitcnt vmcnt(%1)" : "+v"(addr) : "n"(N));
    return addr;
}

// the wait before a row of the march reads its ring slot, k = the number of rows that stored
// since the slot was requested, up to RING - 1 (k <= 0: none): after the slot's loads come the
// LOADS loads of each of the RING - 1 later slots and the STORES stores of each of those k rows
template <int RING, int LOADS, int STORES, int K = 0>
__device__ __forceinline__ unsigned wait_slot(int k, unsigned at)
{
    constexpr int N = (RING - 1) * LOADS + K * STORES;
    if constexpr (K == RING - 1) {
        return wait_vmem<N>(at);
    } else {
        if (k <= K) return wait_vmem<N>(at);
        return wait_slot<RING, LOADS, STORES, K + 1>(k, at);
    }
}

// lane shifts inside the wavefront (lane l <- lane l -+ 1; the first / last lane reads 0), in a
// form the compiler folds into the consuming FP32 instruction
template <int CTRL>
__device__ __forceinline__ float dshift(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dshift(double v)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// staged quantities of one side: 0 vx, 1 vy, 2 sxx, 3 syy, 4 sxy
// slot form per sweep: x: N (vx, sxx), S (vy, sxy); y: N (vy, syy), S (vx, sxy)
constexpr int MQ_XV[2] = {0, 1}, MQ_XS[2] = {2, 4};
constexpr int MQ_Y[4] = {1, 0, 3, 4};  // y-sweep quantities: vy vx syy sxy (v of slots N S, s of N S)

// predictor states of one side and sweep: W- (lo face) and W+ (hi face), slot form
template <typename T>
struct Pred2 {
    T mv[2], ms[2], pv[2], ps[2];
};

// what one row of the march hands to the next (r = the row that row loaded)
template <typename T>
struct MarchCarry {
    T wn[5];               // W of row r + 1 (requested ahead, without the ring)
    T w1[5], w2[5], w3[5]; // W of rows r, r - 1, r - 2 (the next row's r - 1, r - 2, r - 3)
    T dly[4];              // y differences W(r) - W(r - 1) of the y-sweep quantities
    T dy2[4];              // y slopes of row r - 1
    T fy[4];               // y face values at r - 3/2
    Pred2<T> xp;           // x predictor of row r - 2, not yet transversely corrected
    T h1v[2], h1s[2];      // predictor state of the lo y face of row r - 2
    T cpu[2];              // right characteristics of row r - 2's uncorrected W+ (y)
    T cpc[2];              // the same of its corrected W+
    T f1v[2], f1s[2];      // final state of the lo y face of row r - 2
};

// grid: two workgroups of one wavefront per strip, side = blockIdx.x & 1 (a scalar, so every
// plane pointer and descriptor of the side stays in scalar registers).  FP32: at least four
// wavefronts per SIMD (at most 128 VGPRs) keep enough loads in flight to cover their latency;
// FP64 needs about twice the registers (two wavefronts per SIMD) and is left to the compiler.
// PL: the launch may hold scenarios with plasticity.  Without (PL false), a row neither reads gam
// nor writes lam and gam, and a scenario with plasticity is an error.
template <typename T, int PK, bool PL>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == sizeof(float) ? 4 : 1)))
k_godunov_march(const KArgs<T> a)
{
    static_assert(PK == PK_PPM_PRIM, "row marching carries PPM with primitive limiting");
    const Tile tl = a.tiles[blockIdx.x >> 1];
    const Patch<T>& pt = a.patches[tl.patch];
    Scen& S = a.scen[pt.scen];
    const DevParams& P = S.P;
    if (!march_interior(tl, P)) {
        if (threadIdx.x == 0 && (blockIdx.x & 1) == 0) report(&S.err, TMK_E_BOX);
        return;
    }
    const Mat<T> m(P);
    const int lane = threadIdx.x;
    const int sd = blockIdx.x & 1;  // side of this wavefront
    const int i = tl.x0 - 3 + lane;
    const int tnx = tl.x1 - tl.x0 + 1;
    const bool valid = lane >= 3 && lane <= tnx + 2;  // column of the strip
    const double dt = S.dt;
    const StepConst<T> K = step_const<T>(S);
    const T dtbydx = K.dtbydx;
    const T scale = K.scale;
    const T rv1 = (T)-1;
    const T rvy = sd ? rv1 : (T)1;  // velocity sign of this side in y
    const T scx = sd ? rv1 * scale : scale;
    if (!PL && P.plast && a.plastic) {
        if (threadIdx.x == 0 && (blockIdx.x & 1) == 0) report(&S.err, TMK_E_ARG);
        return;
    }
    const bool plastic = PL && P.plast && a.plastic;
    const int o = 11 * sd;
    // stencil planes vx vy sxx syy sxy, read with the stencil (the in copies)
    constexpr int QC[5] = {0, 1, 3, 4, 6};
    rsrc_t rin[5], rout[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        rin[q] = plane_rsrc(pt.in[o + QC[q]], pt.pbytes);
        rout[q] = plane_rsrc(pt.out[o + QC[q]], pt.pbytes);
    }
    // pointwise planes, updated in place: szz, then lam gam (plasticity); index c - 6 for c = 9, 10
    rsrc_t rpw[5];
    rpw[0] = plane_rsrc(pt.out[o + 5], pt.pbytes);
    rpw[1] = rpw[2] = rpw[0];  // sxz, syz: not accessed in mode II
#pragma unroll
    for (int c = 9; c <= 10; ++c) rpw[c - 6] = plane_rsrc(pt.out[o + c], pt.pbytes);
    const unsigned col = (unsigned)(i - pt.ulo0) * (unsigned)sizeof(T);
    const unsigned rowb = (unsigned)pt.unx * (unsigned)sizeof(T);
    // values the loop needs from the patch descriptor and the parameters, read once here: after
    // the loop's first stores the compiler could only read them again with vector loads, each
    // waiting for every request in flight
    const int ulo1 = __builtin_amdgcn_readfirstlane(pt.ulo1);
    const bool flux0 = pt.F[0] != nullptr, flux1 = pt.F[1] != nullptr;
    auto row_off = [=](int j) { return col + (unsigned)(j - ulo1) * rowb; };
    const int y0 = tl.y0, y1 = tl.y1;

    const int rlast = y1 + 3;  // the last row loaded (row y1 is updated after it)
    constexpr bool f32 = sizeof(T) == sizeof(float);
    constexpr bool peel = !f32;  // the first two rows of the march as code of their own
    // FP64 without plasticity requests one row ahead into VGPRs instead (its loop reads nothing
    // else from memory, and the ring's extra loads and LDS reads measured slower there)
    constexpr bool ring_fetch = f32 || PL;
    // the ring: an element is WORDS dwords, each plane row of a slot WORDS loads of one dword per
    // lane (FP64: low words, then high words); a slot holds NPL planes (gam only with plasticity)
    // and a row of the strip stores ROW_STORES planes (lam and gam only with plasticity).  As many
    // slots, at most MARCH_RING, as keep the ring's operations that can be in flight at a wait --
    // the slot's own loads, the later slots' and the stores of RING - 1 rows -- within what the
    // 6-bit counter holds (flux-register stores of hierarchy levels come on top)
    constexpr int WORDS = (int)(sizeof(T) / sizeof(float));
    constexpr int NPL = PL ? MARCH_RING_LOADS : MARCH_RING_LOADS - 1;
    constexpr int SLOT_LOADS = NPL * WORDS;
    constexpr int ROW_STORES = PL ? MARCH_ROW_STORES : MARCH_ROW_STORES - 2;
    constexpr int RING_FIT = (63 + ROW_STORES) / (SLOT_LOADS + ROW_STORES);
    constexpr int RING = RING_FIT < MARCH_RING ? RING_FIT : MARCH_RING;
    static_assert(RING >= 2 && RING * SLOT_LOADS + (RING - 1) * ROW_STORES <= 63, "vmcnt counts to 63");
    // one LDS object: the ring, then the plastic return's constants (PlasticConstL for FP32,
    // PlasticConstL64 for FP64), then the lanes' flux-register face masks (below)
    constexpr int RING_FLOATS = (ring_fetch ? RING * SLOT_LOADS : 0) * 64;
    __shared__ float ring[RING_FLOATS + 16 + 4 * 64];
    float* const pconst = &ring[RING_FLOATS];
    if constexpr (f32) {
        if (lane < 14) pconst[lane] = lane < 7 ? P.pbg[lane] : lane < 13 ? P.pf[lane - 7] : P.fG;
    } else {
        if (lane < PD_N) ((double*)pconst)[lane] = plastic_d(P, lane);
    }
    const PlasticConstL pc{(const volatile __attribute__((address_space(3))) float*)&pconst[0]};
    const PlasticConstL64 pc64{(const volatile __attribute__((address_space(3))) double*)(double*)&pconst[0]};
    // Flux registers (levels of a hierarchy): which faces this lane stores on each row of the
    // strip -- bit j - y0 of a 64-bit mask for its lo x face and one for its lo y face, words
    // [0, 64) / [64, 128) and [128, 192) / [192, 256) -- decided here from the level's face masks,
    // with the flux arrays' extents read once: the row loop then reads no descriptor and no mask
    // (a vector load there would wait for every row request in flight).  Strips of such levels
    // have at most 63 rows (the group build caps them).
    volatile __attribute__((address_space(3))) unsigned* const fsel =
        (volatile __attribute__((address_space(3))) unsigned*)(__attribute__((address_space(3))) float*)&ring[
            RING_FLOATS + 16];
    const int fvlo0 = __builtin_amdgcn_readfirstlane(pt.vlo0), fvlo1 = __builtin_amdgcn_readfirstlane(pt.vlo1);
    const int fvnx = __builtin_amdgcn_readfirstlane(pt.vhi0) - fvlo0 + 1;
    const int fvny = __builtin_amdgcn_readfirstlane(pt.vhi1) - fvlo1 + 1;
    // the flux arrays' addresses in scalar registers (uniform): the stores then take a scalar
    // base and a vector offset, and hold no 64-bit vector address through the loop
    auto uniform_ptr = [](const double* q) {
        const unsigned long long u = (unsigned long long)(uintptr_t)q;
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
        return (__attribute__((address_space(1))) double*)(uintptr_t)(((unsigned long long)hi << 32) | lo);
    };
    __attribute__((address_space(1))) double* const Fo0 = uniform_ptr(pt.F[0]);
    __attribute__((address_space(1))) double* const Fo1 = uniform_ptr(pt.F[1]);
    if ((flux0 || flux1) && y1 - y0 + 1 > 63) {  // the masks hold 64 rows of faces
        if (lane == 0 && sd == 0) report(&S.err, TMK_E_BOX);
        return;
    }
    if (flux0 || flux1) {
        unsigned long long m0 = 0, m1 = 0;
        const bool lx = flux0 && lane >= 3 && lane <= tnx + 3, ly = flux1 && valid;
        if (!pt.fmask) {
            m0 = lx ? ~0ull : 0ull;
            m1 = ly ? ~0ull : 0ull;
        } else if (lx || ly) {
            const __attribute__((address_space(1))) unsigned char* mk = gmem(pt.fmask);
            const int mw = fvnx + 2;  // the masks cover the descriptor's valid box grown by one
            for (int j = y0; j <= y1 + 1; ++j) {
                const size_t q = (size_t)(i - fvlo0 + 1) + (size_t)(j - fvlo1 + 1) * mw;
                const unsigned b = mk[q];
                if (lx && j <= y1 && register_face(mk[q - 1], b)) m0 |= 1ull << (j - y0);
                if (ly && register_face(mk[q - mw], b)) m1 |= 1ull << (j - y0);
            }
        }
        fsel[lane] = (unsigned)m0;
        fsel[64 + lane] = (unsigned)(m0 >> 32);
        fsel[128 + lane] = (unsigned)m1;
        fsel[192 + lane] = (unsigned)(m1 >> 32);
    }
    // bit k (0 .. 63) of this lane's mask w (0: lo x face, 1: lo y face)
    auto fbit = [&](int w, int k) { return ((fsel[128 * w + 64 * (k >> 5) + lane] >> (k & 31)) & 1u) != 0; };
    // this lane's dword in slot 0, plane 0 (an LDS byte address)
    const unsigned ring_lane =
        (unsigned)(uintptr_t)(__attribute__((address_space(3))) float*)&ring[0] + 4u * (unsigned)lane;
    // the rows row(q) reads into slot (q - y0) mod RING: every row issues exactly SLOT_LOADS loads
    // (rows past the strip and lanes outside it read past the planes)
    auto request = [&](const int q) {
        const unsigned sl = (unsigned)(q - y0) % (unsigned)RING;
        const bool live = q <= rlast;
        const unsigned offs = live ? row_off(q) : kDropOffset;
        const unsigned offp = (live && valid && q - 3 >= y0) ? row_off(q - 3) : kDropOffset;
        auto put = [&](rsrc_t rs, int p, unsigned off) {
#pragma unroll
            for (int w = 0; w < WORDS; ++w)
                lds_row(rs, &ring[((sl * NPL + p) * WORDS + w) * 64], off + 4u * (unsigned)w);
        };
#pragma unroll
        for (int p = 0; p < 5; ++p) put(rin[p], p, offs);
        put(rpw[0], 5, offp);
        if constexpr (PL) put(rpw[4], 6, plastic ? offp : kDropOffset);
    };

    // warm-up: rows y0 - 3 .. y0 give the y slopes of rows y0 - 2, y0 - 1 and the face value at
    // y0 - 3/2 (what the first main iteration carries in)
    MarchCarry<T> A, B;
#pragma unroll
    for (int q = 0; q < 5; ++q) A.w1[q] = A.w2[q] = A.w3[q] = (T)0;
#pragma unroll
    for (int q = 0; q < 4; ++q) A.dly[q] = A.dy2[q] = A.fy[q] = (T)0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        T w0[5];
        const unsigned off = row_off(y0 - 3 + r);
#pragma unroll
        for (int q = 0; q < 5; ++q) w0[q] = bld<T>(rin[q], off);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T dry = w0[MQ_Y[q]] - A.w1[MQ_Y[q]];
            const T dy1 = mc_slope(A.dly[q], dry);
            const T fy1 = ppm_face(A.w2[MQ_Y[q]], A.w1[MQ_Y[q]], A.dy2[q], dy1);
            A.dly[q] = dry;
            A.dy2[q] = dy1;
            A.fy[q] = fy1;
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            A.w2[q] = A.w1[q];
            A.w1[q] = w0[q];
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        A.xp.mv[k] = A.xp.ms[k] = A.xp.pv[k] = A.xp.ps[k] = (T)0;
        A.h1v[k] = A.h1s[k] = A.cpu[k] = A.cpc[k] = A.f1v[k] = A.f1s[k] = (T)0;
    }
    if constexpr (ring_fetch) {
#pragma unroll
        for (int q = 1; q < RING; ++q) request(y0 + q);
    } else {
        const unsigned off = row_off(y0 + 1);
#pragma unroll
        for (int q = 0; q < 5; ++q) A.wn[q] = bld<T>(rin[q], off);
    }

    // one row of the march: after loading row r, the predictors of row r - 2 and the update of
    // row r - 3, from what c carries in; n receives what the next row needs.  The first row
    // computes row y0 - 1's face states (used) and row y0 - 2's (not used).  With the ring, loads
    // run RING - 1 rows ahead: row r + RING - 1's stencil planes (ghost rows of the array at the
    // last rows) and row r + RING - 4's pointwise planes are requested before the row's
    // arithmetic, so the wavefront computes while they arrive; without it (FP64 without
    // plasticity), row r + 1's stencil planes into VGPRs.
    // The first two rows of the march only feed the faces of the rows after them: they skip row
    // r - 3's final x faces, update and stores.  FP64 runs them as separate code (head = true),
    // so that the loop's rows have no branch around the update; FP32 runs every row in the loop
    // and branches (this measured faster for FP32: its loop body stays one copy).
    auto row = [&](auto head, const int r, const MarchCarry<T>& c, MarchCarry<T>& n) {
        constexpr bool H = decltype(head)::value;
        const int j3 = r - 3;
        const bool strip_row = !H && (peel || j3 >= y0);  // the same in every lane
        // row r - 3's pointwise planes: szz, and gam with plasticity (in mode II sxz and syz
        // stay zero and the plastic multiplier lam is only written); a lane outside the strip,
        // or a row before it, gets an offset at or past every plane's bound (planes hold at most
        // 2^32 - 256 bytes), far enough below 2^32 that the second dword of an 8-byte access
        // does not wrap: the bound check makes its loads read 0 and drops its stores
        const unsigned off3 = (valid && strip_row) ? row_off(j3) : kDropOffset;
        T w0[5], szz, gam;
        if constexpr (ring_fetch) {
            // request the row RING - 1 ahead, then wait for this row's slot.  Issued after that
            // slot's loads: the RING - 1 later slots (SLOT_LOADS loads each) and the stores of the
            // rows since (ROW_STORES each).  The slots of the first rows were requested before
            // the march, and its first two rows (rows y0 + 1, y0 + 2) store nothing, so the rows
            // that stored since this slot's request number r - (y0 + 3), at most RING - 1.
            request(r + RING - 1);
            const unsigned sl = (unsigned)(r - y0) % (unsigned)RING;
            const unsigned at = wait_slot<RING, SLOT_LOADS, ROW_STORES>(
                r - (y0 + 3), ring_lane + sl * (unsigned)(SLOT_LOADS * 64 * sizeof(float)));
            const __attribute__((address_space(3))) float* slot =
                (const __attribute__((address_space(3))) float*)(uintptr_t)at;
            // plane p of the slot (FP64: its low and high words, 64 lanes apart)
            auto sval = [&](int p) -> T {
                if constexpr (WORDS == 1) {
                    return slot[64 * p];
                } else {
                    const unsigned long long lo = __float_as_uint(slot[64 * (2 * p)]);
                    const unsigned long long hi = __float_as_uint(slot[64 * (2 * p + 1)]);
                    return __longlong_as_double((long long)((hi << 32) | lo));
                }
            };
#pragma unroll
            for (int q = 0; q < 5; ++q) w0[q] = sval(q);
            szz = sval(5);
            if constexpr (PL) gam = plastic ? sval(6) : (T)0;
            else gam = (T)0;
        } else {
#pragma unroll
            for (int q = 0; q < 5; ++q) w0[q] = c.wn[q];
            const unsigned offn = row_off(r + 1);
#pragma unroll
            for (int q = 0; q < 5; ++q) n.wn[q] = bld<T>(rin[q], offn);
            szz = bld<T>(rpw[0], off3);
            gam = (T)0;
        }
        // y: differences, slopes of row r - 1, face values at r - 3/2
        T dry[4], dy1[4], fy1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            dry[q] = w0[MQ_Y[q]] - c.w1[MQ_Y[q]];
            dy1[q] = mc_slope(c.dly[q], dry[q]);
            fy1[q] = ppm_face(c.w2[MQ_Y[q]], c.w1[MQ_Y[q]], c.dy2[q], dy1[q]);
        }
        // y predictor of row r - 2 (its y faces r - 5/2 and r - 3/2)
        Pred2<T> yp;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const T wv = c.w2[MQ_Y[k]], ws = c.w2[MQ_Y[2 + k]];
            const T dvm = incr(c.fy[k], wv), dvp = incr(fy1[k], wv);
            const T dsm = incr(c.fy[2 + k], ws), dsp = incr(fy1[2 + k], ws);
            ppm_slot<T>(m, k, rvy, dtbydx, true, false, wv, ws, dvm, dvp, dsm, dsp, yp.mv[k], yp.ms[k], yp.pv[k],
                        yp.ps[k]);
        }
        // x predictor of row r - 2: slopes and face values once per lane, neighbours by DPP
        Pred2<T> xn;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            T inc[2][2];  // [v|s][m|p]
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const T wc = c.w2[h == 0 ? MQ_XV[k] : MQ_XS[k]];
                const T wr = dshift<DPP_FROM_NEXT>(wc);  // W(i + 1)
                const T dr = wr - wc;
                const T dl = dshift<DPP_FROM_PREV>(dr);  // W(i) - W(i - 1)
                const T d0 = mc_slope(dl, dr);
                const T fh = ppm_face(wc, wr, d0, dshift<DPP_FROM_NEXT>(d0));  // face value at i + 1/2
                const T fl = dshift<DPP_FROM_PREV>(fh);                        // at i - 1/2
                inc[h][0] = incr(fl, wc);
                inc[h][1] = incr(fh, wc);
            }
            ppm_slot<T>(m, k, (T)1, dtbydx, true, false, c.w2[MQ_XV[k]], c.w2[MQ_XS[k]], inc[0][0], inc[0][1],
                        inc[1][0], inc[1][1], xn.mv[k], xn.ms[k], xn.pv[k], xn.ps[k]);
        }
        // predictor state of the lo x face of row r - 2 (its left lane's W+ and its W-)
        T h0v[2], h0s[2];
        {
            T cpx[2], wpL[2], wmR[2];
            right_chars_side<T, 2>(m, (T)1, xn.pv, xn.ps, cpx);
#pragma unroll
            for (int k = 0; k < 2; ++k) wpL[k] = dshift<DPP_FROM_PREV>(cpx[k]);
            left_chars_side<T, 2>(m, (T)1, xn.mv, xn.ms, wmR);
            riemann_side<T, 2>(m, (T)1, wpL, wmR, h0v, h0s);
        }
        // predictor state of the lo y face of row r - 2 (row r - 3's uncorrected W+, its W-)
        {
            T wmR[2];
            left_chars_side<T, 2>(m, rvy, yp.mv, yp.ms, wmR);
            riemann_side<T, 2>(m, rvy, c.cpu, wmR, n.h1v, n.h1s);
            right_chars_side<T, 2>(m, rvy, yp.pv, yp.ps, n.cpu);
        }
        // y predictor corrected with the x-face states of its cell (GETADWDXF, dir2 = x)
        {
            T hv[2], hs[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                hv[k] = dshift<DPP_FROM_NEXT>(h0v[k]);
                hs[k] = dshift<DPP_FROM_NEXT>(h0s[k]);
            }
            transverse_side<T, 2>(m, scale, h0v, h0s, hv, hs, yp.mv, yp.ms);
            transverse_side<T, 2>(m, scale, h0v, h0s, hv, hs, yp.pv, yp.ps);
        }
        // final state of the lo y face of row r - 2 (row r - 3's corrected W+, its corrected W-)
        {
            T wmR[2];
            left_chars_side<T, 2>(m, rvy, yp.mv, yp.ms, wmR);
            riemann_side<T, 2>(m, rvy, c.cpc, wmR, n.f1v, n.f1s);
            right_chars_side<T, 2>(m, rvy, yp.pv, yp.ps, n.cpc);
        }
        const int j2 = r - 2;
        // flux registers (multi-level): the lo y face of row r - 2, the x faces of row r - 3
        // (below); the lane's column, hidden from the compiler here, so that it builds each face's
        // address in place rather than keeping a 64-bit per-lane address alive through the loop
        if (flux1 && j2 >= y0 && fbit(1, j2 - y0)) {
            int ix = i - fvlo0;
            asm volatile("" : "+v"(ix));
            put_face_flux<T, 2>(Fo1, (size_t)ix + (size_t)(j2 - fvlo1) * fvnx, (size_t)fvnx * (fvny + 1), m, 1, n.f1v,
                                n.f1s, rvy, sd);
        }
        // row r - 3, a row of the strip: x predictor corrected with the y-face states below and
        // above it, its final x-face state and its update
        if (strip_row) {
            T f0v[2], f0s[2];
            Pred2<T> xp = c.xp;
            transverse_side<T, 2>(m, scx, c.h1v, c.h1s, n.h1v, n.h1s, xp.mv, xp.ms);
            transverse_side<T, 2>(m, scx, c.h1v, c.h1s, n.h1v, n.h1s, xp.pv, xp.ps);
            T cpx[2], wpL[2], wmR[2];
            right_chars_side<T, 2>(m, (T)1, xp.pv, xp.ps, cpx);
#pragma unroll
            for (int k = 0; k < 2; ++k) wpL[k] = dshift<DPP_FROM_PREV>(cpx[k]);
            left_chars_side<T, 2>(m, (T)1, xp.mv, xp.ms, wmR);
            riemann_side<T, 2>(m, (T)1, wpL, wmR, f0v, f0s);
            if (flux0 && fbit(0, j3 - y0)) {
                int ix = i - fvlo0;
                asm volatile("" : "+v"(ix));
                put_face_flux<T, 2>(Fo0, (size_t)ix + (size_t)(j3 - fvlo1) * (fvnx + 1), (size_t)(fvnx + 1) * fvny, m,
                                    0, f0v, f0s, (T)1, sd);
            }
            // update of row r - 3: every row of the strip issues exactly ROW_STORES stores (the
            // ring's waits count them)
            Faces1<T, 2> F;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                F.xlv[k] = f0v[k];
                F.xls[k] = f0s[k];
                F.xhv[k] = dshift<DPP_FROM_NEXT>(f0v[k]);
                F.xhs[k] = dshift<DPP_FROM_NEXT>(f0s[k]);
                F.ylv[k] = c.f1v[k];
                F.yls[k] = c.f1s[k];
                F.yhv[k] = n.f1v[k];
                F.yhs[k] = n.f1s[k];
            }
            // every lane runs the update (a lane outside the strip has off3 past the planes)
            T u[11];
            u[0] = c.w3[0];
            u[1] = c.w3[1];
            u[3] = c.w3[2];
            u[4] = c.w3[3];
            u[6] = c.w3[4];
            u[2] = u[7] = u[8] = u[9] = (T)0;
            u[5] = szz;
            u[10] = gam;
            if constexpr (f32)
                cell_update_side<T, 2>(pc, m, dt, K.sc, plastic, rvy, F, u);
            else
                cell_update_side<T, 2>(pc64, m, dt, K.sc, plastic, rvy, F, u);
#pragma unroll
            for (int q = 0; q < 5; ++q) bst<T>(rout[q], off3, u[QC[q]]);
            bst<T>(rpw[0], off3, u[5]);
            if constexpr (PL) {
                const unsigned offl = plastic ? off3 : kDropOffset;
                bst<T>(rpw[3], offl, u[9]);
                bst<T>(rpw[4], offl, u[10]);
            }
        }
        // carried to the next row
        n.xp = xn;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            n.dly[q] = dry[q];
            n.dy2[q] = dy1[q];
            n.fy[q] = fy1[q];
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            n.w3[q] = c.w2[q];
            n.w2[q] = c.w1[q];
            n.w1[q] = w0[q];
        }
    };

    // two rows per trip, the carried state alternating between A and B (no copies between rows;
    // the DPP lane shifts are convergent, so the compiler does not unroll the loop itself)
    // (rlast >= y0 + 3: a strip has at least one row)
    int r0 = y0 + 1;
    if constexpr (peel) {
        row(std::true_type{}, y0 + 1, A, B);
        row(std::true_type{}, y0 + 2, B, A);
        r0 = y0 + 3;
    }
    for (int r = r0;; r += 2) {
        row(std::false_type{}, r, A, B);
        if (r + 1 > rlast) break;
        row(std::false_type{}, r + 1, B, A);
        if (r + 2 > rlast) break;
    }
    // the last rows requested rows past the strip (loads the bound check drops, which still write
    // their zeros to the ring): every ring load lands before the wavefront ends and its LDS can
    // go to another workgroup (the last row's ROW_STORES stores are the only operations after
    // the last request)
    if constexpr (ring_fetch) (void)wait_vmem<ROW_STORES>(ring_lane);
}

}  // namespace tmk
