// tmk_physics.hpp -- pointwise physics of the per-patch Godunov update, device side.
//
// Every function restates one reference Fortran routine on the reduced
// "slot" form the fused kernel carries.  For a sweep direction d the elastic
// system only couples three velocity/traction pairs per side:
//   slot N (normal)   : (v_n, s_nn)   dir 0: (vx, sxx)   dir 1: (vy, syy)
//   slot S (in-plane) : (v_t, s_nt)   dir 0: (vy, sxy)   dir 1: (vx, sxy)
//   slot A (antiplane): (v_z, s_nz)   dir 0: (vz, sxz)   dir 1: (vz, syz)
// These are exactly the components read by RIEMANNF's wave characteristics,
// GETFLUXF and GETADWDXF (LinElastPhysicsF.ChF:265-353, :601-765, :791-901);
// the zero-speed invariants (syy - sxx lam/(2G+lam), ...) never reach a flux.
// Side 2 (the mirrored half space, LEINDEX.H:23) flips the velocity sign in
// direction 1 (rv = -1, LinElastPhysicsF.ChF:285-301).
//
// Expressions keep the reference's operation order so the FP64 path rounds
// like the reference when compiled without FMA contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmk {

// The plane pointers come out of the patch table in memory, so the compiler cannot tell that
// they address global memory and would emit flat loads and stores, whose 64-bit addresses cost
// vector instructions and which also count against the LDS wait counter.  gmem states it.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

// Element idx of a plane as the plane's base plus a 32-bit byte offset: the hardware adds that
// offset to a base held in scalar registers, with no 64-bit address arithmetic per access.
// Every plane the library addresses holds at most 4 GiB (checked where levels and the
// drop-in path's scratch are sized).
template <typename T>
__device__ __forceinline__ T gld(const T* base, unsigned idx)
{
    return *(const __attribute__((address_space(1))) T*)((const __attribute__((address_space(1))) char*)gmem(base) +
                                                          idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void gst(T* base, unsigned idx, T v)
{
    *(__attribute__((address_space(1))) T*)((__attribute__((address_space(1))) char*)gmem(base) + idx * (unsigned)sizeof(T)) = v;
}

// ---------------------------------------------------------------------------
// parameters (device copy of tmk_params plus derived constants)
// ---------------------------------------------------------------------------
struct DevParams {
    double cs, cp, G, rho, lam, K, Zp, Zs;
    double iZp, iZs, irho;  // correctly rounded reciprocals 1/Zp, 1/Zs, 1/rho
    double i2Zp, i2Zs;      // 1/(2 Zp), 1/(2 Zs): the reciprocals halved (exact)
    // the same constants rounded to FP32 on the host, for the FP32-bulk kernels
    float frho, flam, fG, fZp, fZs, fden, flam2G, fiZp, fiZs, firho, fi2Zp, fi2Zs, fcp, fcs;
    double den;    // 2G + lam   (LinElastPhysicsF.ChF:427 form)
    double lam2G;  // lam + 2G   (LinElastPhysicsF.ChF:318 form)
    double bg[22];  // the uniform background state (both sides alike)
    // the background in the form the FP32 Drucker-Prager return uses: mean stress, deviators
    // xx yy zz, then sxy sxz syz (rounded from FP64 on the host)
    float pbg[7];
    // FP32 plasticity constants: mu, cohesion b, hardening h, h + G + mu beta K, eta, K beta
    float pf[6];
    int plast;
    double p_mu, p_beta, p_eta, p_b, p_h;
    int bc[4];
    int fault_kind;
    double r0, nsig, rs_a, rs_b, rs_V0, rs_f0, rs_L, rs_fw, rs_Vw, rs_fexp;
    double x0, rs_thr, ramp_x, ramp_w, ramp_a, ramp_Vw;
    int state_law;
    double fs, fd, dw, sm, sw_thr;
    int npatch;
    double xc[8], xw[8], taup[8];
    double fric_c0, fric_w0, outside_f;
    int face_lo, nface;
    const double *f_a, *f_b, *f_L, *f_fs, *f_fd, *f_dc, *f_syy, *f_tau;
    int pred, fourth, prim_lim, char_lim;
    double dx;
    int dlo[2], dhi[2];
};

// FP32 bulk stores the difference from the uniform background, U - bg, instead of U.  The
// background stresses (around 1e8 Pa) are far larger than the waves the update moves (1e5 to
// 1e7 Pa), and an FP32 copy of the absolute value would round every cell at the ulp of the
// background, an error that the rupture's nucleation amplifies.  The elastic update is linear
// and its limiters see only differences, so the same kernels advance U - bg; the background
// enters where the update is not translation invariant: the absorbing and free-surface face
// states (bc_state), the fault solve (the FP64 traces get bg added and the face states bg
// removed), the Drucker-Prager return (plastic_flow_delta) and the flux outputs (the flux of the
// background added in FP64).  FP64 stores U itself, so it rounds like the reference.
template <typename T>
__host__ __device__ constexpr bool stores_delta()
{
    return sizeof(T) == sizeof(float);
}

// FP32 bulk arithmetic regroups some expressions into fewer instructions (constants folded
// first, differences taken before products); FP64 keeps the reference's operation order so that
// it rounds like the reference.
template <typename T>
__host__ __device__ constexpr bool regroup()
{
    return sizeof(T) == sizeof(float);
}

// f - w: the reference's (0 - w) + f in FP64
template <typename T>
__device__ __forceinline__ T incr(T f, T w)
{
    if constexpr (regroup<T>())
        return f - w;
    else
        return ((T)0 - w) + f;
}

template <typename T, int NS>
struct Slots {
    T v[2][NS];
    T s[2][NS];
};

// a / b for b > 0 given y = RN(1/b).  FP64: q0 = RN(a y) is within two ulps of a/b; one
// Newton-Markstein step q1 = RN(q0 + (a - b q0) y) is faithful (its error is below half an ulp
// plus |q0 - a/b| u); a second step from a faithful quotient and a correctly rounded reciprocal
// returns RN(a/b) exactly (Markstein, IBM J. Res. Dev. 34 (1990) 111-119; Muller et al.,
// Handbook of Floating-Point Arithmetic, 2nd ed., sec. 4.7).  The remainders a - b q are exact
// in one fma.  copysign keeps the sign of a zero dividend.  The result equals the IEEE quotient
// bit for bit, at five fma/mul instead of the 10-instruction scaled division sequence.
// FP32 (bulk arithmetic of the reduced-precision mode): a y, within 1.5 ulps.
__device__ __forceinline__ double divr(double a, double b, double y)
{
    double q = a * y;
    double r = fma(-q, b, a);
    q = fma(r, y, q);
    r = fma(-q, b, a);
    return copysign(fma(r, y, q), a);
}
__device__ __forceinline__ float divr(float a, float, float y) { return a * y; }

// FP64 value or its host-rounded FP32 copy
template <typename T>
__device__ __forceinline__ T pick(double d, float f)
{
    if constexpr (sizeof(T) == sizeof(double))
        return (T)d;
    else
        return (T)f;
}

// elastic constants in the bulk arithmetic type
template <typename T>
struct Mat {
    T rho, lam, G, Zp, Zs, den, lam2G;
    T iZp, iZs, irho, i2Zp, i2Zs;  // reciprocals (1/(2Z) = (1/Z)/2 exactly)
    T cp, cs;
    __device__ __forceinline__ explicit Mat(const DevParams& p)
        : rho(pick<T>(p.rho, p.frho)), lam(pick<T>(p.lam, p.flam)), G(pick<T>(p.G, p.fG)),
          Zp(pick<T>(p.Zp, p.fZp)), Zs(pick<T>(p.Zs, p.fZs)), den(pick<T>(p.den, p.fden)),
          lam2G(pick<T>(p.lam2G, p.flam2G)), iZp(pick<T>(p.iZp, p.fiZp)), iZs(pick<T>(p.iZs, p.fiZs)),
          irho(pick<T>(p.irho, p.firho)), i2Zp(pick<T>(p.i2Zp, p.fi2Zp)), i2Zs(pick<T>(p.i2Zs, p.fi2Zs)),
          cp(pick<T>(p.cp, p.fcp)), cs(pick<T>(p.cs, p.fcs)) {}
    __device__ __forceinline__ T Z(int k) const { return k == 0 ? Zp : Zs; }
    // x / Z(k), x / (2 Z(k)), x / rho, rounded like the division
    __device__ __forceinline__ T divZ(T x, int k) const { return k == 0 ? divr(x, Zp, iZp) : divr(x, Zs, iZs); }
    __device__ __forceinline__ T div2Z(T x, int k) const
    {
        return k == 0 ? divr(x, (T)2 * Zp, i2Zp) : divr(x, (T)2 * Zs, i2Zs);
    }
    __device__ __forceinline__ T divrho(T x) const { return divr(x, rho, irho); }
    __device__ __forceinline__ T i2Z(int k) const { return k == 0 ? i2Zp : i2Zs; }
};

// component index of (slot, velocity/stress) for direction d, side 1
__host__ __device__ constexpr int slot_comp(int d, int k, int is_s)
{
    return d == 0 ? (k == 0 ? (is_s ? 3 : 0) : k == 1 ? (is_s ? 6 : 1) : (is_s ? 7 : 2))
                  : (k == 0 ? (is_s ? 4 : 1) : k == 1 ? (is_s ? 6 : 0) : (is_s ? 8 : 2));
}

// ---------------------------------------------------------------------------
// slopes and PPM (restated from the published method; see oracle for the same form)
// ---------------------------------------------------------------------------
// van Leer: lim = min(2|l|, 2|r|, |c|), zero at extrema, with the sign of c.  2 min(|l|, |r|)
// equals min(2|l|, 2|r|) exactly (doubling is exact and monotone) and copysign(lim, c) equals
// lim * copysign(1, c) for lim >= 0, signed zeros included.
template <typename T>
__device__ __forceinline__ T vl_limit(T c, T l, T r)
{
    T lim = (T)2 * fmin(fabs(l), fabs(r));
    lim = fmin(lim, fabs(c));
    if (l * r < (T)0) lim = (T)0;
    return copysign(lim, c);
}

// vl_limit for c = (l + r) / 2 (every slope vl_slope and the row-marching kernel limit): then
// sign(c) = sign(l) = sign(r) whenever l r > 0 and the limited slope is minmod(2 l, 2 r, c).  FP32
// forms it from medians with zero: med3(a, b, 0) is whichever of a, b lies nearer zero when they
// share a sign and 0 otherwise (doubling is exact).
template <typename T>
__device__ __forceinline__ T vl_limit_centred(T c, T l, T r)
{
    if constexpr (regroup<T>()) {
        const T lr = (T)2 * __builtin_amdgcn_fmed3f(l, r, (T)0);
        return __builtin_amdgcn_fmed3f(lr, c, (T)0);
    } else {
        return vl_limit(c, l, r);
    }
}

// side classification of stencil cell: 0 centred, 1 domain lo, 2 domain hi
__device__ __forceinline__ int cell_side(int c, int dlo, int dhi) { return c == dlo ? 1 : (c == dhi ? 2 : 0); }

// second-order slope differences at stencil position q (0..4, centre 2)
template <typename T>
__device__ __forceinline__ void slope_diffs(const T* w, int q, int side, T& C, T& L, T& R)
{
    if (side == 0) {
        const T dr = w[q + 1] - w[q];
        const T dl = w[q] - w[q - 1];
        R = dr; L = dl; C = (T)0.5 * (dr + dl);
    } else if (side == 1) {
        const T dr = w[q + 1] - w[q];
        C = L = R = dr;
    } else {
        const T dl = w[q] - w[q - 1];
        C = L = R = dl;
    }
}

template <typename T>
__device__ __forceinline__ T vl_slope(const T* w, int q, int side, bool limit)
{
    T C, L, R;
    slope_diffs(w, q, side, C, L, R);
    return limit ? vl_limit_centred(C, L, R) : C;
}

// PPM face increments (face value minus cell value) at the lo and hi faces of
// the centre cell, Colella & Woodward (1984) eq. 1.6/1.9 on a uniform grid
template <typename T>
__device__ __forceinline__ void ppm_incr(const T* w, int ci, int dlo, int dhi, bool limit, T& dm, T& dp)
{
    const T sixth = (T)(1.0 / 6.0);
    const int s0 = cell_side(ci, dlo, dhi);
    const T d0 = vl_slope(w, 2, s0, limit);
    T fl, fh;
    if (ci == dlo) {
        fl = w[2] - (T)0.5 * d0;
    } else {
        const T dmm = vl_slope(w, 1, cell_side(ci - 1, dlo, dhi), limit);
        fl = (T)0.5 * (w[2] + w[1]) + sixth * (dmm - d0);
    }
    if (ci == dhi) {
        fh = w[2] + (T)0.5 * d0;
    } else {
        const T dpp = vl_slope(w, 3, cell_side(ci + 1, dlo, dhi), limit);
        fh = (T)0.5 * (w[3] + w[2]) + sixth * (d0 - dpp);
    }
    dm = incr(fl, w[2]);
    dp = incr(fh, w[2]);
}

// branch-free: both candidates are formed and one is selected (the same values as the
// branching form, without divergent control flow)
template <typename T>
__device__ __forceinline__ void ppm_limit(T& m, T& p)
{
    if constexpr (regroup<T>()) {
        // m, -2p share a sign exactly when m p < 0: the median with zero is then the one nearer
        // zero, copysign(min(|m|, 2|p|), m), and 0 otherwise
        const T nm = __builtin_amdgcn_fmed3f(m, (T)-2 * p, (T)0);
        const T np = __builtin_amdgcn_fmed3f(p, (T)-2 * m, (T)0);
        m = nm;
        p = np;
        return;
    }
    const bool keep = m * p < (T)0;
    const T nm = copysign(fmin((T)2 * fabs(p), fabs(m)), m);
    const T np = copysign(fmin((T)2 * fabs(m), fabs(p)), p);
    m = keep ? nm : (T)0;
    p = keep ? np : (T)0;
}

// The PPM predictor of slot k of one side from its face increments (face value minus cell value
// of v and s at the lo (m) and hi (p) faces): limiter in primitive or characteristic variables,
// CHARANALYSISF (LinElastPhysicsF.ChF:424-447), tracing of the wave moving toward lo on the lo
// face and toward hi on the hi face, CHARSYNTHESISF (:526-550), then W+- = increment + W.
template <typename T>
__device__ __forceinline__ void ppm_slot(const Mat<T>& m, int k, T rvs, T dtbydx, bool limit, bool charl, T wv, T ws,
                                         T dvm, T dvp, T dsm, T dsp, T& Wmv, T& Wms, T& Wpv, T& Wps)
{
    if (limit && !charl) {
        ppm_limit(dvm, dvp);
        ppm_limit(dsm, dsp);
    }
    if constexpr (regroup<T>()) {
        // the same tracing and synthesis with the constants folded: m + sig ((p - m) - (p + m)
        // (3 - 2 sig)) / 2 = m (1 - c1 - c2) + p (c1 - c2), c1 = sig / 2, c2 = c1 (3 - 2 sig);
        // the velocity of a characteristic pair is (cm - cp) rvs / (2 Z)
        const T zr = rvs * m.Z(k);
        T cm_m = fma(dvm, zr, dsm), cp_m = fma(-dvm, zr, dsm);
        T cm_p = fma(dvp, zr, dsp), cp_p = fma(-dvp, zr, dsp);
        if (charl) {
            ppm_limit(cm_m, cm_p);
            ppm_limit(cp_m, cp_p);
        }
        const T sig = (k == 0 ? m.cp : m.cs) * dtbydx;
        const T c1 = (T)0.5 * sig, c2 = c1 * ((T)3 - (T)2 * sig);
        const T ta = ((T)1 - c1) - c2, tb = c1 - c2;
        const T tm = fma(cm_m, ta, cm_p * tb);  // traced left-moving wave on the lo face
        const T tp = fma(cp_p, ta, cp_m * tb);  // traced right-moving wave on the hi face
        const T h = rvs * m.i2Z(k);
        Wmv = fma(tm - cp_m, h, wv);
        Wpv = fma(cm_p - tp, h, wv);
        Wms = fma(tm + cp_m, (T)0.5, ws);
        Wps = fma(cm_p + tp, (T)0.5, ws);
        return;
    }
    const T Z = m.Z(k);
    const T vm = rvs * dvm;
    const T vp = rvs * dvp;
    T cm_m = dsm + vm * Z, cp_m = dsm - vm * Z;
    T cm_p = dsp + vp * Z, cp_p = dsp - vp * Z;
    if (charl) {
        ppm_limit(cm_m, cm_p);
        ppm_limit(cp_m, cp_p);
    }
    // PPM tracing: left-moving waves (lambda<0) on the lo face, right-moving on the hi face
    const T sig = (k == 0 ? m.cp : m.cs) * dtbydx;
    {
        const T mm = cm_m, pp = cm_p;
        cm_m = mm + sig * ((pp - mm) - (pp + mm) * ((T)3 - (T)2 * sig)) / (T)2;
    }
    {
        const T mm = cp_m, pp = cp_p;
        cp_p = pp + sig * ((mm - pp) - (pp + mm) * ((T)3 - (T)2 * sig)) / (T)2;
    }
    const T vmi = m.divZ(rvs * (cm_m - cp_m) / (T)2, k);
    const T vpi = m.divZ(rvs * (cm_p - cp_p) / (T)2, k);
    const T smi = (cm_m + cp_m) / (T)2;
    const T spi = (cm_p + cp_p) / (T)2;
    Wmv = vmi + wv;
    Wpv = vpi + wv;
    Wms = smi + ws;
    Wps = spi + ws;
}

// ---------------------------------------------------------------------------
// normal predictor for one cell, one direction and one side (LinElastPatchGodunov.cpp:471-702)
// w[slot][v|s][5]: primitive stencil along d; outputs W- and W+ (primitive) of that side.
// rvs is the side's velocity sign in this direction (-1 on side 2 in y, else 1: the
// reference's rv, LinElastPhysicsF.ChF:285-301).  A product with rvs = 1 is exact, so side 1
// rounds like the reference's unsigned expressions.  Each slot runs PPM/PLM, characteristic
// analysis, tracing and synthesis to the end before the next starts (fewer live values).
// ---------------------------------------------------------------------------
// PK >= 0 fixes the predictor at compile time: PK_PPM_PRIM is PPM with primitive limiting and no
// characteristic limiting; PK_RUNTIME reads pred and the limiter flags from P.
constexpr int PK_RUNTIME = -1;
constexpr int PK_PPM_PRIM = 2;

template <typename T, int NS, int PK = PK_RUNTIME>
__device__ __forceinline__ void normal_pred_side(const DevParams& P, const Mat<T>& m, int pred_rt, T rvs,
                                                 const T (&w)[NS][2][5], int ci, int dlo, int dhi, T dtbydx,
                                                 T (&Wmv)[NS], T (&Wms)[NS], T (&Wpv)[NS], T (&Wps)[NS])
{
    const int pred = PK >= 0 ? PK : pred_rt;
    if (pred == 0) {  // CTUNormalPred: W+- = W
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            Wmv[k] = Wpv[k] = w[k][0][2];
            Wms[k] = Wps[k] = w[k][1][2];
        }
        return;
    }
    const bool limit = PK >= 0 ? true : (P.char_lim || P.prim_lim);
    const bool charl = PK >= 0 ? false : P.char_lim != 0;
    // CHARSYNTHESISF (LinElastPhysicsF.ChF:526-550), then W+- = increment + W
    auto synth = [&](int k, T cm_m, T cp_m, T cm_p, T cp_p) {
        const T vmi = m.divZ(rvs * (cm_m - cp_m) / (T)2, k);
        const T vpi = m.divZ(rvs * (cm_p - cp_p) / (T)2, k);
        const T smi = (cm_m + cp_m) / (T)2;
        const T spi = (cm_p + cp_p) / (T)2;
        Wmv[k] = vmi + w[k][0][2];
        Wpv[k] = vpi + w[k][0][2];
        Wms[k] = smi + w[k][1][2];
        Wps[k] = spi + w[k][1][2];
    };
    if (pred == 2) {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            T dvm, dvp, dsm, dsp;
            ppm_incr(w[k][0], ci, dlo, dhi, limit, dvm, dvp);
            ppm_incr(w[k][1], ci, dlo, dhi, limit, dsm, dsp);
            ppm_slot<T>(m, k, rvs, dtbydx, limit, charl, w[k][0][2], w[k][1][2], dvm, dvp, dsm, dsp, Wmv[k], Wms[k],
                        Wpv[k], Wps[k]);
        }
    } else {  // PLM
        const bool os = charl || P.fourth;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            T d2[2], ol[2], orr[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const T* ww = w[k][q];
                const int s0 = cell_side(ci, dlo, dhi);
                T d;
                if (P.fourth) {
                    const T vl0 = vl_slope(ww, 2, s0, limit);
                    if (s0 != 0) {
                        d = vl0;
                    } else {
                        const T vlm = vl_slope(ww, 1, cell_side(ci - 1, dlo, dhi), limit);
                        const T vlp = vl_slope(ww, 3, cell_side(ci + 1, dlo, dhi), limit);
                        const T hi = ww[3] - (T)0.25 * vlp;
                        const T lo = ww[1] + (T)0.25 * vlm;
                        d = (T)(2.0 / 3.0) * (hi - lo);
                    }
                } else {
                    d = vl_slope(ww, 2, s0, limit);
                }
                T C, L, R;
                slope_diffs(ww, 2, s0, C, L, R);
                d2[q] = d;
                ol[q] = os ? L : (T)0;
                orr[q] = os ? R : (T)0;
            }
            T dv = d2[0], ds = d2[1];
            const T Z = m.Z(k);
            if (limit && !charl && P.fourth) {
                dv = vl_limit(dv, ol[0], orr[0]);
                ds = vl_limit(ds, ol[1], orr[1]);
            }
            const T vd = rvs * dv;
            T cm_d = ds + vd * Z;
            T cp_d = ds - vd * Z;
            if (charl) {
                const T vl = rvs * ol[0];
                const T vr = rvs * orr[0];
                const T cml = ol[1] + vl * Z, cpl = ol[1] - vl * Z;
                const T cmr = orr[1] + vr * Z, cpr = orr[1] - vr * Z;
                cm_d = vl_limit(cm_d, cml, cmr);
                cp_d = vl_limit(cp_d, cpl, cpr);
            }
            // PLM characteristic predictor, lmin = -cp, lmax = +cp
            const T lk = k == 0 ? m.cp : m.cs;
            const T lmin = -m.cp, lmax = m.cp;
            // wave k-minus (lambda = -lk < 0), wave k-plus (lambda = +lk > 0)
            const T cm_m = -((T)0.5 * ((T)1 + dtbydx * (-lk)) * cm_d);
            const T cm_p = (T)0.5 * ((T)1 - dtbydx * lmax) * cm_d;
            const T cp_m = -((T)0.5 * ((T)1 + dtbydx * lmin) * cp_d);
            const T cp_p = (T)0.5 * ((T)1 - dtbydx * lk) * cp_d;
            synth(k, cm_m, cp_m, cm_p, cp_p);
        }
    }
}

// both sides (side 2 with velocity sign rv)
template <typename T, int NS, int PK = PK_RUNTIME>
__device__ __forceinline__ void normal_pred(const DevParams& P, const Mat<T>& m, int pred_rt, T rv,
                                            const T (&w)[2][NS][2][5], int ci, int dlo, int dhi,
                                            T dtbydx, Slots<T, NS>& Wm, Slots<T, NS>& Wp)
{
    normal_pred_side<T, NS, PK>(P, m, pred_rt, (T)1, w[0], ci, dlo, dhi, dtbydx, Wm.v[0], Wm.s[0], Wp.v[0], Wp.s[0]);
    normal_pred_side<T, NS, PK>(P, m, pred_rt, rv, w[1], ci, dlo, dhi, dtbydx, Wm.v[1], Wm.s[1], Wp.v[1], Wp.s[1]);
}

// outgoing-to-the-right characteristic of a left state (RIEMANNF wxxp, LinElastPhysicsF.ChF:677-679,733-735)
template <typename T, int NS>
__device__ __forceinline__ void right_chars_side(const Mat<T>& m, T rvs, const T (&v)[NS], const T (&s)[NS],
                                                 T (&c)[NS])
{
#pragma unroll
    for (int k = 0; k < NS; ++k) c[k] = s[k] - (rvs * m.Z(k)) * v[k];
}
// left-moving characteristic of a right state (RIEMANNF wxxm, :681-683,738-740)
template <typename T, int NS>
__device__ __forceinline__ void left_chars_side(const Mat<T>& m, T rvs, const T (&v)[NS], const T (&s)[NS],
                                                T (&c)[NS])
{
#pragma unroll
    for (int k = 0; k < NS; ++k) c[k] = s[k] + (rvs * m.Z(k)) * v[k];
}
// RIEMANNF face state from the two characteristics (:693-702, :750-759)
template <typename T, int NS>
__device__ __forceinline__ void riemann_side(const Mat<T>& m, T rvs, const T (&wp)[NS], const T (&wm)[NS],
                                             T (&v)[NS], T (&s)[NS])
{
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        if constexpr (regroup<T>()) {
            v[k] = (wm[k] - wp[k]) * (rvs * m.i2Z(k));
            s[k] = (wm[k] + wp[k]) * (T)0.5;
        } else {
            v[k] = m.divZ(rvs * (wm[k] - wp[k]) / (T)2, k);
            s[k] = (wm[k] + wp[k]) / (T)2;
        }
    }
}

template <typename T, int NS>
__device__ __forceinline__ void right_chars(const Mat<T>& m, T rv, const Slots<T, NS>& L, T (&c)[2][NS])
{
    right_chars_side<T, NS>(m, (T)1, L.v[0], L.s[0], c[0]);
    right_chars_side<T, NS>(m, rv, L.v[1], L.s[1], c[1]);
}
template <typename T, int NS>
__device__ __forceinline__ void left_chars(const Mat<T>& m, T rv, const Slots<T, NS>& R, T (&c)[2][NS])
{
    left_chars_side<T, NS>(m, (T)1, R.v[0], R.s[0], c[0]);
    left_chars_side<T, NS>(m, rv, R.v[1], R.s[1], c[1]);
}
template <typename T, int NS>
__device__ __forceinline__ void riemann_state(const Mat<T>& m, T rv, const T (&wp)[2][NS], const T (&wm)[2][NS],
                                              Slots<T, NS>& F)
{
    riemann_side<T, NS>(m, (T)1, wp[0], wm[0], F.v[0], F.s[0]);
    riemann_side<T, NS>(m, rv, wp[1], wm[1], F.v[1], F.s[1]);
}

// LINELASTOUTBCF / LINELASTFREEBCF (LinElastPhysicsF.ChF:906-1144) on the face state.
// lohisign -1: inside chars are the left-moving wm; +1: right-moving wp.
template <typename T, int NS>
__device__ __forceinline__ void bc_state(const DevParams& P, const Mat<T>& m, int d, T rv, int type, int lohisign,
                                         const T (&cin)[2][NS], Slots<T, NS>& F)
{
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const T Z = m.Z(k);
            T wm, wp;
            const int cv = slot_comp(d, k, 0) + 11 * s, cs_ = slot_comp(d, k, 1) + 11 * s;
            if (type == 0) {
                // absorbing: the incoming characteristic is the background's (zero for U - bg)
                if constexpr (stores_delta<T>()) {
                    if (lohisign == -1) { wm = cin[s][k]; wp = (T)0; }
                    else { wp = cin[s][k]; wm = (T)0; }
                } else {
                    const T bv = (T)P.bg[cv], bs = (T)P.bg[cs_];
                    const T rZ = s ? rv * Z : Z;
                    if (lohisign == -1) { wm = cin[s][k]; wp = bs - rZ * bv; }
                    else { wp = cin[s][k]; wm = bs + rZ * bv; }
                }
            } else {
                // free surface: zero traction, wm + wp = 0 for U; for U - bg the two
                // characteristics of the background add 2 bs
                T off = (T)0;
                if constexpr (stores_delta<T>()) off = (T)(2.0 * P.bg[cs_]);
                if (lohisign == -1) { wm = cin[s][k]; wp = -wm - off; }
                else { wp = cin[s][k]; wm = -wp - off; }
            }
            F.v[s][k] = s ? m.div2Z(rv * (wm - wp), k) : m.div2Z(wm - wp, k);
            F.s[s][k] = (wm + wp) / (T)2;
        }
}

// GETADWDXF applied as W(d1) += A(d2) (WHalf(d2)[+1] - WHalf(d2)[0]) (LinElastPhysicsF.ChF:791-901,
// LinElastPatchGodunov.cpp:408-416) on one side.  sc = -(1/2) dt/dx, times rv(d2) on side 2.
template <typename T, int NS>
__device__ __forceinline__ void transverse_side(const Mat<T>& m, T sc, const T (&Hlov)[NS], const T (&Hlos)[NS],
                                                const T (&Hhiv)[NS], const T (&Hhis)[NS], T (&Wv)[NS], T (&Ws)[NS])
{
    const T dNv = Hhiv[0] - Hlov[0], dNs = Hhis[0] - Hlos[0];
    const T dSv = Hhiv[1] - Hlov[1], dSs = Hhis[1] - Hlos[1];
    if constexpr (regroup<T>()) {
        // sc / rho, sc lam, sc G are uniform: one fma per component
        const T kr = -(sc * m.irho);
        Wv[0] = fma(kr, dSs, Wv[0]);
        Ws[0] = fma(-(sc * m.lam), dNv, Ws[0]);
        Wv[1] = fma(kr, dNs, Wv[1]);
        Ws[1] = fma(-(sc * m.G), dSv, Ws[1]);
        if (NS == 3) Wv[NS - 1] = fma(kr, Hhis[NS - 1] - Hlos[NS - 1], Wv[NS - 1]);
        return;
    }
    Wv[0] = Wv[0] + (-m.divrho(sc * dSs));
    Ws[0] = Ws[0] + (-(sc * dNv * m.lam));
    Wv[1] = Wv[1] + (-m.divrho(sc * dNs));
    Ws[1] = Ws[1] + (-(sc * dSv * m.G));
    if (NS == 3) {
        const T dAs = Hhis[NS - 1] - Hlos[NS - 1];
        Wv[NS - 1] = Wv[NS - 1] + (-m.divrho(sc * dAs));
    }
}

template <typename T, int NS>
__device__ __forceinline__ void transverse_add(const Mat<T>& m, T sc1, T sc2, const Slots<T, NS>& Hlo,
                                               const Slots<T, NS>& Hhi, Slots<T, NS>& W)
{
    transverse_side<T, NS>(m, sc1, Hlo.v[0], Hlo.s[0], Hhi.v[0], Hhi.s[0], W.v[0], W.s[0]);
    transverse_side<T, NS>(m, sc2, Hlo.v[1], Hlo.s[1], Hhi.v[1], Hhi.s[1], W.v[1], W.s[1]);
}

// ---------------------------------------------------------------------------
// fault physics (FP64 always)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double face_val(const double* arr, const DevParams& P, int i, double dflt)
{
    if (!arr) return dflt;
    int k = i - P.face_lo;
    k = k < 0 ? 0 : (k >= P.nface ? P.nface - 1 : k);
    return gmem(arr)[k];
}

__device__ __forceinline__ double gaussianfunc(double w, double r) { return exp(-(0.5 * ((r / w) * (r / w)))); }
__device__ __forceinline__ double smoothfunc(double w, double r)
{
    if (r >= w) return 0.0;
    if (r <= 0.0) return 1.0;
    return exp(r * r / (r * r - w * w));
}

// RSNEWTON (RSIBCF.ChF:336-401): bracketed Newton on the slip rate.  Each lane
// keeps its own iterate; the wave runs until its last lane has converged.
__device__ __forceinline__ double rs_newton(const DevParams& P, double phi, double psi, double sn, double c_rs_a,
                                            double& Vio)
{
    const double atolG = 2.220446049250313080847e-16 * 1e4;
    double V = Vio;
    double Vmin = 0.0, Vmax = phi / P.Zs;
    if (V < Vmin || V > Vmax) V = (Vmin + Vmax) / 2.0;
    const double alpha = sn * c_rs_a;
    const double beta = exp(psi / c_rs_a) / (2.0 * P.rs_V0);
    const double eta = 0.5 * P.Zs;
    double fval = alpha * asinh(beta * V) + eta * V - phi;
    for (int it = 1; it <= 100; ++it) {
        if (fval < 0.0) Vmin = V; else Vmax = V;
        const double dfval = alpha * beta / sqrt(beta * beta * (V * V) + 1.0) + eta;
        const double dV = -(fval / dfval);
        V = V + dV;
        if (V < Vmin || V > Vmax) V = (Vmin + Vmax) / 2.0;
        fval = alpha * asinh(beta * V) + eta * V - phi;
        if (fabs(fval) <= atolG) break;
        if (fabs(dV) <= 0.0) break;
    }
    Vio = V;
    return alpha * asinh(beta * V);
}

// inside state of a fault face (W- of the cell above, physical components)
struct FaultIn {
    double vx[2], vy[2], vz[2], syy[2], sxy[2], syz[2];
};
struct FaultOut {
    double vx[2], vy[2], vz[2], syy[2], sxy[2], syz[2];
};

__device__ __forceinline__ double ramp_of(const DevParams& P, double x)
{
    return fmin(1.0, fmax(0.0, (x - P.ramp_x) / P.ramp_w));
}

// RSFAULTBCF (RSIBCF.ChF:158-323)
__device__ __forceinline__ void rs_fault(const DevParams& P, int i, double psi, const FaultIn& in, FaultOut& o)
{
    const double Zp = P.Zp, Zs = P.Zs;
    const double x = (i + 0.5) * P.dx - P.x0;
    const double wpxy1 = -in.sxy[0] - Zs * in.vx[0];
    const double wpyy1 = -in.syy[0] - Zp * in.vy[0];
    const double wpxy2 = in.sxy[1] - Zs * in.vx[1];
    const double wpyy2 = in.syy[1] - Zp * in.vy[1];
    const double wpyz1 = 0.0, wpyz2 = 0.0;
    double tau_xy_0 = P.nsig * gaussianfunc(P.r0, x);
    if (P.f_tau) tau_xy_0 = tau_xy_0 + face_val(P.f_tau, P, i, 0.0);
    const double tau_yz_0 = 0.0;
    const double phi_xy = 0.5 * (wpxy2 - wpxy1) + tau_xy_0;
    const double phi_yz = 0.5 * (wpyz2 - wpyz1) + tau_yz_0;
    const double phi = sqrt(phi_xy * phi_xy + phi_yz * phi_yz);
    const double syy = -((wpyy1 - wpyy2) / 2.0);
    const double vy = -((wpyy1 + wpyy2) / 2.0 / Zp);
    const double dvx = in.vx[0] - in.vx[1], dvz = in.vz[0] - in.vz[1];
    double V = sqrt(dvx * dvx + dvz * dvz);
    const double c_rs_a = face_val(P.f_a, P, i, P.rs_a) + P.ramp_a * ramp_of(P, x);
    double sn = -syy;
    if (P.f_syy) sn = -(syy + face_val(P.f_syy, P, i, 0.0));
    const double tau = rs_newton(P, phi, psi, sn, c_rs_a, V);
    const double sxy = (phi_xy / phi) * tau - tau_xy_0;
    const double syz = (phi_yz / phi) * tau - tau_yz_0;
    const double Vx = (phi_xy / phi) * V;
    const double Vz = (phi_yz / phi) * V;
    const double vx1 = -((sxy + wpxy1) / Zs);
    const double vz1 = -((syz + wpyz1) / Zs);
    o.vx[0] = vx1; o.vy[0] = vy; o.vz[0] = vz1; o.syy[0] = syy; o.sxy[0] = sxy; o.syz[0] = syz;
    o.vx[1] = vx1 - Vx; o.vy[1] = vy; o.vz[1] = vz1 - Vz; o.syy[1] = syy; o.sxy[1] = sxy; o.syz[1] = syz;
}

// RSSETBND psi update (RSIBCF.ChF:125-146) and the aging-law extension
__device__ __forceinline__ double rs_psi_update(const DevParams& P, int i, double psi, double V, double dt)
{
    const double x = (i + 0.5) * P.dx - P.x0;
    const double c_ramp = ramp_of(P, x);
    const double c_rs_a = face_val(P.f_a, P, i, P.rs_a) + P.ramp_a * c_ramp;
    const double L = face_val(P.f_L, P, i, P.rs_L);
    const double b = face_val(P.f_b, P, i, P.rs_b);
    if (P.state_law == 0) {
        const double f = c_rs_a * asinh(V * exp(psi / c_rs_a) / (2.0 * P.rs_V0));
        double fss = P.rs_f0 - (b - c_rs_a) * log(V / P.rs_V0);
        if (P.rs_fexp > 0.0) {
            const double c_rs_Vw = P.rs_Vw + P.ramp_Vw * c_ramp;
            const double r = V / c_rs_Vw;
            const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
            fss = P.rs_fw + (fss - P.rs_fw) / pow(1.0 + r8, 0.125);
        }
        return psi - dt * (V / L) * (f - fss);
    }
    return psi + dt * (b * P.rs_V0 / L) * (exp((P.rs_f0 - psi) / b) - V / P.rs_V0);
}

// SWFAULTBCF (SWIBCF.ChF:79-267), side 2 at IX2_* = 11+n; slip = fault slip-total
__device__ __forceinline__ void sw_fault(const DevParams& P, int i, double slip, const FaultIn& in, FaultOut& o)
{
    const double Zp = P.Zp, Zs = P.Zs;
    const double x = (i + 0.5) * P.dx - P.fric_c0;
    const double wpxy1 = -in.sxy[0] - Zs * in.vx[0];
    const double wpyy1 = -in.syy[0] - Zp * in.vy[0];
    const double wpyz1 = -in.syz[0] - Zs * in.vz[0];
    const double wpxy2 = in.sxy[1] - Zs * in.vx[1];
    const double wpyy2 = in.syy[1] - Zp * in.vy[1];
    const double wpyz2 = in.syz[1] - Zs * in.vz[1];
    double tau_0 = 0.0;
    for (int ip = 0; ip < P.npatch; ++ip) {
        const double valx = smoothfunc(P.sm, fabs(x - P.xc[ip]) - P.xw[ip]);
        tau_0 = tau_0 + P.taup[ip] * valx;
    }
    if (P.f_tau) tau_0 = tau_0 + face_val(P.f_tau, P, i, 0.0);
    const double ta = 0.5 * (wpxy2 - wpxy1) + tau_0;
    const double tb = 0.5 * (wpyz2 - wpyz1);
    const double tau_lock = sqrt(ta * ta + tb * tb);
    const double syy = -((wpyy1 - wpyy2) / 2.0);
    const double vy = -((wpyy1 + wpyy2) / 2.0 / Zp);
    const double valx = (fabs(x) >= P.fric_w0) ? 0.0 : 1.0;
    const double fs = face_val(P.f_fs, P, i, P.fs);
    const double fd = face_val(P.f_fd, P, i, P.fd);
    const double dw = face_val(P.f_dc, P, i, P.dw);
    double fric = fs * valx + (1.0 - valx) * P.outside_f;
    fric = fric - (fric - fd) * fmin(1.0, slip / dw);
    double sn_syy = syy;
    if (P.f_syy) sn_syy = syy + face_val(P.f_syy, P, i, 0.0);
    const double tau = fmin(tau_lock, -sn_syy * fric) - tau_0;
    const double phi_xy = 0.5 * (wpxy2 - wpxy1);
    const double phi_yz = 0.5 * (wpyz2 - wpyz1);
    const double phi = sqrt(phi_xy * phi_xy + phi_yz * phi_yz);
    const double sxy = (phi_xy / phi) * tau;
    const double syz = (phi_yz / phi) * tau;
    o.vx[0] = -((sxy + wpxy1) / Zs);
    o.vz[0] = -((syz + wpyz1) / Zs);
    o.vy[0] = vy; o.syy[0] = syy; o.sxy[0] = sxy; o.syz[0] = syz;
    o.vx[1] = (sxy - wpxy2) / Zs;
    o.vz[1] = (syz - wpyz2) / Zs;
    o.vy[1] = vy; o.syy[1] = syy; o.sxy[1] = sxy; o.syz[1] = syz;
}

// PLASTICFLOW + INVARIANTS + YIELD (PlasticPhysicsF.ChF:74-181)
// the FP64 Drucker-Prager constants, from the parameters or (plastic_flow's PP = PlasticConstL64)
// held in LDS by a kernel that keeps no copy of them in registers across its loop
enum { PD_B, PD_H, PD_MU, PD_G, PD_K, PD_BETA, PD_ETA, PD_N };
struct PlasticConstL64 {
    const volatile __attribute__((address_space(3))) double* c;
};
__device__ __forceinline__ double plastic_d(const DevParams& P, int k)
{
    return k == PD_B ? P.p_b : k == PD_H ? P.p_h : k == PD_MU ? P.p_mu : k == PD_G ? P.G : k == PD_K ? P.K
         : k == PD_BETA ? P.p_beta : P.p_eta;
}
__device__ __forceinline__ double plastic_d(const PlasticConstL64& P, int k) { return P.c[k]; }

template <typename T, typename PP>
__device__ __forceinline__ void plastic_flow(const PP& P, T dt, T& sxx, T& syy, T& szz, T& sxy, T& sxz, T& syz,
                                             T& plam, T& pgam)
{
    T sigma = (sxx + syy + szz) / (T)3;
    const T a = sxx - syy, b = syy - szz, c = szz - sxx;
    const T J2 = (a * a + b * b + c * c) / (T)6 + sxy * sxy + sxz * sxz + syz * syz;
    T tau = sqrt(J2);
    const T bb = (T)plastic_d(P, PD_B) + pgam * (T)plastic_d(P, PD_H);
    const T Y = tau - fmax(-(T)plastic_d(P, PD_MU) * sigma + bb, (T)0);
    if (Y <= (T)0) { plam = (T)0; return; }
    const T G = (T)plastic_d(P, PD_G), K = (T)plastic_d(P, PD_K);
    plam = Y / (dt * ((T)plastic_d(P, PD_H) + G + (T)plastic_d(P, PD_MU) * (T)plastic_d(P, PD_BETA) * K) +
                (T)plastic_d(P, PD_ETA));
    pgam = pgam + dt * plam;
    const T sig_e = sigma;
    tau = tau - dt * plam * G;
    sigma = sigma - dt * plam * K * (T)plastic_d(P, PD_BETA);
    const T factor = tau / (tau + dt * plam * G);
    sxx = sigma + factor * (sxx - sig_e);
    syy = sigma + factor * (syy - sig_e);
    szz = sigma + factor * (szz - sig_e);
    sxy = factor * sxy;
    sxz = factor * sxz;
    syz = factor * syz;
}

// plastic_flow on the stored difference from the background (FP32 bulk, stores_delta).  With
// s = bg + d, mean stress sigma = sigma_bg + mean(d) and deviator dev = dev_bg + dev(d) are
// formed from the difference and the background's own (pbg) without ever rounding an absolute
// stress to FP32 where it is stored.  The return scales the deviator by factor = 1 - g with
// g = dt lam G / (tau_new + dt lam G) and moves the mean by -dt lam K beta, so the new difference
// is d - dt lam K beta - g dev (diagonal) and d - g s (shear): the same flow rule as plastic_flow
// written as an increment.  A cell that does not yield keeps its values.
// the FP32 Drucker-Prager constants held in LDS (pbg[7], pf[6], then G), for a kernel that keeps
// no copy of them in VGPRs across its loop; volatile, so that every use reads them again
struct PlasticConstL {
    const volatile __attribute__((address_space(3))) float* c;
};
__device__ __forceinline__ float plastic_pbg(const DevParams& P, int k) { return P.pbg[k]; }
__device__ __forceinline__ float plastic_pf(const DevParams& P, int k) { return P.pf[k]; }
__device__ __forceinline__ float plastic_G(const DevParams& P) { return P.fG; }
__device__ __forceinline__ float plastic_pbg(const PlasticConstL& P, int k) { return P.c[k]; }
__device__ __forceinline__ float plastic_pf(const PlasticConstL& P, int k) { return P.c[7 + k]; }
__device__ __forceinline__ float plastic_G(const PlasticConstL& P) { return P.c[13]; }

// PP: DevParams or PlasticConstL
template <typename T, typename PP>
__device__ __forceinline__ void plastic_flow_delta(const PP& P, T dt, T& dxx, T& dyy, T& dzz, T& dxy, T& dxz,
                                                   T& dyz, T& plam, T& pgam)
{
    const T md = (dxx + dyy + dzz) * (T)(1.0 / 3.0);
    const T exx = (dxx - md) + (T)plastic_pbg(P, 1), eyy = (dyy - md) + (T)plastic_pbg(P, 2), ezz = (dzz - md) + (T)plastic_pbg(P, 3);
    const T sxy = dxy + (T)plastic_pbg(P, 4), sxz = dxz + (T)plastic_pbg(P, 5), syz = dyz + (T)plastic_pbg(P, 6);
    const T sigma = md + (T)plastic_pbg(P, 0);
    const T a = exx - eyy, b = eyy - ezz, c = ezz - exx;
    const T J2 = (a * a + b * b + c * c) * (T)(1.0 / 6.0) + sxy * sxy + sxz * sxz + syz * syz;
    const T bb = (T)plastic_pf(P, 1) + pgam * (T)plastic_pf(P, 2);
    const T ys = fmax(-(T)plastic_pf(P, 0) * sigma + bb, (T)0);
    // yield iff sqrt(J2) > ys >= 0, i.e. J2 > ys^2: the root only where the cell yields
    if (!(J2 > ys * ys)) { plam = (T)0; return; }
    const T tau = sqrt(J2);
    const T Y = tau - ys;
    plam = Y / (dt * (T)plastic_pf(P, 3) + (T)plastic_pf(P, 4));
    pgam = pgam + dt * plam;
    const T dG = dt * plam * (T)plastic_G(P);
    const T g = dG / tau;  // 1 - tau_new / (tau_new + dt lam G), tau_new = tau - dt lam G
    const T dm = dt * plam * (T)plastic_pf(P, 5);
    dxx = dxx - (dm + g * exx);
    dyy = dyy - (dm + g * eyy);
    dzz = dzz - (dm + g * ezz);
    dxy = dxy - g * sxy;
    dxz = dxz - g * sxz;
    dyz = dyz - g * syz;
}

}  // namespace tmk
