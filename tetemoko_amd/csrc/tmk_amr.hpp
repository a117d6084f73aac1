// tmk_amr.hpp -- device kernels of the multi-level step: coarse-fine ghost fill of
// the state and of the fault data, flux-register increments, reflux, average-down.
//
// Every kernel works from a host-built list (one entry per ghost cell, register
// face, or coarse cell) whose fields are box indices and offsets into the boxes'
// ghosted planes; the plane pointers come from the levels' Patch descriptors at
// the parity the launch needs.  One thread per (entry, component); no atomics on
// data, every sum runs in a fixed order, so the results do not depend on scheduling.
//
// Reference behaviour followed (what each operation computes):
//   coarse-fine fill of U   Chombo PiecewiseLinearFillPatch as used by
//                           LinElastLevelGodunov::step (LinElastLevelGodunov.cpp:226-264);
//                           the same steps as D1PiecewiseLinearFillPatch.cpp:289-590 in 2-D
//   coarse-fine fill of B   D1PiecewiseLinearFillPatch::fillInterp (:289-326), slopes unlimited
//   flux registers          LevelFluxRegister incrementCoarse/incrementFine
//                           (LinElastLevelGodunov.cpp:334-355), reflux (AMRLevelLinElast.cpp:499-500)
//   average-down            CoarseAverage / D1CoarseAverage (AMRLevelLinElast.cpp:503-509)
#pragma once

#include "tmk_kernels.hpp"

#define TMK_AMR_MAXR 8

namespace tmk {

// one fine ghost cell of one box, filled from the coarse level
struct CfCell {
    int dpatch, doff;        // fine box and offset in its ghosted plane
    int cpatch[9], coff[9];  // coarse 3x3 neighbourhood (jj = -1..1 outer, ii = -1..1 inner); -1: not read
    int flags;               // bits 0-1 x slope, bits 2-3 y slope (0 central, 1 one-sided low,
                             // 2 one-sided high), bit 4: limited
    int m;                   // member (scenario) of the hierarchy batch: its clocks
    double coef0, coef1;     // -1/2 + (offset + 1/2) / ref_ratio per direction
};

// one fine ghost fault face of one box
struct CfFace {
    int dpatch, doff;        // fine box, offset in its fault-data row
    int cpatch[3], coff[3];  // coarse faces ic-1, ic, ic+1 (-1: not read)
    int kind;                // 0 central, 1 one-sided low, 2 one-sided high
    int limit;               // 1: limit the slope against the three faces (D1FineInterp)
    int m, pad;              // member of the hierarchy batch
    double coef;
};

// one register face: the uncovered coarse cell, its face flux on the coarse level,
// and the ref_ratio fine faces along it
struct RegFace {
    int cpatch, cface;       // coarse box holding the uncovered cell, face offset in its F[d]
    int d;
    int fpatch;              // fine box whose F[d] holds the fine faces
    int fface[TMK_AMR_MAXR]; // face offsets in that box's F[d]
    int m;                   // member of the hierarchy batch: its dt on both levels
};

// one uncovered coarse cell with register faces e0 .. e1-1 (in the order dir 0 lo, dir 0 hi,
// dir 1 lo, dir 1 hi)
struct RefluxCell {
    int cpatch, coff;
    int e0, e1;
};

// one coarse cell under the fine level, averaged from the r x r fine cells from foff
struct AvgCell {
    int cpatch, coff;
    int fpatch, foff;
};

// one coarse fault face under the fine level
struct AvgFace {
    int cpatch, coff;
    int fpatch, foff;
};

// the time-interpolation weight of a fine step inside the coarse step, from the device
// clocks, as LinElastLevelGodunov::step computes it (LinElastLevelGodunov.cpp:234-258):
// alpha = (t - tCoarseOld) / (tCoarseNew - tCoarseOld), snapped to 0 or 1 within
// 0.04 dt / ref_ratio.  Returns false when alpha lies outside [0, 1].
__device__ __forceinline__ bool amr_alpha(const Scen& fine, const Scen& crs, int r, double& alpha)
{
    const double tn = crs.time;
    const double to = tn - crs.dt;
    double a = (fine.time - to) / (tn - to);
    const double eps = 0.04 * fine.dt / (double)r;
    if (fabs(a) < eps) a = 0.0;
    if (fabs(1.0 - a) < eps) a = 1.0;
    alpha = a;
    return a >= 0.0 && a <= 1.0;
}

// new * alpha + old * (1 - alpha), with the exact copies at 0 and 1
// (D1PiecewiseLinearFillPatch::timeInterp, :328-395)
template <typename T>
__device__ __forceinline__ T time_interp(T o, T n, double alpha)
{
    if (alpha == 1.0) return n;
    if (alpha == 0.0) return o;
    return (T)((double)n * alpha + (double)o * (1.0 - alpha));
}

// comps: the planes to fill (the stencil planes; the others are never read in ghosts)
// alpha_fixed >= 0: that weight (regrid: 1, the coarse level's current data) instead of the
// clocks'
template <typename T>
__global__ void k_amr_fill_state(const CfCell* __restrict__ cells, int ncells, const int* __restrict__ comps,
                                 int ncomp, const Patch<T>* __restrict__ fine, const Patch<T>* __restrict__ cold,
                                 const Patch<T>* __restrict__ cnew, const Scen* fscen, const Scen* cscen, int r,
                                 double alpha_fixed)
{
    // component-major: the lanes of a wave fill neighbouring ghost cells of one plane, whose
    // coarse neighbourhoods overlap
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ncells * ncomp) return;
    const int e = (int)(t % ncells), c = comps[t / ncells];
    const CfCell cc = cells[e];
    double alpha = alpha_fixed;
    if (alpha_fixed < 0.0 && !amr_alpha(fscen[cc.m], cscen[cc.m], r, alpha)) {
        report(&((Scen*)fscen)[cc.m].err, TMK_E_AMR_TIME);
        return;
    }
    T v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        v[k] = (T)0;
        if (cc.cpatch[k] >= 0)
            v[k] = time_interp<T>(gmem(cold[cc.cpatch[k]].in[c])[cc.coff[k]], gmem(cnew[cc.cpatch[k]].in[c])[cc.coff[k]],
                                  alpha);
    }
    const T c0 = v[4];
    T s[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const int kind = (cc.flags >> (2 * d)) & 3;
        const T lo = d == 0 ? v[3] : v[1], hi = d == 0 ? v[5] : v[7];
        s[d] = kind == 0 ? (T)0.5 * (hi - lo) : kind == 1 ? c0 - lo : hi - c0;
    }
    if (cc.flags & 16) {
        T smax = c0, smin = c0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            smax = fmax(smax, v[k]);
            smin = fmin(smin, v[k]);
        }
        const T dsum = (T)0.5 * (fabs(s[0]) + fabs(s[1]));
        if (dsum > (T)0) {
            const T eta = fmin(fmin(smax - c0, c0 - smin) / dsum, (T)1);
            s[0] = eta * s[0];
            s[1] = eta * s[1];
        }
    }
    T val = c0 + (T)cc.coef0 * s[0];
    val = val + (T)cc.coef1 * s[1];
    gmem((T*)fine[cc.dpatch].in[c])[cc.doff] = val;
}

template <typename T>
__global__ void k_amr_fill_fault(const CfFace* __restrict__ faces, int nfaces, int nb, const Patch<T>* __restrict__ fine,
                                 const Patch<T>* __restrict__ cold, const Patch<T>* __restrict__ cnew,
                                 const Scen* fscen, const Scen* cscen, int r, double alpha_fixed)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfaces * nb) return;
    const int e = t / nb, c = t % nb;
    const CfFace f = faces[e];
    if (c >= fine[f.dpatch].bnb) return;  // members of a batch may have different fault laws
    double alpha = alpha_fixed;
    if (alpha_fixed < 0.0 && !amr_alpha(fscen[f.m], cscen[f.m], r, alpha)) return;
    double v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = 0.0;
        if (f.cpatch[k] >= 0) {
            const Patch<T>& po = cold[f.cpatch[k]];
            const Patch<T>& pn = cnew[f.cpatch[k]];
            v[k] = time_interp<double>(gmem(po.bin)[f.coff[k] + (size_t)c * po.bnx], gmem(pn.bin)[f.coff[k] + (size_t)c * pn.bnx],
                                       alpha);
        }
    }
    double s = f.kind == 0 ? 0.5 * (v[2] - v[0]) : f.kind == 1 ? v[1] - v[0] : v[2] - v[1];
    if (f.limit) {
        const double smax = fmax(fmax(v[0], v[1]), v[2]), smin = fmin(fmin(v[0], v[1]), v[2]);
        const double dsum = 0.5 * fabs(s);
        if (dsum > 0.0) s = fmin(fmin(smax - v[1], v[1] - smin) / dsum, 1.0) * s;
    }
    const Patch<T>& pd = fine[f.dpatch];
    gmem((double*)pd.bin)[f.doff + (size_t)c * pd.bnx] = v[1] + f.coef * s;
}

// plane size of the flux array of direction d (faces of the patch's valid box)
template <typename T>
__device__ __forceinline__ size_t flux_plane(const Patch<T>& p, int d)
{
    const int fnx = p.vhi0 - p.vlo0 + 1 + (d == 0), fny = p.vhi1 - p.vlo1 + 1 + (d == 1);
    return (size_t)fnx * fny;
}

// reg = -dt_c F_c (the register is cleared by this write: LevelFluxRegister::setToZero
// then incrementCoarse); dt_c is the member's coarse step
template <typename T>
__global__ void k_amr_reg_coarse(const RegFace* __restrict__ faces, int nfaces, double* __restrict__ reg,
                                 const Patch<T>* __restrict__ crs, const Scen* __restrict__ cscen)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfaces * TMK_NCOMP) return;
    const int c = t / nfaces, e = t % nfaces;
    const RegFace f = faces[e];
    const double dtc = cscen[f.m].dt;
    const Patch<T>& p = crs[f.cpatch];
    const double F = gmem(p.F[f.d])[f.cface + (size_t)c * flux_plane(p, f.d)];
    double acc = 0.0;
    acc = acc - dtc * F;
    reg[(size_t)c * nfaces + e] = acc;
}

// reg += (dt_f / r) F_f over the r fine faces, in order; dt_f is the member's fine step
template <typename T>
__global__ void k_amr_reg_fine(const RegFace* __restrict__ faces, int nfaces, double* __restrict__ reg,
                               const Patch<T>* __restrict__ fine, int r, const Scen* __restrict__ fscen)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfaces * TMK_NCOMP) return;
    const int c = t / nfaces, e = t % nfaces;
    const RegFace f = faces[e];
    const double w = fscen[f.m].dt * (1.0 / (double)r);
    const Patch<T>& p = fine[f.fpatch];
    const size_t fp = flux_plane(p, f.d);
    const __attribute__((address_space(1))) double* F = gmem(p.F[f.d]) + (size_t)c * fp;
    double acc = reg[(size_t)c * nfaces + e];
    for (int m = 0; m < r; ++m) acc = acc + w * F[f.fface[m]];
    reg[(size_t)c * nfaces + e] = acc;
}

// U += s_e reg_e for the register faces of each uncovered coarse cell, in order
template <typename T>
__global__ void k_amr_reflux(const RefluxCell* __restrict__ cells, int ncells, const double* __restrict__ reg,
                             const double* __restrict__ scale, int nfaces, const Patch<T>* __restrict__ crs)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells * TMK_NCOMP) return;
    const int c = t / ncells, k = t % ncells;
    const RefluxCell rc = cells[k];
    __attribute__((address_space(1))) T* u = gmem((T*)crs[rc.cpatch].in[c]) + rc.coff;
    double v = (double)*u;
    for (int e = rc.e0; e < rc.e1; ++e) v = v + scale[e] * reg[(size_t)c * nfaces + e];
    *u = (T)v;
}

// coarse = (sum of the r x r fine cells, i fastest) / r^2, for the components in comps
template <typename T>
__global__ void k_amr_avg_state(const AvgCell* __restrict__ cells, int ncells, const int* __restrict__ comps,
                                int ncomp, const Patch<T>* __restrict__ crs, const Patch<T>* __restrict__ fine, int r,
                                double scale)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells * ncomp) return;
    const int c = comps[t / ncells], k = t % ncells;
    const AvgCell a = cells[k];
    const Patch<T>& pf = fine[a.fpatch];
    const __attribute__((address_space(1))) T* f = gmem(pf.in[c]) + a.foff;
    double acc = 0.0;
    for (int jj = 0; jj < r; ++jj)
        for (int ii = 0; ii < r; ++ii) acc = acc + (double)f[ii + (size_t)jj * pf.unx];
    gmem((T*)crs[a.cpatch].in[c])[a.coff] = (T)(acc * scale);
}

// coarse fault face = (sum of the r fine faces) / r (D1AVERAGE)
template <typename T>
__global__ void k_amr_avg_fault(const AvgFace* __restrict__ faces, int nfaces, int nb, const Patch<T>* __restrict__ crs,
                                const Patch<T>* __restrict__ fine, int r, double scale)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfaces * nb) return;
    const int c = t / nfaces, k = t % nfaces;
    const AvgFace a = faces[k];
    const Patch<T>& pf = fine[a.fpatch];
    const Patch<T>& pc = crs[a.cpatch];
    if (c >= pc.bnb) return;  // members of a batch may have different fault laws
    const __attribute__((address_space(1))) double* f = gmem(pf.bin) + a.foff + (size_t)c * pf.bnx;
    double acc = 0.0;
    for (int ii = 0; ii < r; ++ii) acc = acc + f[ii];
    gmem((double*)pc.bin)[a.coff + (size_t)c * pc.bnx] = acc * scale;
}

}  // namespace tmk
