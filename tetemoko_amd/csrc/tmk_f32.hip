// tmk_f32.hip -- the FP32-bulk update kernels, compiled apart from the rest of the library.
//
// TMK_PREC_FP32_BULK is a reduced-precision mode with a stated tolerance (state 1e-5, slip and
// rupture time against the FP64 path), so its kernels are built with FMA contraction
// (-ffp-contract=fast), which removes about a tenth of their vector instructions.  Every FP64
// kernel, whose results must round like the reference, lives in tmk_api.hip, built without
// contraction.  tmk_api.hip declares these instantiations extern and launches them.
#include "tmk_kernels.hpp"
#include "tmk_interior.hpp"
#include "tmk_march.hpp"

namespace tmk {

#define TMK_INSTANTIATE(T, NS, BX, BY, PK) \
    template __global__ void k_godunov_interior<T, NS, BX, BY, PK>(const KArgs<T>);

TMK_ALL_UPDATE_KERNELS(TMK_INSTANTIATE, float)
template __global__ void k_godunov_bnd<float, 2, PK_RUNTIME>(const KArgs<float>);
template __global__ void k_godunov_bnd<float, 3, PK_RUNTIME>(const KArgs<float>);
template __global__ void k_godunov_bnd<float, 2, PK_PPM_PRIM>(const KArgs<float>);
template __global__ void k_godunov_bnd<float, 3, PK_PPM_PRIM>(const KArgs<float>);
template __global__ void k_godunov_march<float, PK_PPM_PRIM, true>(const KArgs<float>);
template __global__ void k_godunov_march<float, PK_PPM_PRIM, false>(const KArgs<float>);

}  // namespace tmk
