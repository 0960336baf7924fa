// klsp_contract_f64_batch.hip -- FP64 instantiations of the M(omega)
// contraction for two and three omegas per pass (complex128 W, FP64 MFMA
// 16x16x4).  Wave tiles keep ~8 accumulator blocks per omega within the
// VGPRs at 2 waves per SIMD; built with VGPR-form MFMA accumulators.
#include "klsp_contract.hpp"

namespace klsp {

template <int NW> struct F64BatchShape;
template <> struct F64BatchShape<2> { static constexpr int MB = 2, NB = 4; };
template <> struct F64BatchShape<3> { static constexpr int MB = 2, NB = 2; };

const void* contract_f64_batch_kernel(int nw, int* MB, int* NB)
{
    if (nw == 2) {
        *MB = F64BatchShape<2>::MB;
        *NB = F64BatchShape<2>::NB;
        return (const void*)k_contract<true, F64BatchShape<2>::MB, F64BatchShape<2>::NB, 2>;
    }
    *MB = F64BatchShape<3>::MB;
    *NB = F64BatchShape<3>::NB;
    return (const void*)k_contract<true, F64BatchShape<3>::MB, F64BatchShape<3>::NB, 3>;
}

}  // namespace klsp
