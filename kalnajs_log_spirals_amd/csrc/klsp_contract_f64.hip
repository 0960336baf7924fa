// klsp_contract_f64.hip -- FP64 instantiations of the M(omega) contraction
// (complex128 W, FP64 MFMA 16x16x4) and their host launchers.
#include "klsp_contract.hpp"

namespace klsp {

// Wave tiles (MB x NB blocks) per omega-batch size NW: ~8 accumulator blocks
// per omega keeps the kernel at 2 waves per SIMD for a single omega.
template <int NW> struct F64Shape;
template <> struct F64Shape<1> { static constexpr int MB = 4, NB = 2; };
template <> struct F64Shape<2> { static constexpr int MB = 2, NB = 4; };
template <> struct F64Shape<3> { static constexpr int MB = 2, NB = 2; };

template <int NW>
static const void* kernel_f64()
{
    return (const void*)k_contract<true, F64Shape<NW>::MB, F64Shape<NW>::NB, NW>;
}

static const void* kernel_for(int nw)
{
    return nw == 1 ? kernel_f64<1>() : nw == 2 ? kernel_f64<2>() : kernel_f64<3>();
}

ContractShape contract_f64_shape(int nw)
{
    static int occ[4] = {0, 0, 0, 0};
    const int i = nw < 1 ? 1 : (nw > 3 ? 3 : nw);
    if (occ[i] == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_for(i), 64, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 4;
        occ[i] = per_cu;
    }
    if (i == 1) return {F64Shape<1>::MB, F64Shape<1>::NB, occ[i]};
    if (i == 2) return {F64Shape<2>::MB, F64Shape<2>::NB, occ[i]};
    return {F64Shape<3>::MB, F64Shape<3>::NB, occ[i]};
}

hipError_t contract_f64_launch(int nw, const ContractArgs& a, hipStream_t stream)
{
    const unsigned grid = (unsigned)(a.n_gangs * a.wrows * a.wcols);
    ContractArgs args = a;
    void* params[] = {(void*)&args};
    return hipLaunchKernel(kernel_for(nw), dim3(grid), dim3(64), params, 0, stream);
}

}  // namespace klsp
