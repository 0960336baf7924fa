// klsp_contract_f32.hip -- FP32 instantiations of the M(omega) contraction
// (complex64 W, FP32 MFMA 16x16x4 with FP64 K-tile sums) and their host
// launchers.  Built with -mllvm -amdgpu-mfma-vgpr-form: keeping the FP32 and
// FP64 accumulators in VGPRs fits the kernel in 2 waves per SIMD.
#include "klsp_contract.hpp"

namespace klsp {

// Wave tiles (MB x NB blocks) per omega-batch size NW: ~8 accumulator blocks
// per omega keeps the kernel at 2 waves per SIMD for a single omega.
template <int NW> struct F32Shape;
template <> struct F32Shape<1> { static constexpr int MB = 4, NB = 2; };
template <> struct F32Shape<2> { static constexpr int MB = 2, NB = 2; };
template <> struct F32Shape<3> { static constexpr int MB = 2, NB = 1; };

template <int NW>
static const void* kernel_f32()
{
    return (const void*)k_contract<false, F32Shape<NW>::MB, F32Shape<NW>::NB, NW>;
}

static const void* kernel_for(int nw)
{
    return nw == 1 ? kernel_f32<1>() : nw == 2 ? kernel_f32<2>() : kernel_f32<3>();
}

ContractShape contract_f32_shape(int nw)
{
    static std::atomic<int> occ[4] = {0, 0, 0, 0};
    const int i = nw < 1 ? 1 : (nw > 3 ? 3 : nw);
    if (occ[i] == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_for(i), 64, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 4;
        occ[i] = per_cu;
    }
    if (i == 1) return {F32Shape<1>::MB, F32Shape<1>::NB, occ[i]};
    if (i == 2) return {F32Shape<2>::MB, F32Shape<2>::NB, occ[i]};
    return {F32Shape<3>::MB, F32Shape<3>::NB, occ[i]};
}

hipError_t contract_f32_launch(int nw, const ContractArgs& a, hipStream_t stream)
{
    const unsigned grid = a.units ? (unsigned)a.n_units : (unsigned)(a.n_gangs * a.wrows * a.wcols);
    ContractArgs args = a;
    void* params[] = {(void*)&args};
    return hipLaunchKernel(kernel_for(nw), dim3(grid), dim3(64), params, 0, stream);
}

}  // namespace klsp
