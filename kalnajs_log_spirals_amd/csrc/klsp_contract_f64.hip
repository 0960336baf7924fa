// klsp_contract_f64.hip -- FP64 instantiation of the M(omega) contraction for
// one omega (complex128 W, FP64 MFMA 16x16x4) and the host launchers of all
// FP64 contractions.  The one-omega kernel uses 4 x 4 block wave tiles: its
// 16 accumulator blocks fill the 256 AGPRs (one wave per SIMD), so each
// K-step loads 8 operand segments for 64 MFMAs.  Built with the default
// (AGPR) accumulator form; the batched kernels (klsp_contract_f64_batch.hip)
// keep theirs in VGPRs.
#include "klsp_contract.hpp"

namespace klsp {

template <int NW> struct F64Shape;
template <> struct F64Shape<1> { static constexpr int MB = 4, NB = 4; };

const void* contract_f64_batch_kernel(int nw, int* MB, int* NB);   // nw = 2, 3

static const void* kernel_for(int nw, int* MB, int* NB)
{
    if (nw <= 1) {
        *MB = F64Shape<1>::MB;
        *NB = F64Shape<1>::NB;
        return (const void*)k_contract<true, F64Shape<1>::MB, F64Shape<1>::NB, 1>;
    }
    return contract_f64_batch_kernel(nw, MB, NB);
}

ContractShape contract_f64_shape(int nw)
{
    static std::atomic<int> occ[4] = {0, 0, 0, 0};
    const int i = nw < 1 ? 1 : (nw > 3 ? 3 : nw);
    int MB = 0, NB = 0;
    const void* k = kernel_for(i, &MB, &NB);
    if (occ[i] == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 4;
        occ[i] = per_cu;
    }
    return {MB, NB, occ[i]};
}

hipError_t contract_f64_launch(int nw, const ContractArgs& a, hipStream_t stream)
{
    const unsigned grid = a.units ? (unsigned)a.n_units : (unsigned)(a.n_gangs * a.wrows * a.wcols);
    ContractArgs args = a;
    void* params[] = {(void*)&args};
    int MB = 0, NB = 0;
    return hipLaunchKernel(kernel_for(nw, &MB, &NB), dim3(grid), dim3(64), params, 0, stream);
}

}  // namespace klsp
