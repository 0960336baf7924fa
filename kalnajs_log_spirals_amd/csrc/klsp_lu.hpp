// klsp_lu.hpp -- the step-wise complex LU determinant (klsp_lu.hip), shared
// by the drivers in klsp_device.hip (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace klsp {

// Batches of at most this many matrices take the step-wise LU; larger ones
// fill the device with one workgroup per matrix (k_lu_det).
constexpr int kLuStepMaxMatrices = 24;
// ... of matrices of at most this order (a thread per panel row).
constexpr int kLuStepMaxN = 512;

// Device bytes of per-matrix bookkeeping the step-wise LU needs (row
// interchanges of the last two panels, U's diagonal, swap count, zero-pivot
// flag); the caller provides nd times this in `aux`.
size_t lu_step_aux_bytes();

// det(I - M_b) for b < nd (coef == nullptr), or det(I - sum_j coef[b*nt + j]
// M_j) (a point of a Taylor model), of N x N column-major complex matrices,
// by LU with partial pivoting (LAPACK zgetrf's pivot rule) into det_out[b].
// A holds nd N x N working copies.  One launch per 16-column panel: the
// workgroups of a launch each own one column block, apply the last panel's
// row interchanges and rank-16 update to it, and the owner of the next panel
// factors it.  Enqueued on st; N <= kLuStepMaxN.
hipError_t launch_lu_steps(hipStream_t st, const double2* M, const double2* coef, int nt,
                           double2* A, int N, int nd, void* aux, double2* det_out);

}  // namespace klsp
