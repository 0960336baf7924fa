// klsp_internal.hpp -- declarations shared by the host and device halves of
// libklsp.so (not part of the public ABI).
#pragma once

#include <cstdint>

namespace klsp {

// Godfrey-Lanczos complex Gamma (host), out = {re, im}.
void host_gamma(double re, double im, double* out);

// Largest N_alpha the LU kernel accepts (its row/column caches live in LDS).
constexpr int kMaxN = 1024;

// Device tiling constants of the M(omega) contraction.
constexpr int kTile = 64;       // output tile (beta x alpha) per workgroup
constexpr int kColPad = 16;     // alpha padding of the device W rows
constexpr int kKStep = 4;       // K per MFMA step (16x16x4 f64/f32 MFMA)
constexpr int kRowPad = 32;     // zero rows past Kp in W and c (prefetch slack)
constexpr int kMaxBatch = 3;    // omegas sharing one pass over W
constexpr int kMaxTerms = 4;    // Taylor terms of one BF16x3 pass
constexpr int kMaxW = kMaxTerms > kMaxBatch ? kMaxTerms : kMaxBatch;  // weight sets per pass

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace klsp
