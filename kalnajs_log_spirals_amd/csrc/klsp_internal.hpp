// klsp_internal.hpp -- declarations shared by the host and device halves of
// libklsp.so (not part of the public ABI).
#pragma once

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "klsp.h"

namespace klsp {

// Status of the C++ exception being handled (call only inside a catch
// clause): the C entry points end in a function-try-block whose handler
// returns this, so no exception crosses the C ABI.  Host allocation failure
// is KLSP_ERR_OOM, a size beyond the host's limits KLSP_ERR_INVALID_ARG,
// anything else (e.g. a worker thread that could not start) KLSP_ERR_HOST.
// The reason goes to *msg when msg is not null.
inline int exception_status(std::string* msg) noexcept
{
    int code = KLSP_ERR_HOST;
    try {
        try {
            throw;
        } catch (const std::bad_alloc&) {
            code = KLSP_ERR_OOM;
            if (msg) *msg = "klsp: host allocation failed";
        } catch (const std::length_error&) {
            code = KLSP_ERR_INVALID_ARG;
            if (msg) *msg = "klsp: size beyond the host's limits";
        } catch (const std::exception& e) {
            if (msg) *msg = std::string("klsp: ") + e.what();
        } catch (...) {
            if (msg) *msg = "klsp: unknown host failure";
        }
    } catch (...) {
        // the message itself could not be stored; the status still says why
    }
    return code;
}

// Godfrey-Lanczos complex Gamma (host), out = {re, im}.
void host_gamma(double re, double im, double* out);

// Largest N_alpha the LU kernel accepts (its row/column caches live in LDS).
constexpr int kMaxN = 1024;

// Device tiling constants of the M(omega) contraction.
constexpr int kTile = 64;       // output tile (beta x alpha) per workgroup
constexpr int kColPad = 16;     // alpha padding of the device W rows
constexpr int kKStep = 4;       // K per MFMA step (16x16x4 f64/f32 MFMA)
constexpr int kRowPad = 64;     // zero rows past Kp in W and c (prefetch slack)
constexpr int kMaxBatch = 3;    // omegas sharing one pass over W
constexpr int kMaxTerms = 4;    // Taylor terms of one BF16x3 pass
constexpr int kMaxW = kMaxTerms > kMaxBatch ? kMaxTerms : kMaxBatch;  // weight sets per pass

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace klsp
