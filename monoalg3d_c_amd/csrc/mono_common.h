// Shared definitions for the MI355X (gfx950) monodomain kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#define MONO_CHECK(expr)                                                                              \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if(_e != hipSuccess) {                                                                        \
            mono_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
            return -1;                                                                                \
        }                                                                                             \
    } while(0)

// last-error text for the C ABI (mono_last_error)
void mono_set_error(const char *fmt, ...);

namespace mono {

constexpr int kWave = 64;

// Ionic model identifiers (stable; part of the C ABI).
enum Model : int { kTT2006 = 0, kORd2011 = 1, kMS2003 = 2, kCRN1998 = 3 };
constexpr int kNumModels = 4;

// number of state variables per model (ten_tusscher_2006.h:6, Ohara_Rudy_2011.h:7,
// mitchell_shaeffer_2003.h:7, courtemanche_ramirez_nattel_1998.h:6)
__host__ __device__ constexpr int model_neq(int m) {
    return m == kTT2006 ? 19 : (m == kORd2011 ? 41 : (m == kMS2003 ? 2 : 21));
}

// 27-point stencil slot of the neighbour offset (ox, oy, oz); 13 is the cell itself.
__host__ __device__ constexpr int slot_of(int ox, int oy, int oz) { return (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1); }

// Device-resident CG scalars.  Written by block 0 of the kernel that
// completes each reduction; r.z of iteration k lives in rrs[k & 1] so the
// kernel that produces the new value never overwrites the one its other blocks
// are still reading.
struct CgScalars {
    double rrs[2];    // r.z (r.r without Jacobi) by iteration parity
    double rr_true;   // r.r of the current residual (stopping test, reported error)
    double pAp;
    double alpha;
    double beta;
    double rr0;       // r.r at the start of the solve
    double tol_abs;   // stop when rr_true <= tol_abs
    int32_t iters;    // iteration count as the reference counts it (starts at 1)
    int32_t max_its;
    int32_t done;     // 1 once converged or max_its reached
    int32_t done_iter;  // iteration (kernel index k) at which `done` was decided
};

}  // namespace mono
