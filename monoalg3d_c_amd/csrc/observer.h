// Activation-time / APD observer (host interface of observer.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mono {

struct Observer {
    float act_thr = -30.0f, apd_thr = -83.0f, time_thr = 10.0f;
    int max_events = 0;  // activation times and APDs kept per cell
    int vm_double = 0;   // 1: keep the last Vm in FP64 (0: float, as the reference)
    int64_t n = 0;
    float *last_v = nullptr, *was_active = nullptr;
    double *last_vd = nullptr;
    int32_t *n_act = nullptr, *n_apd = nullptr;
    float *act = nullptr, *apd = nullptr;  // [max_events][n]

    int init(int64_t n, float act_thr, float apd_thr, float time_thr, int max_events, int vm_double, hipStream_t s);
    int release();
    // one save call at simulation time t on the potential v
    int observe(const double *v, const uint8_t *active, double t, hipStream_t s);
};

// First action potential of every cell for precision studies (no reference
// counterpart): activation time (upward crossing of a threshold, linearly
// interpolated between samples), peak and APD90 (time from activation to the
// downward crossing of V90 = Vmax - 0.9 (Vmax - Vrest) after the peak, also
// interpolated; Vrest = the cell's potential at set-up), sampled on the FP64
// potential at every call.
struct ApdTracker {
    int64_t n = 0;
    double thr = -30.0;
    double *prev = nullptr, *vrest = nullptr, *lat = nullptr, *vmax = nullptr, *apd90 = nullptr;
    double tprev = 0.0;
    int8_t *phase = nullptr;  // 0 resting, 1 activated, 2 repolarised

    int init(int64_t n, double thr, const double *v, double t, hipStream_t s);
    int release();
    int sample(const double *v, const uint8_t *active, double t, hipStream_t s);
};

}  // namespace mono
