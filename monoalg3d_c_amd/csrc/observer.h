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

}  // namespace mono
