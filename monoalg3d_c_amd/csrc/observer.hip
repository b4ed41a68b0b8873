// Activation times and action-potential durations on the device.
//
// Restates save_with_activation_times (save_mesh.c:730-913) per cell, on the
// device, at each save call (every print_rate iterations, before the step):
//   - the first call (t == 0) only records Vm;
//   - an upward crossing of the activation threshold between the previous and
//     the current call records an activation at the current time, unless the
//     cell already has one less than time_threshold ms earlier;
//   - a cell with a pending activation records APD = t - activation time when
//     Vm falls to the APD threshold (or at once for the earlier of two pending
//     activations).
// The reference keeps Vm, counters and times in float (its hash values and
// arrays are float); so does this kernel, except that the previous Vm can be
// kept in FP64 instead (vm_double): with a float copy, a sample that rounds to
// exactly the threshold from below hides that cell's activation (about one
// cell in a million), which differs between two runs that agree to 1e-6 mV.  Times and APDs beyond max_events per
// cell are counted but not stored (their APDs come out as NaN).
#include "mono_common.h"
#include "observer.h"

#include <vector>

namespace mono {

__global__ __launch_bounds__(256) void k_observe(int64_t n, const double *__restrict__ v, const uint8_t *__restrict__ active, double t,
                                                 float act_thr, float apd_thr, float time_thr, int kmax, float *__restrict__ last_v,
                                                 double *__restrict__ last_vd, float *__restrict__ was_active, int32_t *__restrict__ n_act,
                                                 int32_t *__restrict__ n_apd, float *__restrict__ act, float *__restrict__ apd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    if(active && !active[i]) return;
    const double vi = v[i];
    if(t == 0.0) {
        if(last_vd) last_vd[i] = vi;
        else last_v[i] = (float)vi;
        return;
    }
    int na = n_act[i];
    float wa = was_active[i];
    const bool below = last_vd ? last_vd[i] < (double)act_thr : last_v[i] < act_thr;
    if(below && vi >= (double)act_thr) {
        bool rec = na == 0;
        if(!rec) {
            const double last = na - 1 < kmax ? (double)act[(int64_t)(na - 1) * n + i] : (double)NAN;
            rec = !(t - last <= (double)time_thr);  // a lost time does not block a new activation
        }
        if(rec) {
            if(na < kmax) act[(int64_t)na * n + i] = (float)t;
            na++;
            wa += 1.0f;
        }
    }
    if(wa != 0.0f && (vi <= (double)apd_thr || wa == 2.0f)) {
        const int tmp = (int)wa;
        const int k = na - tmp;
        const double last = k >= 0 && k < kmax ? (double)act[(int64_t)k * n + i] : (double)NAN;
        const int np = n_apd[i];
        if(np < kmax) apd[(int64_t)np * n + i] = (float)(t - last);
        n_apd[i] = np + 1;
        wa = (float)(tmp - 1);
    }
    if(last_vd) last_vd[i] = vi;
    else last_v[i] = (float)vi;
    n_act[i] = na;
    was_active[i] = wa;
}

int Observer::init(int64_t n_, float a, float p, float tt, int kmax, int vmd, hipStream_t s) {
    if(release()) return -1;
    n = n_;
    vm_double = vmd;
    act_thr = a;
    apd_thr = p;
    time_thr = tt;
    max_events = kmax;
    const size_t nn = (size_t)(n > 0 ? n : 1), kk = (size_t)(kmax > 0 ? kmax : 1);
    MONO_CHECK(hipMalloc(&last_v, sizeof(float) * nn));
    MONO_CHECK(hipMalloc(&was_active, sizeof(float) * nn));
    MONO_CHECK(hipMalloc(&n_act, sizeof(int32_t) * nn));
    MONO_CHECK(hipMalloc(&n_apd, sizeof(int32_t) * nn));
    MONO_CHECK(hipMalloc(&act, sizeof(float) * nn * kk));
    MONO_CHECK(hipMalloc(&apd, sizeof(float) * nn * kk));
    // reference defaults: last Vm -100, no activations
    const float m100 = -100.0f;
    std::vector<float> h(nn, m100);
    MONO_CHECK(hipMemcpyAsync(last_v, h.data(), sizeof(float) * nn, hipMemcpyHostToDevice, s));
    if(vmd) {
        MONO_CHECK(hipMalloc(&last_vd, sizeof(double) * nn));
        std::vector<double> hd(nn, -100.0);
        MONO_CHECK(hipMemcpyAsync(last_vd, hd.data(), sizeof(double) * nn, hipMemcpyHostToDevice, s));
        MONO_CHECK(hipStreamSynchronize(s));
    }
    MONO_CHECK(hipMemsetAsync(was_active, 0, sizeof(float) * nn, s));
    MONO_CHECK(hipMemsetAsync(n_act, 0, sizeof(int32_t) * nn, s));
    MONO_CHECK(hipMemsetAsync(n_apd, 0, sizeof(int32_t) * nn, s));
    MONO_CHECK(hipMemsetAsync(act, 0xff, sizeof(float) * nn * kk, s));  // NaN
    MONO_CHECK(hipMemsetAsync(apd, 0xff, sizeof(float) * nn * kk, s));
    MONO_CHECK(hipStreamSynchronize(s));
    return 0;
}

int Observer::release() {
    void *bufs[] = {last_v, was_active, n_act, n_apd, act, apd, last_vd};
    for(void *p : bufs)
        if(p) MONO_CHECK(hipFree(p));
    last_v = was_active = act = apd = nullptr;
    last_vd = nullptr;
    n_act = n_apd = nullptr;
    max_events = 0;
    return 0;
}

int Observer::observe(const double *v, const uint8_t *active, double t, hipStream_t s) {
    if(!last_v) {
        mono_set_error("activation observer not set up");
        return -1;
    }
    if(n == 0) return 0;
    k_observe<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, v, active, t, act_thr, apd_thr, time_thr, max_events, last_v, last_vd,
                                                          was_active, n_act, n_apd, act, apd);
    MONO_CHECK(hipGetLastError());
    return 0;
}

// One sample of the APD90 tracker at time t (previous sample at tp).
__global__ __launch_bounds__(256) void k_apd_sample(int64_t n, const double *__restrict__ v, const uint8_t *__restrict__ active,
                                                    double t, double tp, double thr, double *__restrict__ prev,
                                                    const double *__restrict__ vrest, double *__restrict__ lat,
                                                    double *__restrict__ vmax, double *__restrict__ apd90, int8_t *__restrict__ phase) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    if(active && !active[i]) return;
    const double vi = v[i], vp = prev[i];
    int ph = phase[i];
    if(ph == 0 && vp < thr && vi >= thr) {
        lat[i] = tp + (t - tp) * ((thr - vp) / (vi - vp));
        vmax[i] = vi;
        ph = 1;
    } else if(ph == 1) {
        if(vi > vmax[i]) {
            vmax[i] = vi;
        } else {
            const double v90 = vmax[i] - 0.9 * (vmax[i] - vrest[i]);
            if(vp > v90 && vi <= v90) {
                apd90[i] = tp + (t - tp) * ((v90 - vp) / (vi - vp)) - lat[i];
                ph = 2;
            }
        }
    }
    prev[i] = vi;
    phase[i] = (int8_t)ph;
}

__global__ __launch_bounds__(256) void k_apd_init(int64_t n, const double *__restrict__ v, double *__restrict__ prev,
                                                  double *__restrict__ vrest, double *__restrict__ lat, double *__restrict__ vmax,
                                                  double *__restrict__ apd90, int8_t *__restrict__ phase) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    prev[i] = v[i];
    vrest[i] = v[i];
    lat[i] = NAN;
    vmax[i] = NAN;
    apd90[i] = NAN;
    phase[i] = 0;
}

int ApdTracker::init(int64_t n_, double thr_, const double *v, double t, hipStream_t s) {
    if(release()) return -1;
    n = n_;
    thr = thr_;
    tprev = t;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    double **bufs[] = {&prev, &vrest, &lat, &vmax, &apd90};
    for(double **b : bufs) MONO_CHECK(hipMalloc(b, sizeof(double) * nn));
    MONO_CHECK(hipMalloc(&phase, nn));
    if(n > 0) {
        k_apd_init<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, v, prev, vrest, lat, vmax, apd90, phase);
        MONO_CHECK(hipGetLastError());
    }
    MONO_CHECK(hipStreamSynchronize(s));
    return 0;
}

int ApdTracker::release() {
    void *bufs[] = {prev, vrest, lat, vmax, apd90, phase};
    for(void *p : bufs)
        if(p) MONO_CHECK(hipFree(p));
    prev = vrest = lat = vmax = apd90 = nullptr;
    phase = nullptr;
    return 0;
}

int ApdTracker::sample(const double *v, const uint8_t *active, double t, hipStream_t s) {
    if(!prev) {
        mono_set_error("APD90 tracker not set up");
        return -1;
    }
    if(n > 0) {
        k_apd_sample<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, v, active, t, tprev, thr, prev, vrest, lat, vmax, apd90, phase);
        MONO_CHECK(hipGetLastError());
    }
    tprev = t;
    return 0;
}

}  // namespace mono
