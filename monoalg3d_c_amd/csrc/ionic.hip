// Ionic step kernels: one lane per cell, structure-of-arrays state in HBM.
//
// Replaces SOLVE_MODEL_ODES / SET_ODE_INITIAL_CONDITIONS_GPU of the reference
// (ten_tusscher_2006_RS_GPU.cu:8-156, Ohara_Rudy_2011.cu:8-139,
// mitchell_shaeffer_2003.cu:7-128) and fuses UPDATE_MONODOMAIN
// (update_monodomain.c:19-56) into the epilogue: the kernel reads the FP64
// potential v written by the previous diffusion solve (the Vm mirror of
// monodomain_solver.c:1012-1047 becomes a plain read), advances all states by
// num_steps ODE steps, and writes the right-hand side b = alpha * V* for the
// next solve.  The stimulus is evaluated on the fly from the cell centre and a
// per-step amplitude computed on the host (ode_solver.c:288-341).
//
// State layout: sv[s * pitch + cell], s = 0..neq-1, in T (double or float).
// Row 0 (V) is only written when `write_vm_row` is set (the reference's plugin
// contract that row 0 holds Vm after SOLVE_MODEL_ODES).
#include "ionic.h"
#include "ionic_crn1998.h"
#include "ionic_models.h"
#include "mono_common.h"

namespace mono {

template <int MODEL> struct ModelTraits;
template <> struct ModelTraits<kTT2006> { static constexpr int neq = 19; };
template <> struct ModelTraits<kORd2011> { static constexpr int neq = 41; };
template <> struct ModelTraits<kMS2003> { static constexpr int neq = 2; };
template <> struct ModelTraits<kCRN1998> { static constexpr int neq = kCRNNeq; };

// dy = f(y) for the models integrated by (adaptive) forward Euler
template <int MODEL, typename T> __device__ __forceinline__ void model_rhs(const T *y, T *dy, T stim) {
    static_assert(MODEL == kCRN1998, "adaptive Euler runs the models given by their full RHS");
    crn1998_rhs<T>(y, dy, stim);
}

template <int MODEL, typename T>
__device__ __forceinline__ T model_step(T *y, T stim, T dt, const float *scale) {
    if constexpr(MODEL == kTT2006) {
        return tt2006_step<T>(y, stim, dt, T(scale[0]), T(scale[1]), T(scale[2]));
    } else if constexpr(MODEL == kORd2011) {
        return ord2011_step<T>(y, stim, dt);
    } else if constexpr(MODEL == kCRN1998) {
        return crn1998_step<T>(y, stim, dt);
    } else {
        return ms2003_step<T>(y, stim, dt, T(scale[0]));
    }
}

// stimulus for cell i: per-cell array, else (inside the stimulus window) the
// per-cell mask or the x-limit rule on the cell centre
template <typename T> __device__ __forceinline__ double cell_stim(const IonicLaunch<T> &a, int64_t i) {
    double st = 0.0;
    if(a.stim_cell) {
        st = a.stim_cell[i];
    } else if(a.stim_amp != 0.0) {
        if(a.stim_mask) {
            if(a.stim_mask[i]) st = a.stim_amp;
        } else {
            const int ix = (int)(i % a.nx);
            const double cx = ix * a.dx + a.dx / 2.0;
            if(cx < a.stim_x_limit) st = a.stim_amp;
        }
    }
    return st;
}

// The state rows, v and b stream through the step kernel once per step with
// no reuse inside it: they are read and written with the non-temporal hint,
// so they do not displace the solver's vectors from the caches (measured at
// 16M cells: FP32 ORd kernel 1.10 -> 0.93 ms, and the solve that follows
// slightly faster).
template <typename T> __device__ __forceinline__ T stream_load(const T *p) { return __builtin_nontemporal_load(p); }
template <typename T> __device__ __forceinline__ void stream_store(T *p, T v) { __builtin_nontemporal_store(v, p); }

// Writes each new ORd state to its HBM row as soon as the model has it final.
template <typename T> struct RowStore {
    T *sv;
    size_t pitch;
    int64_t i;
    __device__ __forceinline__ void operator()(int s, T val) const { stream_store(&sv[(size_t)s * pitch + i], val); }
};

// ONE: the launch advances one ODE step per PDE step (num_steps == 1, every
// configuration with dt_ode = dt_pde).  Without the loop-carried state array
// the compiler keeps far fewer registers live, and ORd stores each gate right
// after its update: the FP64 ORd kernel drops from 256 VGPRs with 236 B/lane
// of scratch spills to ~180 VGPRs without (4.1 -> 2.6 ms at 16M cells).  The
// FP32 kernels are HBM-bound and keep the loop form: the one-step FP32 ORd
// kernel runs at 6 waves/SIMD and measured 4% slower than the two-wave one.
// The per-instance scaling table of a batch, staged in LDS (dynamic shared
// memory of ninst * 4 floats, given at launch); every thread of the block
// takes part before any returns.
template <int MODEL, typename T> __device__ __forceinline__ void stage_instances(const IonicLaunch<T> &a, float *tab) {
    if constexpr(MODEL == kTT2006 || MODEL == kMS2003) {
        if(a.inst) {
            for(int j = threadIdx.x; j < 4 * a.ninst; j += blockDim.x) tab[j] = a.inst_scale[j];
            __syncthreads();
        }
    }
}

template <int MODEL, typename T, bool ONE, bool ACT>
__device__ __forceinline__ void ionic_step_body(const IonicLaunch<T> &a, const float *tab) {
    constexpr int NEQ = ModelTraits<MODEL>::neq;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= a.n) return;
    // ACT (a launch with inactive cells): every lane steps its cell and
    // writes its rows, so each row store covers whole lines.  (Stores with
    // some lanes masked off made the launch up to 1.6x slower at the same
    // counted traffic: 10% inactive cells, 16M-cell ORd box.)  An inactive
    // cell steps as an unstimulated resting cell: its potential is its state
    // row 0 (the potential v of an inactive cell is 0, outside the model's
    // range), and its right-hand side is 0.
    bool live = true;
    if constexpr(ACT) live = a.active[i] != 0;

    T y[NEQ];
#pragma unroll
    for(int s = 1; s < NEQ; s++) y[s] = stream_load(&a.sv[(size_t)s * a.pitch + i]);
    double V;
    if constexpr(ACT) {
        const double vrow = (double)a.sv[i];
        const double vv = stream_load(&a.v[i]);
        V = (a.first_step || !live) ? vrow : vv;
    } else {
        V = a.first_step ? (double)a.sv[i] : stream_load(&a.v[i]);
    }

    const double st = live ? cell_stim(a, i) : 0.0;
    if constexpr(ONE && MODEL == kORd2011) {
        y[0] = T(V);
        V = V + (double)ord2011_step<T>(y, T(st), a.dt, RowStore<T>{a.sv, a.pitch, i});
        if(a.write_vm_row) stream_store(&a.sv[i], T(V));
        if(a.b) stream_store(&a.b[i], live ? V * a.alpha : 0.0);
        return;
    }
    float scale[4] = {1.f, 1.f, 1.f, 1.f};
    if constexpr(MODEL == kTT2006 || MODEL == kMS2003) {
        if(a.inst) {
            const int id = a.inst[i];
#pragma unroll
            for(int k = 0; k < 4; k++) scale[k] = tab[4 * id + k];
        } else if(a.scale) {
#pragma unroll
            for(int k = 0; k < 4; k++) scale[k] = a.scale[4 * i + k];
        }
    }
    for(int k = 0; k < (ONE ? 1 : a.num_steps); k++) {
        y[0] = T(V);
        const T dv = model_step<MODEL, T>(y, T(st), a.dt, scale);
        V = V + (double)dv;
    }
#pragma unroll
    for(int s = 1; s < NEQ; s++) stream_store(&a.sv[(size_t)s * a.pitch + i], y[s]);
    if(a.write_vm_row) stream_store(&a.sv[i], T(V));
    // fused UPDATE_MONODOMAIN: b = Vm * alpha
    if(a.b) stream_store(&a.b[i], live ? V * a.alpha : 0.0);
}

// Adaptive forward Euler (solve_forward_euler_cpu_adpt, default_solvers.c:61-176)
// for models given by their full RHS.  Each lane advances its cell from its own
// time_new to final_time: Euler steps checked by a second RHS evaluation at the
// new point, error = max_i |dt/2 (k1_i - k2_i)| / max(abs_tol, rel_tol |y_new_i|)
// (+ abs_tol^2), next dt = 0.8 dt / sqrt(error) clamped to [min_dt, max_dt],
// a step rejected while error >= 1 and dt > min_dt, and the reference's time
// bookkeeping (including its t + dt - dt round trip around the second RHS),
// so that the sequence of step sizes is the reference's.  V is carried in
// FP64 (it is the PDE unknown); the other states in T.
template <int MODEL, typename T>
__device__ __forceinline__ void adaptive_body(const IonicLaunch<T> &a) {
    constexpr int NEQ = ModelTraits<MODEL>::neq;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= a.n) return;
    if(a.active && !a.active[i]) return;
    T y[NEQ], k1[NEQ], k2[NEQ], yp[NEQ];
#pragma unroll
    for(int s = 1; s < NEQ; s++) y[s] = a.sv[(size_t)s * a.pitch + i];
    double V = a.first_step ? (double)a.sv[i] : a.v[i];
    const T st = T(cell_stim(a, i));
    double *arow = a.aux;
    double dt = arow[i], tnew = arow[a.pitch + i], pdt;
    const double final_time = a.final_time, min_dt = a.min_dt, max_dt = a.max_dt;
    const double abs_tol = a.abs_tol, rel_tol = a.rel_tol, tiny = pow(abs_tol, 2.0);
    pdt = dt;
    if(tnew + dt > final_time) dt = final_time - tnew;
    bool ok = false;
    y[0] = T(V);
    model_rhs<MODEL, T>(y, k1, st);
    tnew += dt;
    for(int it = 0; it < kMaxAdaptiveSubsteps; it++) {
        // Euler step from the accepted state and the RHS at its end
        const double Vp = (double)k1[0] * dt + V;
#pragma unroll
        for(int s = 1; s < NEQ; s++) yp[s] = k1[s] * T(dt) + y[s];
        yp[0] = T(Vp);
        tnew += dt;
        model_rhs<MODEL, T>(yp, k2, st);
        tnew -= dt;
        double err = 0.0;
#pragma unroll
        for(int s = 0; s < NEQ; s++) {
            const double yn = s == 0 ? Vp : (double)yp[s];
            const double at = fabs(yn) * rel_tol;
            const double tol = (abs_tol > at) ? abs_tol : at;
            const double e = fabs(((dt / 2.0) * ((double)k1[s] - (double)k2[s])) / tol);
            err = (e > err) ? e : err;
        }
        err += tiny;
        pdt = dt;
        dt = 0.8 * dt * sqrt(1.0 / err);
        if(dt < min_dt) dt = min_dt;
        else if(dt > max_dt) dt = max_dt;
        if(tnew + dt > final_time) dt = final_time - tnew;
        if(err >= 1.0 && dt > min_dt) continue;  // rejected: y, V and k1 stay
        V = Vp;
#pragma unroll
        for(int s = 0; s < NEQ; s++) {
            k1[s] = k2[s];
            if(s > 0) y[s] = yp[s];
        }
        if(tnew + pdt >= final_time) {
            if(final_time == tnew) {
                ok = true;
                break;
            } else if(tnew < final_time) {
                dt = pdt = final_time - tnew;
                tnew += pdt;
                ok = true;
                break;
            }
        } else {
            tnew += pdt;
        }
    }
    if(!ok) V = __longlong_as_double(0x7ff8000000000000ll);  // substep cap reached: report NaN
    arow[i] = dt;
    arow[a.pitch + i] = tnew;
    arow[2 * a.pitch + i] = pdt;
#pragma unroll
    for(int s = 1; s < NEQ; s++) a.sv[(size_t)s * a.pitch + i] = y[s];
    if(a.write_vm_row) a.sv[i] = T(V);
    if(a.b) a.b[i] = V * a.alpha;
}

template <int MODEL, typename T>
__global__ __launch_bounds__(256) void k_ionic_adaptive(IonicLaunch<T> a) {
    adaptive_body<MODEL, T>(a);
}

// FP32: the register count the compiler picks.  FP64 models are bound by
// their transcendental and division arithmetic; at two waves per SIMD the
// state loads of one wave overlap the arithmetic of the other (the multi-step
// kernels spill a few registers to scratch, which measures faster than one
// wave per SIMD).
template <int MODEL, bool ACT = false>
__global__ __launch_bounds__(256) void k_ionic_step32(IonicLaunch<float> a) {
    extern __shared__ float inst_tab[];
    stage_instances<MODEL>(a, inst_tab);
    ionic_step_body<MODEL, float, false, ACT>(a, inst_tab);
}
template <int MODEL, bool ONE, bool ACT = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_ionic_step64(IonicLaunch<double> a) {
    extern __shared__ float inst_tab[];
    stage_instances<MODEL>(a, inst_tab);
    ionic_step_body<MODEL, double, ONE, ACT>(a, inst_tab);
}

template <int MODEL, typename T>
__global__ __launch_bounds__(256) void k_ionic_init(T *sv, size_t pitch, double *v, int64_t n, const T *ic) {
    constexpr int NEQ = ModelTraits<MODEL>::neq;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
#pragma unroll
    for(int s = 0; s < NEQ; s++) sv[(size_t)s * pitch + i] = ic[s];
    if(v) v[i] = (double)ic[0];
}

// Default initial conditions (ten_tusscher_2006_RS_CPU.c:28-46, Ohara_Rudy_2011.c:25-65,
// mitchell_shaeffer_2003.c:24-25, courtemanche_ramirez_nattel_1998.c:31-51);
// float literals where the reference has them.
void default_initial_state(int model, double *out) {
    if(model == kTT2006) {
        static const float ic[19] = {-85.23f, 0.00621f, 0.4712f, 0.0095f, 0.00172f, 0.7444f, 0.7045f,
                                     3.373e-5f, 0.7888f, 0.9755f, 0.9953f, 0.999998f, 2.42e-8f, 0.000126f,
                                     3.64f, 0.00036f, 0.9073f, 8.604f, 136.89f};
        for(int s = 0; s < 19; s++) out[s] = (double)ic[s];
    } else if(model == kORd2011) {
        static const double ic[41] = {-85.f, 7, 7, 145, 145, 1.0e-4, 1.0e-4, 1.2, 1.2, 0, 1, 1, 1, 1,
                                      1, 0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 1, 1, 1, 1, 0, 1, 1, 0, 0, 0,
                                      0, 1, 0, 0, 0};
        for(int s = 0; s < 41; s++) out[s] = ic[s];
    } else if(model == kCRN1998) {
        static const float ic[21] = {-8.118000e+01f, 2.908000e-03f, 9.649000e-01f, 9.775000e-01f, 3.043000e-02f,
                                     9.992000e-01f,  4.966000e-03f, 9.986000e-01f, 3.296000e-05f, 1.869000e-02f,
                                     1.367000e-04f,  9.996000e-01f, 7.755000e-01f, 0.0f,          1.000000e+00f,
                                     9.992000e-01f,  1.117000e+01f, 1.390000e+02f, 1.013000e-04f, 1.488000e+00f,
                                     1.488000e+00f};
        for(int s = 0; s < 21; s++) out[s] = (double)ic[s];
    } else {
        out[0] = (double)0.00000820413566106744f;
        out[1] = (double)0.8789655121804799f;
    }
}

// INITIAL_V of the model headers: the PDE potential's initial value
// (ten_tusscher_2006.h:7, Ohara_Rudy_2011.h:8, mitchell_shaeffer_2003.h:8)
double model_initial_v(int model) {
    if(model == kTT2006) return (double)(-85.23f);
    if(model == kORd2011) return (double)(-85.f);
    if(model == kCRN1998) return (double)(-81.180000f);  // courtemanche_ramirez_nattel_1998.h:7
    return 0.00000820413566106744;
}

// FP64 step kernels run one wave per block: the work per cell is long (VALU
// bound), and single-wave blocks fill the SIMDs more evenly than 4-wave ones
// (16M-cell ORd: 2.46 ms at 64 threads, 2.47 at 128, 2.54 at 256, 2.82 at 512).
// The HBM-bound FP32 kernels are fastest at 256 (0.93 ms; 1.02 at 128, 1.24 at 512).
template <bool ONE>
static hipError_t launch_step64(int model, const IonicLaunch<double> &a, hipStream_t s) {
    const int block = a.block > 0 ? a.block : 64;
    const unsigned grid = (unsigned)((a.n + block - 1) / block);
    const size_t lds = a.inst ? sizeof(float) * 4 * (size_t)a.ninst : 0;
    const bool act = a.active != nullptr;
    switch(model) {
    case kTT2006:
        if(act) k_ionic_step64<kTT2006, ONE, true><<<grid, block, lds, s>>>(a);
        else k_ionic_step64<kTT2006, ONE><<<grid, block, lds, s>>>(a);
        break;
    case kORd2011:
        if(act) k_ionic_step64<kORd2011, ONE, true><<<grid, block, 0, s>>>(a);
        else k_ionic_step64<kORd2011, ONE><<<grid, block, 0, s>>>(a);
        break;
    case kMS2003:
        if(act) k_ionic_step64<kMS2003, ONE, true><<<grid, block, lds, s>>>(a);
        else k_ionic_step64<kMS2003, ONE><<<grid, block, lds, s>>>(a);
        break;
    case kCRN1998:
        if(act) k_ionic_step64<kCRN1998, ONE, true><<<grid, block, 0, s>>>(a);
        else k_ionic_step64<kCRN1998, ONE><<<grid, block, 0, s>>>(a);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_step_t(int model, const IonicLaunch<T> &a, hipStream_t s) {
    if(a.n == 0) return hipSuccess;
    if constexpr(sizeof(T) == 8) {
        return a.num_steps == 1 ? launch_step64<true>(model, a, s) : launch_step64<false>(model, a, s);
    } else {
        const int block = a.block > 0 ? a.block : 256;
        const unsigned grid = (unsigned)((a.n + block - 1) / block);
        const size_t lds = a.inst ? sizeof(float) * 4 * (size_t)a.ninst : 0;
        const bool act = a.active != nullptr;
        switch(model) {
        case kTT2006:
            if(act) k_ionic_step32<kTT2006, true><<<grid, block, lds, s>>>(a);
            else k_ionic_step32<kTT2006><<<grid, block, lds, s>>>(a);
            break;
        case kORd2011:
            if(act) k_ionic_step32<kORd2011, true><<<grid, block, 0, s>>>(a);
            else k_ionic_step32<kORd2011><<<grid, block, 0, s>>>(a);
            break;
        case kMS2003:
            if(act) k_ionic_step32<kMS2003, true><<<grid, block, lds, s>>>(a);
            else k_ionic_step32<kMS2003><<<grid, block, lds, s>>>(a);
            break;
        case kCRN1998:
            if(act) k_ionic_step32<kCRN1998, true><<<grid, block, 0, s>>>(a);
            else k_ionic_step32<kCRN1998><<<grid, block, 0, s>>>(a);
            break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
}

template <typename T>
static hipError_t launch_adaptive_t(int model, const IonicLaunch<T> &a, hipStream_t s) {
    const int block = 256;
    const unsigned grid = (unsigned)((a.n + block - 1) / block);
    if(a.n == 0) return hipSuccess;
    if(!a.aux) return hipErrorInvalidValue;
    switch(model) {
    case kCRN1998: k_ionic_adaptive<kCRN1998, T><<<grid, block, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ionic_adaptive(int model, const IonicLaunch<double> &a, hipStream_t s) { return launch_adaptive_t(model, a, s); }
hipError_t launch_ionic_adaptive(int model, const IonicLaunch<float> &a, hipStream_t s) { return launch_adaptive_t(model, a, s); }

hipError_t launch_ionic_step(int model, const IonicLaunch<double> &a, hipStream_t s) { return launch_step_t(model, a, s); }
hipError_t launch_ionic_step(int model, const IonicLaunch<float> &a, hipStream_t s) { return launch_step_t(model, a, s); }

template <typename T>
static hipError_t launch_init_t(int model, T *sv, size_t pitch, double *v, int64_t n, const T *ic_dev, hipStream_t s) {
    const int block = 256;
    const unsigned grid = (unsigned)((n + block - 1) / block);
    if(n == 0) return hipSuccess;
    switch(model) {
    case kTT2006: k_ionic_init<kTT2006, T><<<grid, block, 0, s>>>(sv, pitch, v, n, ic_dev); break;
    case kORd2011: k_ionic_init<kORd2011, T><<<grid, block, 0, s>>>(sv, pitch, v, n, ic_dev); break;
    case kMS2003: k_ionic_init<kMS2003, T><<<grid, block, 0, s>>>(sv, pitch, v, n, ic_dev); break;
    case kCRN1998: k_ionic_init<kCRN1998, T><<<grid, block, 0, s>>>(sv, pitch, v, n, ic_dev); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ionic_init(int model, double *sv, size_t pitch, double *v, int64_t n, const double *ic, hipStream_t s) {
    return launch_init_t(model, sv, pitch, v, n, ic, s);
}
hipError_t launch_ionic_init(int model, float *sv, size_t pitch, double *v, int64_t n, const float *ic, hipStream_t s) {
    return launch_init_t(model, sv, pitch, v, n, ic, s);
}

}  // namespace mono
