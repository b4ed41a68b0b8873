// Per-cell ionic models as device functions, templated on the arithmetic type
// T (double for the FP64 path, float for the FP32 path).
//
// Each model advances one cell by one ODE step of size dt.  The membrane
// potential is handled separately from the other states: the model returns the
// forward-Euler increment dV (already multiplied by dt, in T), and the caller
// adds it to the FP64 potential, so the FP32 path accumulates V in FP64.
// Literal constants keep the reference's float/double rounding (a literal that
// carries an `f` suffix in the reference is float-rounded there even in the
// FP64 build), so the FP64 path reproduces the reference's model constants.
//
// Scaling factors (`scale`) multiply selected conductances; all 1 reproduces
// the reference model exactly.
#pragma once

#include <hip/hip_runtime.h>

namespace mono {

// Every model function is host-and-device, so the same source also runs on the
// host (tests compare it with the oracle there, without a GPU).
#define MONO_HD __host__ __device__ __forceinline__

// exp for the FP64 path: Cody & Waite's argument reduction (W. J. Cody and
// W. Waite, Software Manual for the Elementary Functions, 1980) x = n ln2 + r,
// |r| <= ln2/2, with ln2 split into its nearest double and the remainder
// (both FMA-applied), exp(r) by a degree-11 polynomial: the Taylor series
// through r^13 economized to degree 11 with Chebyshev polynomials on
// [-0.34658, 0.34658] (Lanczos' economization, exact rational arithmetic,
// coefficients then rounded; truncation below 4e-18 relative), scaled by 2^n.
// Within 1 ulp of a correctly rounded exp for every finite argument the
// models produce; no special cases: a non-finite argument gives NaN (a state
// that has left the model's range is reported downstream as a NaN
// potential).  The library exp spends a fifth of its instructions on those
// range checks.
MONO_HD double exp_cw(double x) {
    const double n = rint(x * 0x1.71547652b82fep+0);  // log2(e)
    double r = fma(-n, 0x1.62e42fefa39efp-1, x);     // ln2, nearest double
    r = fma(-n, 0x1.abc9e3b39803fp-56, r);           // ln2 - that double
    double p = 0x1.af780f134c082p-26;
    p = fma(p, r, 0x1.28b3c5e683c14p-22);
    p = fma(p, r, 0x1.71dde78948070p-19);
    p = fma(p, r, 0x1.a01991d94f3c5p-16);
    p = fma(p, r, 0x1.a01a01b7f87b9p-13);
    p = fma(p, r, 0x1.6c16c187efa50p-10);
    p = fma(p, r, 0x1.111111110db8cp-7);
    p = fma(p, r, 0x1.555555554f0fcp-5);
    p = fma(p, r, 0x1.5555555555562p-3);
    p = fma(p, r, 0x1.0000000000011p-1);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

// FP64 path: exp_cw above, the device's libm (OCML) for log/pow/expm1, so
// results agree with the reference's glibc to rounding (divisions, compiled as
// a Newton-refined reciprocal, are within 1 ulp).  FP32 path: the hardware
// approximations (v_exp_f32 / v_log_f32 based, a few ulp), which halve the
// kernel's instruction count; its accuracy is bounded by the activation-time
// tests, not by bitwise agreement.  (Host compiles of the FP32 path use the
// host libm.)
#ifndef MONO_FP32_FAST_MATH
#define MONO_FP32_FAST_MATH 1
#endif
#if MONO_FP32_FAST_MATH && defined(__HIP_DEVICE_COMPILE__)
#define MONO_F32(fast, lib) fast
#else
#define MONO_F32(fast, lib) lib
#endif
template <typename T> MONO_HD T dexp(T x);
template <> MONO_HD double dexp<double>(double x) { return exp_cw(x); }
template <> MONO_HD float dexp<float>(float x) { return MONO_F32(__expf, expf)(x); }
template <typename T> MONO_HD T dexpm1(T x);
template <> MONO_HD double dexpm1<double>(double x) { return expm1(x); }
template <> MONO_HD float dexpm1<float>(float x) { return expm1f(x); }
template <typename T> MONO_HD T dlog(T x);
template <> MONO_HD double dlog<double>(double x) { return log(x); }
template <> MONO_HD float dlog<float>(float x) { return MONO_F32(__logf, logf)(x); }
template <typename T> MONO_HD T dpow(T x, T y);
template <> MONO_HD double dpow<double>(double x, double y) { return pow(x, y); }
template <> MONO_HD float dpow<float>(float x, float y) { return MONO_F32(__powf, powf)(x, y); }
template <typename T> MONO_HD T dsqrt(T x);
template <> MONO_HD double dsqrt<double>(double x) { return sqrt(x); }
template <> MONO_HD float dsqrt<float>(float x) { return sqrtf(x); }

// x / (exp(x) - 1), finite at x = 0 (limit 1).  The FP32 path evaluates the
// Goldman-Hodgkin-Katz flux factors with it: written as the reference writes
// them, v F^2/RT * (...) / (exp(k v F/RT) - 1) is 0/0 whenever a cell's Vm
// rounds to within ~1e-6 mV of 0 in float (millions of plateau cells cross 0 mV),
// while in FP64 that window is ~1e-15 mV.
template <typename T> MONO_HD T x_over_expm1(T x) { return x == T(0) ? T(1) : x / dexpm1<T>(x); }

// Rush-Larsen update toward the steady state: inf + (x - inf) * exp(-dt / tau)
template <typename T> MONO_HD T rl_toward(T inf, T x, T dt, T tau) {
    return inf + (x - inf) * dexp<T>(-dt / tau);
}
// the same update written as inf - (inf - x) * exp(-dt / tau) (ORd form)
template <typename T> MONO_HD T rl_from(T inf, T x, T dt, T tau) {
    return inf - (inf - x) * dexp<T>(-dt / tau);
}
// the ORd form with the decay factor exp(-dt / tau) given: gates that share a
// time constant share one exponential, and a time constant written 1 / D
// gives the factor as exp(-dt * D) with no division
template <typename T> MONO_HD T rl_decay(T inf, T x, T decay) { return inf - (inf - x) * decay; }

// Per-instance conductance scalings (batching / sweeps).  Identity = all 1.
struct IonicScale {
    float g[4];  // TT2006: GKr, GCaL, GKs, unused; MS: tau_close, unused...
};

// ---------------------------------------------------------------------------
// ten Tusscher & Panfilov 2006, epicardial cell
// (models_library/ten_tusscher/ten_tusscher_2006_RS_CPU.c:137-351; the step of
// :100-134: Rush-Larsen for states 1..12, forward Euler for the rest).
// y[0] (V) is read-only here; the V increment is returned.
// ---------------------------------------------------------------------------
template <typename T>
MONO_HD T tt2006_step(T *y, T stim, T dt, T sKr, T sCaL, T sKs) {
    const T V = y[0], Xr1 = y[1], Xr2 = y[2], Xs = y[3], m = y[4], h = y[5], j = y[6];
    const T d = y[7], f = y[8], f2 = y[9], fCass = y[10], s = y[11], r = y[12];
    const T Ca_i = y[13], Ca_SR = y[14], Ca_ss = y[15], Rp = y[16], Na_i = y[17], K_i = y[18];

    const T R = T(8314.472), Tk = T(310.0), F = T(96485.3415f), Cm = T(0.185), V_c = T(0.016404);
    const T Ko = T(5.4), Nao = T(140.0), Cao = T(2.0);
    const T P_kna = T(0.03), K_mk = T(1.0), P_NaK = T(2.724), K_mNa = T(40.0), K_pCa = T(0.0005);
    const T V_rel = T(0.102), k1_prime = T(0.15), max_sr = T(2.5), min_sr = T(1.0), EC = T(1.5);
    const T Vmax_up = T(0.006375);
    const T n_alpha = T(2.5), n_gamma = T(0.35), K_sat = T(0.1), Km_Ca = T(1.38), Km_Nai = T(87.5);
    const T K_NaCa = T(1000.0);
    const T g_to = T(0.294), g_Kr = T(0.153) * sKr, g_Ks = T(0.098) * sKs, g_CaL = T(3.98e-05) * sCaL;
    const T g_Na = T(14.838), g_pK = T(0.0146), g_bca = T(0.000592), g_pCa = T(0.1238);
    const T g_K1 = T(5.405), g_bna = T(0.00029);
    const T RT = R * Tk;

    const T EK = (RT / F) * dlog<T>(Ko / K_i);
    const T EKs = (RT / F) * dlog<T>((Ko + (P_kna * Nao)) / (K_i + (P_kna * Na_i)));
    const T ENa = (RT / F) * dlog<T>(Nao / Na_i);
    const T ECa = ((T(0.5f) * R * Tk) / F) * dlog<T>(Cao / Ca_i);

    const T VmEK = V - EK;
    const T bK1 = ((T(3.0f) * dexp<T>(T(0.0002f) * (VmEK + T(100.0f)))) + dexp<T>(T(0.1f) * (VmEK - T(10.0f)))) /
                  (T(1.0f) + dexp<T>(T(-0.5f) * VmEK));
    const T aK1 = T(0.1f) / (T(1.0f) + dexp<T>(T(0.06f) * (VmEK - T(200.0f))));
    const T xK1 = aK1 / (aK1 + bK1);
    const T IK1 = g_K1 * xK1 * VmEK;
    const T Ito = g_to * r * s * VmEK;
    const T IKr = g_Kr * Xr1 * Xr2 * VmEK * dsqrt<T>(Ko / T(5.4f));
    const T IKs = g_Ks * (Xs * Xs) * (V - EKs);
    const T IpK = (g_pK * VmEK) / (T(1.0f) + dexp<T>((T(25.0f) - V) / T(5.98f)));

    T ICaL;
    const T V15 = V - T(15.0f);
    if(V < T(15.0f - 1.0e-5f) || V > T(15.0f + 1.0e-5f)) {
        const T z = (T(2.0f) * V15 * F) / RT;
        ICaL = (((g_CaL * d * f * f2 * fCass * T(4.0f) * V15 * (F * F)) / RT) * ((T(0.25f) * Ca_ss * dexp<T>(z)) - Cao)) /
               dexpm1<T>(z);
    } else {
        ICaL = g_CaL * d * f * f2 * fCass * T(2.0f) * F * (T(0.25f) * Ca_ss - Cao);
    }
    const T IbCa = g_bca * (V - ECa);
    const T IpCa = (g_pCa * Ca_i) / (Ca_i + K_pCa);

    const T INaK = ((((P_NaK * Ko) / (Ko + K_mk)) * Na_i) / (Na_i + K_mNa)) /
                   (T(1.0f) + (T(0.1245f) * dexp<T>((T(-0.1f) * V * F) / RT)) + (T(0.0353f) * dexp<T>(((-V) * F) / RT)));
    const T INa = g_Na * (m * m * m) * h * j * (V - ENa);
    const T IbNa = g_bna * (V - ENa);
    const T eg1 = dexp<T>(((n_gamma - T(1.0f)) * V * F) / RT);
    const T INaCa = (K_NaCa * ((dexp<T>((n_gamma * V * F) / RT) * (Na_i * Na_i * Na_i) * Cao) -
                               (eg1 * (Nao * Nao * Nao) * Ca_i * n_alpha))) /
                    (((Km_Nai * Km_Nai * Km_Nai) + (Nao * Nao * Nao)) * (Km_Ca + Cao) * (T(1.0f) + (K_sat * eg1)));

    const T xr1_inf = T(1.0f) / (T(1.0f) + dexp<T>((T(-26.0f) - V) / T(7.0f)));
    const T tau_xr1 = (T(450.0f) / (T(1.0f) + dexp<T>((T(-45.0f) - V) / T(10.0f)))) *
                      (T(6.0f) / (T(1.0f) + dexp<T>((V + T(30.0f)) / T(11.5f))));
    const T xr2_inf = T(1.0f) / (T(1.0f) + dexp<T>((V + T(88.0f)) / T(24.0f)));
    const T tau_xr2 = (T(3.0f) / (T(1.0f) + dexp<T>((T(-60.0f) - V) / T(20.0f)))) *
                      (T(1.12f) / (T(1.0f) + dexp<T>((V - T(60.0f)) / T(20.0f))));
    const T xs_inf = T(1.0f) / (T(1.0f) + dexp<T>((T(-5.0f) - V) / T(14.0f)));
    const T tau_xs = ((T(1400.0f) / dsqrt<T>(T(1.0f) + dexp<T>((T(5.0f) - V) / T(6.0f)))) *
                      (T(1.0f) / (T(1.0f) + dexp<T>((V - T(35.0f)) / T(15.0f))))) + T(80.0f);

    const T em = T(1.0f) + dexp<T>((T(-56.86f) - V) / T(9.03f));
    const T m_inf = T(1.0f) / (em * em);
    const T a_m = T(1.0f) / (T(1.0f) + dexp<T>((T(-60.0f) - V) / T(5.0f)));
    const T b_m = (T(0.1f) / (T(1.0f) + dexp<T>((V + T(35.0f)) / T(5.0f)))) + (T(0.1f) / (T(1.0f) + dexp<T>((V - T(50.0f)) / T(200.0f))));
    const T tau_m = a_m * b_m;

    const T eh = T(1.0f) + dexp<T>((V + T(71.55f)) / T(7.43f));
    const T h_inf = T(1.0f) / (eh * eh);
    const bool low = V < T(-40.0f);
    T a_h, b_h, a_j, b_j;
    if(low) {
        a_h = T(0.057f) * dexp<T>((-(V + T(80.0f))) / T(6.8f));
        b_h = (T(2.7f) * dexp<T>(T(0.079f) * V)) + (T(310000.0f) * dexp<T>(T(0.3485f) * V));
        a_j = (((T(-25428.0f) * dexp<T>(T(0.2444f) * V)) - (T(6.948e-06f) * dexp<T>(T(-0.04391f) * V))) * (V + T(37.78f))) /
              (T(1.0f) + dexp<T>(T(0.311f) * (V + T(79.23f))));
        b_j = (T(0.02424f) * dexp<T>(T(-0.01052f) * V)) / (T(1.0f) + dexp<T>(T(-0.1378f) * (V + T(40.14f))));
    } else {
        a_h = T(0.0f);
        b_h = T(0.77f) / (T(0.13f) * (T(1.0f) + dexp<T>((V + T(10.66f)) / T(-11.1f))));
        a_j = T(0.0f);
        b_j = (T(0.6f) * dexp<T>(T(0.057f) * V)) / (T(1.0f) + dexp<T>(T(-0.1f) * (V + T(32.0f))));
    }
    const T tau_h = T(1.0f) / (a_h + b_h);
    const T j_inf = h_inf;
    const T tau_j = T(1.0f) / (a_j + b_j);

    const T d_inf = T(1.0f) / (T(1.0f) + dexp<T>((T(-8.0f) - V) / T(7.5f)));
    const T a_d = (T(1.4f) / (T(1.0f) + dexp<T>((T(-35.0f) - V) / T(13.0f)))) + T(0.25f);
    const T b_d = T(1.4f) / (T(1.0f) + dexp<T>((V + T(5.0f)) / T(5.0f)));
    const T g_d = T(1.0f) / (T(1.0f) + dexp<T>((T(50.0f) - V) / T(20.0f)));
    const T tau_d = (a_d * b_d) + g_d;

    const T e30 = dexp<T>((V + T(30.0f)) / T(10.0f));
    const T v27 = V + T(27.0f);
    const T f_inf = T(1.0f) / (T(1.0f) + dexp<T>((V + T(20.0f)) / T(7.0f)));
    const T tau_f = (T(1102.5f) * dexp<T>((-(v27 * v27)) / T(225.0f))) + (T(200.0f) / (T(1.0f) + dexp<T>((T(13.0f) - V) / T(10.0f)))) +
                    (T(180.0f) / (T(1.0f) + e30)) + T(20.0f);
    const T f2_inf = (T(0.67f) / (T(1.0f) + dexp<T>((V + T(35.0f)) / T(7.0f)))) + T(0.33f);
    const T tau_f2 = (T(562.0f) * dexp<T>((-(v27 * v27)) / T(240.0f))) + (T(31.0f) / (T(1.0f) + dexp<T>((T(25.0f) - V) / T(10.0f)))) +
                     (T(80.0f) / (T(1.0f) + e30));

    const T cq = Ca_ss / T(0.05f);
    const T fCass_inf = (T(0.6f) / (T(1.0f) + cq * cq)) + T(0.4f);
    const T tau_fCass = (T(80.0f) / (T(1.0f) + cq * cq)) + T(2.0f);

    const T v45 = V + T(45.0f);
    const T s_inf = T(1.0f) / (T(1.0f) + dexp<T>((V + T(20.0f)) / T(5.0f)));
    const T tau_s = (T(85.0f) * dexp<T>((-(v45 * v45)) / T(320.0f))) + (T(5.0f) / (T(1.0f) + dexp<T>((V - T(20.0f)) / T(5.0f)))) + T(3.0f);
    const T v40 = V + T(40.0f);
    const T r_inf = T(1.0f) / (T(1.0f) + dexp<T>((T(20.0f) - V) / T(6.0f)));
    const T tau_r = (T(9.5f) * dexp<T>((-(v40 * v40)) / T(1800.0f))) + T(0.8f);

    const T ecs = EC / Ca_SR;
    const T kcasr = max_sr - ((max_sr - min_sr) / (T(1.0f) + ecs * ecs));
    const T k1 = k1_prime / kcasr;
    const T k3 = T(0.06);
    const T css2 = Ca_ss * Ca_ss;
    const T O = (k1 * css2 * Rp) / (k3 + (k1 * css2));
    const T Irel = V_rel * O * (Ca_SR - Ca_ss);
    const T K_up = T(0.00025);
    const T Iup = Vmax_up / (T(1.0f) + ((K_up * K_up) / (Ca_i * Ca_i)));
    const T Ileak = T(0.00036f) * (Ca_SR - Ca_i);
    const T Ixfer = T(0.0038f) * (Ca_ss - Ca_i);
    const T k2 = T(0.045f) * kcasr;
    const T k4 = T(0.005f);
    const T cb = Ca_i + T(0.001f), sb = Ca_SR + T(0.3f), ssb = Ca_ss + T(0.00025f);
    const T bufc = T(1.0f) / (T(1.0f) + ((T(0.2f) * T(0.001f)) / (cb * cb)));
    const T bufsr = T(1.0f) / (T(1.0f) + ((T(10.0f) * T(0.3f)) / (sb * sb)));
    const T bufss = T(1.0f) / (T(1.0f) + ((T(0.4f) * T(0.00025f)) / (ssb * ssb)));
    const T V_sr = T(0.001094f), V_ss = T(5.468e-05f);

    const T dV = -(IK1 + Ito + IKr + IKs + ICaL + INaK + INa + IbNa + INaCa + IbCa + IpK + IpCa + stim);
    const T dRp = ((-k2) * Ca_ss * Rp) + (k4 * (T(1.0f) - Rp));
    const T dCai = bufc * (((((Ileak - Iup) * V_sr) / V_c) + Ixfer) - ((((IbCa + IpCa) - (T(2.0f) * INaCa)) * Cm) / (T(2.0f) * V_c * F)));
    const T dCaSR = bufsr * (Iup - (Irel + Ileak));
    const T dCass = bufss * (((((-ICaL) * Cm) / (T(2.0f) * V_ss * F)) + ((Irel * V_sr) / V_ss)) - ((Ixfer * V_c) / V_ss));
    const T dNai = ((-(INa + IbNa + (T(3.0f) * INaK) + (T(3.0f) * INaCa))) / (V_c * F)) * Cm;
    const T dKi = ((-((IK1 + Ito + IKr + IKs + IpK + stim) - (T(2.0f) * INaK))) / (V_c * F)) * Cm;

    y[1] = rl_toward<T>(xr1_inf, Xr1, dt, tau_xr1);
    y[2] = rl_toward<T>(xr2_inf, Xr2, dt, tau_xr2);
    y[3] = rl_toward<T>(xs_inf, Xs, dt, tau_xs);
    y[4] = rl_toward<T>(m_inf, m, dt, tau_m);
    y[5] = rl_toward<T>(h_inf, h, dt, tau_h);
    y[6] = rl_toward<T>(j_inf, j, dt, tau_j);
    y[7] = rl_toward<T>(d_inf, d, dt, tau_d);
    y[8] = rl_toward<T>(f_inf, f, dt, tau_f);
    y[9] = rl_toward<T>(f2_inf, f2, dt, tau_f2);
    y[10] = rl_toward<T>(fCass_inf, fCass, dt, tau_fCass);
    y[11] = rl_toward<T>(s_inf, s, dt, tau_s);
    y[12] = rl_toward<T>(r_inf, r, dt, tau_r);
    y[13] = dt * dCai + Ca_i;
    y[14] = dt * dCaSR + Ca_SR;
    y[15] = dt * dCass + Ca_ss;
    y[16] = dt * dRp + Rp;
    y[17] = dt * dNai + Na_i;
    y[18] = dt * dKi + K_i;
    return dt * dV;
}

// ---------------------------------------------------------------------------
// Mitchell-Schaeffer 2003 (mitchell_shaeffer_2003.c:56-93): forward Euler on
// (V, h); the stimulus is added.  scale multiplies tau_close.
// ---------------------------------------------------------------------------
template <typename T>
MONO_HD T ms2003_step(T *y, T stim, T dt, T s_tau_close) {
    const T V = y[0], h = y[1];
    const T tau_in = T(0.3), tau_out = T(6.0), V_gate = T(0.13), tau_open = T(120.0);
    const T tau_close = T(150.0) * s_tau_close;
    const T J_in = (h * ((V * V) * (T(1.0) - V))) / tau_in;
    const T J_out = -(V / tau_out);
    const T dV = J_out + J_in + stim;
    const T dh = (V < V_gate ? (T(1.0) - h) / tau_open : -h / tau_close);
    y[1] = dt * dh + h;
    return dt * dV;
}

}  // namespace mono

namespace mono {

template <typename T> MONO_HD T cube(T x) { return x * x * x; }

// State sink of a model step: st(s, value) is called once per state s >= 1 as
// soon as its new value is final.  A kernel that passes a sink writing state
// rows straight to HBM stores each gate right after its update, so the gate is
// dead for the rest of the step instead of held in registers until the end.
struct NullStore {
    template <typename T> MONO_HD void operator()(int, T) const {}
};

// ---------------------------------------------------------------------------
// O'Hara-Rudy 2011, endocardial cell (Ohara_Rudy_2011.c:110-691, celltype 0).
// In-place and in statement order: the gates advance by Rush-Larsen from V_n,
// the currents use the updated gates, V by forward Euler (:570), then CaMK,
// release fluxes and concentrations.  y[0] (V) is read-only; returns dV.
// pow(x, n) with integer n is evaluated by products (<= 2 ulp from pow).
// The new states are written to y[1..40] and also passed to `st`.
// ---------------------------------------------------------------------------
template <typename T, typename St = NullStore>
MONO_HD T ord2011_step(T *y, T stim, T dt, St st = St()) {
    const T v = y[0];
    T nai = y[1], nass = y[2], ki = y[3], kss = y[4], cai = y[5], cass = y[6];
    T cansr = y[7], cajsr = y[8], m = y[9], hf = y[10], hs = y[11], j = y[12], hsp = y[13];
    T jp = y[14], mL = y[15], hL = y[16], hLp = y[17], a = y[18], iF = y[19], iS = y[20];
    T ap = y[21], iFp = y[22], iSp = y[23], d = y[24], ff = y[25], fs = y[26], fcaf = y[27];
    T fcas = y[28], jca = y[29], nca = y[30], ffp = y[31], fcafp = y[32], xrf = y[33];
    T xrs = y[34], xs1 = y[35], xs2 = y[36], xk1 = y[37], Jrelnp = y[38], Jrelp = y[39];
    T CaMKt = y[40];

    const T nao = T(140.0), cao = T(1.8), ko = T(5.4);
    const T BSRmax = T(0.047), KmBSR = T(0.00087), BSLmax = T(1.124), KmBSL = T(0.0087);
    const T cmdnmax = T(0.05), kmcmdn = T(0.00238), trpnmax = T(0.07), kmtrpn = T(0.0005);
    const T csqnmax = T(10.0f), kmcsqn = T(0.8);
    const T aCaMK = T(0.05), bCaMK = T(0.00068), CaMKo = T(0.05), KmCaM = T(0.0015), KmCaMK = T(0.15);
    const T R = T(8314.0), Tk = T(310.0), F = T(96485.0);
    const T L = T(0.01), rad = T(0.0011);
    // float*float literal products are float-rounded as in the reference (:183-190)
    const T vcell = T(1000.0f * 3.14f) * rad * rad * L;
    const T Ageo = T(2.0f * 3.14f) * rad * rad + T(2.0f * 3.14f) * rad * L;
    const T Acap = T(2) * Ageo;
    const T vmyo = T(0.68f) * vcell;
    const T vnsr = T(0.0552f) * vcell;
    const T vjsr = T(0.0048f) * vcell;
    const T vss = T(0.02f) * vcell;
    const T RTF = R * Tk / F;
    const T RT = R * Tk;

    const T ENa = RTF * dlog<T>(nao / nai);
    const T EK = RTF * dlog<T>(ko / ki);
    const T EKs = RTF * dlog<T>((ko + T(0.01833f) * nao) / (ki + T(0.01833f) * nai));
    T CaMKb = CaMKo * (T(1.0f) - CaMKt) / (T(1.0f) + KmCaM / cass);
    T CaMKa = CaMKb + CaMKt;
    const T vffrt = v * F * F / RT;
    const T vfrt = v * F / RT;
    const T fCaMK = (T(1.0f) / (T(1.0f) + KmCaMK / CaMKa));  // fINap, fINaLp, fItop, fICaLp

    // INa
    // Time constants written 1/D enter the Rush-Larsen factor as exp(-dt D)
    // and those written c + 1/D as exp(-dt D / (c D + 1)), one division
    // instead of two; gates with the same time constant share the factor, and
    // exp((v + a) / k) of a second offset a' is the first times exp((a' - a) / k)
    // (a constant).  Each differs from the reference's expression by a few
    // ulp of the exponent or the result.
    const T mss = T(1.0f) / (T(1.0f) + dexp<T>((-(v + T(39.57f))) / T(9.871f)));
    const T Dm = T(6.765f) * dexp<T>((v + T(11.64f)) / T(34.77f)) + T(8.552f) * dexp<T>(-(v + T(77.42f)) / T(5.955f));
    const T Em = dexp<T>(-dt * Dm);  // m and mL share tm = 1 / Dm
    m = rl_decay<T>(mss, m, Em);
    st(9, m);
    const T eh = dexp<T>((v + T(82.90f)) / T(6.086f));
    const T hss = T(1.0f) / (T(1) + eh);
    const T Dhf = T(1.432e-5f) * dexp<T>(-(v + T(1.196f)) / T(6.285f)) + T(6.149f) * dexp<T>((v + T(0.5096f)) / T(20.27f));
    const T Dhs = T(0.009794f) * dexp<T>(-(v + T(17.95f)) / T(28.05f)) + T(0.3343f) * dexp<T>((v + T(5.730f)) / T(56.66f));
    const T Ahf = T(0.99);
    const T Ahs = T(1.0f) - Ahf;
    hf = rl_decay<T>(hss, hf, dexp<T>(-dt * Dhf));
    st(10, hf);
    hs = rl_decay<T>(hss, hs, dexp<T>(-dt * Dhs));
    st(11, hs);
    const T h = Ahf * hf + Ahs * hs;
    const T jss = hss;
    const T Dj = T(0.02136f) * dexp<T>(-(v + T(100.6f)) / T(8.281f)) + T(0.3052f) * dexp<T>((v + T(0.9941f)) / T(38.45f));
    const T rtj = Dj / (T(2.038f) * Dj + T(1.0f));  // 1 / tj, tj = 2.038 + 1 / Dj
    j = rl_decay<T>(jss, j, dexp<T>(-dt * rtj));
    st(12, j);
    const T hssp = T(1.0f) / (T(1.0f) + eh * dexp<T>((T(89.1f) - T(82.90f)) / T(6.086f)));
    hsp = rl_decay<T>(hssp, hsp, dexp<T>(-dt * Dhs / T(3.0f)));  // thsp = 3 ths
    st(13, hsp);
    const T hp = Ahf * hf + Ahs * hsp;
    jp = rl_decay<T>(jss, jp, dexp<T>(-dt * rtj / T(1.46f)));  // tjp = 1.46 tj
    st(14, jp);
    const T GNa = T(75);
    const T INa = GNa * (v - ENa) * m * m * m * ((T(1.0f) - fCaMK) * h * j + fCaMK * hp * jp);

    // INaL
    const T mLss = T(1.0f) / (T(1.0f) + dexp<T>((-(v + T(42.85f))) / T(5.264f)));
    mL = rl_decay<T>(mLss, mL, Em);
    st(15, mL);
    const T ehL = dexp<T>((v + T(87.61f)) / T(7.488f));
    const T hLss = T(1.0f) / (T(1.0f) + ehL);
    const T thL = T(200.0);
    hL = rl_from<T>(hLss, hL, dt, thL);
    st(16, hL);
    const T hLssp = T(1.0f) / (T(1.0f) + ehL * dexp<T>((T(93.81f) - T(87.61f)) / T(7.488f)));
    const T thLp = T(3.0f) * thL;
    hLp = rl_from<T>(hLssp, hLp, dt, thLp);
    st(17, hLp);
    const T GNaL = T(0.0075);
    const T INaL = GNaL * (v - ENa) * mL * ((T(1.0f) - fCaMK) * hL + fCaMK * hLp);

    // Ito
    const T ea = dexp<T>((-(v - T(14.34f))) / T(14.82f));
    const T ass = T(1.0f) / (T(1.0f) + ea);
    // ta = 1.0515 / Da: a and ap share the factor
    const T Da = T(1.0f) / (T(1.2089f) * (T(1.0f) + dexp<T>(-(v - T(18.4099f)) / T(29.3814f)))) +
                 T(3.5f) / (T(1.0f) + dexp<T>((v + T(100.0f)) / T(29.3814f)));
    const T Ea = dexp<T>(-dt * Da / T(1.0515f));
    a = rl_decay<T>(ass, a, Ea);
    st(18, a);
    const T iss = T(1.0f) / (T(1.0f) + dexp<T>((v + T(43.94f)) / T(5.711f)));
    const T DiF = T(0.3933f) * dexp<T>((-(v + T(100.0f))) / T(100.0f)) + T(0.08004f) * dexp<T>((v + T(50.0f)) / T(16.59f));
    const T DiS = T(0.001416f) * dexp<T>((-(v + T(96.52f))) / T(59.05f)) + T(1.780e-8f) * dexp<T>((v + T(114.1f)) / T(8.079f));
    const T rtiF = DiF / (T(4.562f) * DiF + T(1.0f));   // 1 / tiF
    const T rtiS = DiS / (T(23.62f) * DiS + T(1.0f));   // 1 / tiS
    const T AiF = T(1.0f) / (T(1.0f) + dexp<T>((v - T(213.6f)) / T(151.2f)));
    const T AiS = T(1.0f) - AiF;
    iF = rl_decay<T>(iss, iF, dexp<T>(-dt * rtiF));
    st(19, iF);
    iS = rl_decay<T>(iss, iS, dexp<T>(-dt * rtiS));
    st(20, iS);
    const T ii = AiF * iF + AiS * iS;
    const T assp = T(1.0f) / (T(1.0f) + ea * dexp<T>((T(24.34f) - T(14.34f)) / T(14.82f)));
    ap = rl_decay<T>(assp, ap, Ea);
    st(21, ap);
    const T dti_develop = T(1.354f) + T(1.0e-4f) / (dexp<T>((v - T(167.4f)) / T(15.89f)) + dexp<T>(-(v - T(12.23f)) / T(0.2154f)));
    const T e20 = dexp<T>((v + T(70.0f)) / T(20.0f));
    const T dti_recover = T(1.0f) - T(0.5f) / (T(1.0f) + e20);
    // tiFp = dti tiF, tiSp = dti tiS with dti = dti_develop * dti_recover
    const T rdti = T(1.0f) / (dti_develop * dti_recover);
    iFp = rl_decay<T>(iss, iFp, dexp<T>(-dt * rtiF * rdti));
    st(22, iFp);
    iSp = rl_decay<T>(iss, iSp, dexp<T>(-dt * rtiS * rdti));
    st(23, iSp);
    const T ip = AiF * iFp + AiS * iSp;
    const T Gto = T(0.02);
    const T Ito = Gto * (v - EK) * ((T(1.0f) - fCaMK) * a * ii + fCaMK * ap * ip);

    // ICaL, ICaNa, ICaK
    const T dss = T(1.0f) / (T(1.0f) + dexp<T>((-(v + T(3.940f))) / T(4.230f)));
    const T Dd = dexp<T>(T(-0.05f) * (v + T(6.0f))) + dexp<T>(T(0.09f) * (v + T(14.0f)));
    d = rl_decay<T>(dss, d, dexp<T>(-dt * (Dd / (T(0.6f) * Dd + T(1.0f)))));  // td = 0.6 + 1 / Dd
    st(24, d);
    const T fss = T(1.0f) / (T(1.0f) + dexp<T>((v + T(19.58f)) / T(3.696f)));
    // exp((v + 20) / 10) and exp(-(v + 20) / 10) are reciprocals; Afcaf's
    // exp((v - 10) / 10) is the first times exp(-3)
    const T e10 = dexp<T>((v + T(20.0f)) / T(10.0f));
    const T Dff = T(0.0045f) * (T(1.0f) / e10) + T(0.0045f) * e10;
    const T Dfs = T(0.000035f) * dexp<T>(-(v + T(5.0f)) / T(4.0f)) + T(0.000035f) * dexp<T>((v + T(5.0f)) / T(6.0f));
    const T rtff = Dff / (T(7.0f) * Dff + T(1.0f));       // 1 / tff
    const T Aff = T(0.6);
    const T Afs = T(1.0f) - Aff;
    ff = rl_decay<T>(fss, ff, dexp<T>(-dt * rtff));
    st(25, ff);
    fs = rl_decay<T>(fss, fs, dexp<T>(-dt * (Dfs / (T(1000.0f) * Dfs + T(1.0f)))));
    st(26, fs);
    const T f = Aff * ff + Afs * fs;
    const T fcass = fss;
    const T e7 = dexp<T>((v - T(4.0f)) / T(7.0f));
    const T Dfcaf = T(0.04f) * (T(1.0f) / e7) + T(0.04f) * e7;
    const T Dfcas = T(0.00012f) * dexp<T>(-v / T(3.0f)) + T(0.00012f) * dexp<T>(v / T(7.0f));
    const T rtfcaf = Dfcaf / (T(7.0f) * Dfcaf + T(1.0f));  // 1 / tfcaf
    const T Afcaf = T(0.3f) + T(0.6f) / (T(1.0f) + e10 * dexp<T>((T(-10.0f) - T(20.0f)) / T(10.0f)));
    const T Afcas = T(1.0f) - Afcaf;
    fcaf = rl_decay<T>(fcass, fcaf, dexp<T>(-dt * rtfcaf));
    st(27, fcaf);
    fcas = rl_decay<T>(fcass, fcas, dexp<T>(-dt * (Dfcas / (T(100.0f) * Dfcas + T(1.0f)))));
    st(28, fcas);
    const T fca = Afcaf * fcaf + Afcas * fcas;
    const T tjca = T(75.0);
    jca = rl_from<T>(fcass, jca, dt, tjca);
    st(29, jca);
    ffp = rl_decay<T>(fss, ffp, dexp<T>(-dt * rtff / T(2.5f)));  // tffp = 2.5 tff
    st(31, ffp);
    const T fp = Aff * ffp + Afs * fs;
    fcafp = rl_decay<T>(fcass, fcafp, dexp<T>(-dt * rtfcaf / T(2.5f)));  // tfcafp = 2.5 tfcaf
    st(32, fcafp);
    const T fcap = Afcaf * fcafp + Afcas * fcas;
    const T Kmn = T(0.002);
    const T k2n = T(1000.0);
    const T km2n = jca * T(1.0f);
    const T kq = T(1.0f) + Kmn / cass;
    const T kq2 = kq * kq;
    const T anca = T(1.0f) / (k2n / km2n + kq2 * kq2);
    nca = anca * k2n / km2n - (anca * k2n / km2n - nca) * dexp<T>(-km2n * dt);
    st(30, nca);
    const T e1 = dexp<T>(vfrt);
    const T e2 = e1 * e1;  // exp(2 vfrt)
    T PhiCaL, PhiCaNa, PhiCaK;
    if constexpr(sizeof(T) == 8) {
        PhiCaL = T(4.0f) * vffrt * (cass * e2 - T(0.341f) * cao) / (e2 - T(1.0f));
        PhiCaNa = T(1.0f) * vffrt * (T(0.75f) * nass * e1 - T(0.75f) * nao) / (e1 - T(1.0f));
        PhiCaK = T(1.0f) * vffrt * (T(0.75f) * kss * e1 - T(0.75f) * ko) / (e1 - T(1.0f));
    } else {
        // 4 vffrt / (e2 - 1) = 2 F * (2 vfrt) / expm1(2 vfrt), likewise for e1
        const T g2 = T(2.0f) * F * x_over_expm1<T>(T(2.0f) * vfrt), g1 = F * x_over_expm1<T>(vfrt);
        PhiCaL = g2 * (cass * e2 - T(0.341f) * cao);
        PhiCaNa = g1 * (T(0.75f) * nass * e1 - T(0.75f) * nao);
        PhiCaK = g1 * (T(0.75f) * kss * e1 - T(0.75f) * ko);
    }
    const T zca = T(2.0f);
    const T PCa = T(0.0001);
    const T PCap = T(1.1f) * PCa;
    const T PCaNa = T(0.00125f) * PCa;
    const T PCaK = T(3.574e-4f) * PCa;
    const T PCaNap = T(0.00125f) * PCap;
    const T PCaKp = T(3.574e-4f) * PCap;
    const T gate_np = (f * (T(1.0f) - nca) + jca * fca * nca);
    const T gate_p = (fp * (T(1.0f) - nca) + jca * fcap * nca);
    const T ICaL = (T(1.0f) - fCaMK) * PCa * PhiCaL * d * gate_np + fCaMK * PCap * PhiCaL * d * gate_p;
    const T ICaNa = (T(1.0f) - fCaMK) * PCaNa * PhiCaNa * d * gate_np + fCaMK * PCaNap * PhiCaNa * d * gate_p;
    const T ICaK = (T(1.0f) - fCaMK) * PCaK * PhiCaK * d * gate_np + fCaMK * PCaKp * PhiCaK * d * gate_p;

    // IKr
    const T xrss = T(1.0f) / (T(1.0f) + dexp<T>((-(v + T(8.337f))) / T(6.789f)));
    const T Dxrf = T(0.3652f) * dexp<T>((v - T(31.66f)) / T(3.869f)) + T(4.123e-5f) * dexp<T>((-(v - T(47.78f))) / T(20.38f));
    const T Dxrs = T(0.06629f) * dexp<T>((v - T(34.70f)) / T(7.355f)) + T(1.128e-5f) * dexp<T>((-(v - T(29.74f))) / T(25.94f));
    const T Axrf = T(1.0f) / (T(1.0f) + dexp<T>((v + T(54.81f)) / T(38.21f)));
    const T Axrs = T(1.0f) - Axrf;
    xrf = rl_decay<T>(xrss, xrf, dexp<T>(-dt * (Dxrf / (T(12.98f) * Dxrf + T(1.0f)))));
    st(33, xrf);
    xrs = rl_decay<T>(xrss, xrs, dexp<T>(-dt * (Dxrs / (T(1.865f) * Dxrs + T(1.0f)))));
    st(34, xrs);
    const T xr = Axrf * xrf + Axrs * xrs;
    const T rkr = T(1.0f) / ((T(1.0f) + dexp<T>((v + T(55.0f)) / T(75.0f))) * (T(1.0f) + dexp<T>((v - T(10.0f)) / T(30.0f))));
    const T GKr = T(0.046);
    const T IKr = GKr * dsqrt<T>(ko / T(5.4f)) * xr * rkr * (v - EK);

    // IKs
    const T xs1ss = T(1.0f) / (T(1.0f) + dexp<T>((-(v + T(11.60f))) / T(8.932f)));
    const T Dxs1 = T(2.326e-4f) * dexp<T>((v + T(48.28f)) / T(17.80f)) + T(0.001292f) * dexp<T>((-(v + T(210.0f))) / T(230.0f));
    xs1 = rl_decay<T>(xs1ss, xs1, dexp<T>(-dt * (Dxs1 / (T(817.3f) * Dxs1 + T(1.0f)))));  // txs1 = 817.3 + 1 / Dxs1
    st(35, xs1);
    const T xs2ss = xs1ss;
    // txs2 = 1 / Dxs2; exp((v - 50) / 20) is dti_recover's exp((v + 70) / 20) times exp(-6)
    const T Dxs2 = T(0.01f) * (e20 * dexp<T>((T(-50.0f) - T(70.0f)) / T(20.0f))) + T(0.0193f) * dexp<T>((-(v + T(66.54f))) / T(31.0f));
    xs2 = rl_decay<T>(xs2ss, xs2, dexp<T>(-dt * Dxs2));
    st(36, xs2);
    const T KsCa = T(1.0f) + T(0.6f) / (T(1.0f) + dpow<T>(T(3.8e-5f) / cai, T(1.4f)));
    const T GKs = T(0.0034);
    const T IKs = GKs * KsCa * xs1 * xs2 * (v - EKs);

    // IK1
    const T xk1ss = T(1.0f) / (T(1.0f) + dexp<T>(-(v + T(2.5538f) * ko + T(144.59f)) / (T(1.5692f) * ko + T(3.8115f))));
    const T Dxk1 = dexp<T>((-(v + T(127.2f))) / T(20.36f)) + dexp<T>((v + T(236.8f)) / T(69.33f));
    xk1 = rl_decay<T>(xk1ss, xk1, dexp<T>(-dt * Dxk1 / T(122.2f)));  // txk1 = 122.2 / Dxk1
    st(37, xk1);
    const T rk1 = T(1.0f) / (T(1.0f) + dexp<T>((v + T(105.8f) - T(2.6f) * ko) / T(9.493f)));
    const T GK1 = T(0.1908);
    const T IK1 = GK1 * dsqrt<T>(ko) * rk1 * xk1 * (v - EK);

    // INaCa (bulk and subspace share everything except [Na] and [Ca])
    const T kna1 = T(15.0), kna2 = T(5.0), kna3 = T(88.12), kasymm = T(12.5), wna = T(6.0e4), wca = T(6.0e4);
    const T wnaca = T(5.0e3), kcaon = T(1.5e6), kcaoff = T(5.0e3), qna = T(0.5224), qca = T(0.1670);
    const T hca = dexp<T>((qca * v * F) / RT);
    const T hna = dexp<T>((qna * v * F) / RT);
    const T zna = T(1.0f);
    const T Gncx = T(0.0008);
    const T h7 = T(1.0f) + nao / kna3 * (T(1.0f) + T(1.0f) / hna);
    const T h8 = nao / (kna3 * hna * h7);
    const T h9 = T(1.0f) / h7;
    const T h10 = kasymm + T(1.0f) + nao / kna1 * (T(1.0f) + nao / kna2);
    const T h11 = nao * nao / (h10 * kna1 * kna2);
    const T h12 = T(1.0f) / h10;
    const T k1 = h12 * cao * kcaon;
    const T k2 = kcaoff;
    const T k3p = h9 * wca;
    const T k3pp = h8 * wnaca;
    const T k3 = k3p + k3pp;
    const T k5 = kcaoff;
    const T k8 = h8 * h11 * wna;
    const T KmCaAct = T(150.0e-6);
    T INaCa_part[2];
#pragma unroll
    for(int sub = 0; sub < 2; sub++) {
        const T na = sub == 0 ? nai : nass;
        const T ca = sub == 0 ? cai : cass;
        const T h1 = T(1) + na / kna3 * (T(1) + hna);
        const T h2 = (na * hna) / (kna3 * h1);
        const T h3 = T(1.0f) / h1;
        const T h4 = T(1.0f) + na / kna1 * (T(1) + na / kna2);
        const T h5 = na * na / (h4 * kna1 * kna2);
        const T h6 = T(1.0f) / h4;
        const T k4p = h3 * wca / hca;
        const T k4pp = h2 * wnaca;
        const T k4 = k4p + k4pp;
        const T k6 = h6 * ca * kcaon;
        const T k7 = h5 * h2 * wna;
        const T x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
        const T x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
        const T x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
        const T x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
        const T rxs = T(1.0f) / (x1 + x2 + x3 + x4);
        const T E1 = x1 * rxs, E2 = x2 * rxs, E3 = x3 * rxs, E4 = x4 * rxs;
        const T cr = KmCaAct / ca;
        const T allo = T(1.0f) / (T(1.0f) + cr * cr);
        const T JncxNa = T(3.0f) * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
        const T JncxCa = E2 * k2 - E1 * k1;
        INaCa_part[sub] = (sub == 0 ? T(0.8f) : T(0.2f)) * Gncx * allo * (zna * JncxNa + zca * JncxCa);
    }
    const T INaCa_i = INaCa_part[0];
    const T INaCa_ss = INaCa_part[1];
    const T INaCa = INaCa_i + INaCa_ss;

    // INaK
    const T k1p = T(949.5), k1m = T(182.4), k2p = T(687.2), k2m = T(39.4), k3p_ = T(1899.0), k3m = T(79300.0);
    const T k4p_ = T(639.0), k4m = T(40.0), Knai0 = T(9.073), Knao0 = T(27.78), delta = T(-0.1550f);
    const T Knai = Knai0 * dexp<T>((delta * v * F) / (T(3.0f) * R * Tk));
    const T Knao = Knao0 * dexp<T>(((T(1.0f) - delta) * v * F) / (T(3.0f) * R * Tk));
    const T Kki = T(0.5), Kko = T(0.3582), MgADP = T(0.05), MgATP = T(9.8), Kmgatp = T(1.698e-7), H = T(1.0e-7);
    const T eP = T(4.2), Khp = T(1.698e-7), Knap = T(224.0), Kxkur = T(292.0);
    const T P = eP / (T(1.0f) + H / Khp + nai / Knap + ki / Kxkur);
    const T nn = T(1.0f) + nai / Knai, kk = T(1.0f) + ki / Kki, on = T(1.0f) + nao / Knao, ok = T(1.0f) + ko / Kko;
    const T den_i = cube<T>(nn) + kk * kk - T(1.0f);
    const T den_o = cube<T>(on) + ok * ok - T(1.0f);
    const T a1 = (k1p * cube<T>(nai / Knai)) / den_i;
    const T b1 = k1m * MgADP;
    const T a2 = k2p;
    const T b2 = (k2m * cube<T>(nao / Knao)) / den_o;
    const T kq_ = ko / Kko;
    const T a3 = (k3p_ * (kq_ * kq_)) / den_o;
    const T b3 = (k3m * P * H) / (T(1.0f) + MgATP / Kmgatp);
    const T a4 = (k4p_ * MgATP / Kmgatp) / (T(1.0f) + MgATP / Kmgatp);
    const T kiq = ki / Kki;
    const T b4 = (k4m * (kiq * kiq)) / den_i;
    const T x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
    const T x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
    const T x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
    const T x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
    const T rxs = T(1.0f) / (x1 + x2 + x3 + x4);
    const T E1 = x1 * rxs, E2 = x2 * rxs, E3 = x3 * rxs, E4 = x4 * rxs;
    const T zk = T(1.0f);
    const T JnakNa = T(3.0f) * (E1 * a3 - E2 * b3);
    const T JnakK = T(2.0f) * (E4 * b1 - E3 * a1);
    const T Pnak = T(30);
    const T INaK = Pnak * (zna * JnakNa + zk * JnakK);

    // background currents and sarcolemmal pump
    const T xkb = T(1.0f) / (T(1.0f) + dexp<T>(-(v - T(14.48f)) / T(18.34f)));
    const T GKb = T(0.003);
    const T IKb = GKb * xkb * (v - EK);
    const T PNab = T(3.75e-10);
    const T ev = e1;
    const T PCab = T(2.5e-8);
    T INab, ICab;
    if constexpr(sizeof(T) == 8) {
        INab = PNab * vffrt * (nai * ev - nao) / (ev - T(1.0f));
        ICab = PCab * T(4.0f) * vffrt * (cai * e2 - T(0.341f) * cao) / (e2 - T(1.0f));
    } else {
        INab = PNab * F * x_over_expm1<T>(vfrt) * (nai * ev - nao);
        ICab = PCab * T(2.0f) * F * x_over_expm1<T>(T(2.0f) * vfrt) * (cai * e2 - T(0.341f) * cao);
    }
    const T GpCa = T(0.0005);
    const T IpCa = GpCa * cai / (T(0.0005f) + cai);

    const T dv = -dt * (INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCa + INaK + INab + IKb + IpCa + ICab + stim);

    // CaMK (cass has not been updated yet, so CaMKb is unchanged)
    CaMKb = CaMKo * (T(1.0f) - CaMKt) / (T(1.0f) + KmCaM / cass);
    CaMKa = CaMKb + CaMKt;
    CaMKt += dt * (aCaMK * CaMKb * (CaMKb + CaMKt) - bCaMK * CaMKt);

    const T JdiffNa = (nass - nai) / T(2.0f);
    const T JdiffK = (kss - ki) / T(2.0f);
    const T Jdiff = (cass - cai) / T(0.2f);
    const T bt = T(4.75);
    const T a_rel = T(0.5f) * bt;
    const T q = T(1.5f) / cajsr;
    const T q2 = q * q, q4 = q2 * q2, q8 = q4 * q4;
    const T Jrel_inf = a_rel * (-ICaL) / (T(1.0f) + q8);
    T tau_rel = bt / (T(1.0f) + T(0.0123f) / cajsr);
    if(tau_rel < T(0.005)) tau_rel = T(0.005);
    Jrelnp = rl_from<T>(Jrel_inf, Jrelnp, dt, tau_rel);
    st(38, Jrelnp);
    const T btp = T(1.25f) * bt;
    const T a_relp = T(0.5f) * btp;
    const T Jrel_infp = a_relp * (-ICaL) / (T(1.0f) + q8);
    T tau_relp = btp / (T(1.0f) + T(0.0123f) / cajsr);
    if(tau_relp < T(0.005f)) tau_relp = T(0.005f);
    Jrelp = rl_from<T>(Jrel_infp, Jrelp, dt, tau_relp);
    st(39, Jrelp);
    const T fJ = (T(1.0f) / (T(1.0f) + KmCaMK / CaMKa));
    const T Jrel = (T(1.0f) - fJ) * Jrelnp + fJ * Jrelp;

    const T Jupnp = T(0.004375f) * cai / (cai + T(0.00092f));
    const T Jupp = T(2.75f * 0.004375f) * cai / (cai + T(0.00092f) - T(0.00017f));
    const T Jleak = T(0.0039375f) * cansr / T(15.0f);
    const T Jup = (T(1.0f) - fJ) * Jupnp + fJ * Jupp - Jleak;
    const T Jtr = (cansr - cajsr) / T(100.0f);

    nai += dt * (-(INa + INaL + T(3.0) * INaCa_i + T(3.0) * INaK + INab) * Acap / (F * vmyo) + JdiffNa * vss / vmyo);
    nass += dt * (-(ICaNa + T(3.0) * INaCa_ss) * Acap / (F * vss) - JdiffNa);
    ki += dt * (-(Ito + IKr + IKs + IK1 + IKb + stim - T(2.0f) * INaK) * Acap / (F * vmyo) + JdiffK * vss / vmyo);
    kss += dt * (-(ICaK)*Acap / (F * vss) - JdiffK);
    // buffering factors 1 / (1 + a / c1^2 + b / c2^2) over one division
    const T c1 = kmcmdn + cai, c2 = kmtrpn + cai;
    const T c1s = c1 * c1, c2s = c2 * c2;
    const T Bcai = (c1s * c2s) / (c1s * c2s + cmdnmax * kmcmdn * c2s + trpnmax * kmtrpn * c1s);
    cai += dt * (Bcai * (-(IpCa + ICab - T(2.0f) * INaCa_i) * Acap / (T(2.0f) * F * vmyo) - Jup * vnsr / vmyo + Jdiff * vss / vmyo));
    const T c3 = KmBSR + cass, c4 = KmBSL + cass;
    const T c3s = c3 * c3, c4s = c4 * c4;
    const T Bcass = (c3s * c4s) / (c3s * c4s + BSRmax * KmBSR * c4s + BSLmax * KmBSL * c3s);
    cass += dt * (Bcass * (-(ICaL - T(2.0f) * INaCa_ss) * Acap / (T(2.0f) * F * vss) + Jrel * vjsr / vss - Jdiff));
    cansr += dt * (Jup - Jtr * vjsr / vnsr);
    const T c5s = (kmcsqn + cajsr) * (kmcsqn + cajsr);
    const T Bcajsr = c5s / (c5s + csqnmax * kmcsqn);
    cajsr += dt * (Bcajsr * (Jtr - Jrel));
    st(1, nai); st(2, nass); st(3, ki); st(4, kss); st(5, cai); st(6, cass); st(7, cansr); st(8, cajsr);
    st(40, CaMKt);

    y[1] = nai;    y[2] = nass;   y[3] = ki;      y[4] = kss;    y[5] = cai;
    y[6] = cass;   y[7] = cansr;  y[8] = cajsr;   y[9] = m;      y[10] = hf;
    y[11] = hs;    y[12] = j;     y[13] = hsp;    y[14] = jp;    y[15] = mL;
    y[16] = hL;    y[17] = hLp;   y[18] = a;      y[19] = iF;    y[20] = iS;
    y[21] = ap;    y[22] = iFp;   y[23] = iSp;    y[24] = d;     y[25] = ff;
    y[26] = fs;    y[27] = fcaf;  y[28] = fcas;   y[29] = jca;   y[30] = nca;
    y[31] = ffp;   y[32] = fcafp; y[33] = xrf;    y[34] = xrs;   y[35] = xs1;
    y[36] = xs2;   y[37] = xk1;   y[38] = Jrelnp; y[39] = Jrelp; y[40] = CaMKt;
    return dv;
}

}  // namespace mono
