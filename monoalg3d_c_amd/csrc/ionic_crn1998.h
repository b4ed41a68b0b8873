// Courtemanche-Ramirez-Nattel 1998 human atrial model (21 states), the model
// of the reference's default-solver family used here for the forward-Euler and
// adaptive-Euler integrators
// (courtemanche_ramirez_nattel_1998_common.inc.c:1-164; states and defaults
// courtemanche_ramirez_nattel_1998.c:31-51).  Unlike TT2006 / ORd it has no
// Rush-Larsen gates: the integrator advances every state from the full RHS.
//
// FP64 keeps the reference's float-rounded parameters and its pow() calls;
// FP32 writes pow(x, -1), pow(x, 2) and pow(x, 3) as a division and products.
#pragma once

#include "ionic_models.h"

namespace mono {

constexpr int kCRNNeq = 21;

// FP64: the device libm's exp, not the models' exp_cw.  The adaptive
// integrator's accept/reject test turns last-bit differences into different
// step sequences, and exp_cw rounds differently from glibc's exp for ~7% of
// arguments: with it, the 400 ms two-cell adaptive trace drifted 0.15 ms from
// the reference's activation time, with the library exp it stays within the
// 0.1 ms its test allows.  The model is on no timed path.
template <typename T> __device__ __forceinline__ T crn_exp(T x) {
    if constexpr(sizeof(T) == 8) return exp(x);
    else return dexp<T>(x);
}
template <typename T> __device__ __forceinline__ T p_inv(T x) {
    if constexpr(sizeof(T) == 8) return pow(x, -1.0);
    else return T(1) / x;
}
template <typename T> __device__ __forceinline__ T p_sq(T x) {
    if constexpr(sizeof(T) == 8) return pow(x, 2.0);
    else return x * x;
}
template <typename T> __device__ __forceinline__ T p_cube(T x) {
    if constexpr(sizeof(T) == 8) return pow(x, 3.0);
    else return x * x * x;
}

// dy = f(y); stim (pA) enters dV/dt as -(sum of currents + stim) / Cm
template <typename T> __device__ __forceinline__ void crn1998_rhs(const T *y, T *dy, T stim) {
    const T V = y[0], m = y[1], h = y[2], j = y[3], oa = y[4], oi = y[5], ua = y[6], ui = y[7];
    const T xr = y[8], xs = y[9], d = y[10], f = y[11], fCa = y[12], u = y[13], v = y[14], w = y[15];
    const T Na_i = y[16], K_i = y[17], Ca_i = y[18], Ca_up = y[19], Ca_rel = y[20];

    const T Cm = T(1.0e+02f), g_Na = T(7.8f), R = T(8.3143f), Tk = T(3.1e+02f), F = T(9.64867e+01f);
    const T Na_o = T(1.4e+02f), K_o = T(5.4f), g_K1 = T(9.0e-02f), g_to = T(1.652e-01f), K_Q10 = T(3.0f);
    const T g_Kr = T(2.9411765e-02f), g_Ks = T(1.2941176e-01f), g_Ca_L = T(1.2375e-01f), i_NaK_max = T(5.9933874e-01f);
    const T Km_Na_i = T(1.0e+01f), Km_K_o = T(1.5f), Ca_o = T(1.8f), g_B_Na = T(6.744375e-04f), g_B_Ca = T(1.131e-03f);
    const T I_NaCa_max = T(1.6e+03f), gam = T(3.5e-01f), K_mNa = T(8.75e+01f), K_mCa = T(1.38f);
    const T K_sat = T(1.0e-01f), i_CaP_max = T(2.75e-01f), K_rel = T(3.0e+01f), tau_tr = T(1.8e+02f);
    const T I_up_max = T(5.0e-03f), K_up = T(9.2e-04f), Ca_up_max = T(1.5e+01f), CMDN_max = T(5.0e-02f);
    const T Km_CMDN = T(2.38e-03f), TRPN_max = T(7.000000000000001e-02f), Km_TRPN = T(5.0e-04f), CSQN_max = T(1.0e+01f);
    const T Km_CSQN = T(8.0e-01f), V_cell = T(2.01e+04f);
    const T RTF = (R * Tk) / F;
    const T vs = V - T(-1.0e+01);

    // reversal potentials
    const T E_Na = RTF * dlog<T>(Na_o / Na_i);
    const T E_K = RTF * dlog<T>(K_o / K_i);
    const T E_Ca = ((R * Tk) / (T(2.0) * F)) * dlog<T>(Ca_o / Ca_i);

    // I_Na gates m, h, j
    const T a_m = V == T(-4.713e+01) ? T(3.2) : (T(3.2e-01) * (V + T(4.713e+01))) / (T(1.0) - crn_exp<T>(T(-1.0e-01) * (V + T(4.713e+01))));
    const T b_m = T(8.0e-02) * crn_exp<T>((-V) / T(1.1e+01));
    const bool lo = V < T(-4.0e+01);
    const T a_h = lo ? T(1.35e-01) * crn_exp<T>((V + T(8.0e+01)) / T(-6.8)) : T(0.0);
    const T b_h = lo ? (T(3.56) * crn_exp<T>(T(7.9e-02) * V)) + (T(3.1e+05) * crn_exp<T>(T(3.5e-01) * V))
                     : T(1.0) / (T(1.3e-01) * (T(1.0) + crn_exp<T>((V + T(1.066e+01)) / T(-1.11e+01))));
    const T a_j = lo ? (((T(-1.2714e+05) * crn_exp<T>(T(2.444e-01) * V)) - (T(3.474e-05) * crn_exp<T>(T(-4.391e-02) * V))) * (V + T(3.778e+01))) /
                           (T(1.0) + crn_exp<T>(T(3.11e-01) * (V + T(7.923e+01))))
                     : T(0.0);
    const T b_j = lo ? (T(1.212e-01) * crn_exp<T>(T(-1.052e-02) * V)) / (T(1.0) + crn_exp<T>(T(-1.378e-01) * (V + T(4.014e+01))))
                     : (T(3.0e-01) * crn_exp<T>(T(-2.535e-07) * V)) / (T(1.0) + crn_exp<T>(T(-1.0e-01) * (V + T(3.2e+01))));

    // transient outward and ultrarapid K gates (oa/ua share their rates)
    const T a_oa = T(6.5e-01) * p_inv<T>(crn_exp<T>(vs / T(-8.5)) + crn_exp<T>((vs - T(4.0e+01)) / T(-5.9e+01)));
    const T b_oa = T(6.5e-01) * p_inv<T>(T(2.5) + crn_exp<T>((vs + T(7.2e+01)) / T(1.7e+01)));
    const T oa_inf = p_inv<T>(T(1.0) + crn_exp<T>((vs + T(1.047e+01)) / T(-1.754e+01)));
    const T a_oi = p_inv<T>(T(1.853e+01) + (T(1.0) * crn_exp<T>((vs + T(1.037e+02)) / T(1.095e+01))));
    const T b_oi = p_inv<T>(T(3.556e+01) + (T(1.0) * crn_exp<T>((vs - T(8.74)) / T(-7.44))));
    const T oi_inf = p_inv<T>(T(1.0) + crn_exp<T>((vs + T(3.31e+01)) / T(5.3)));
    const T g_Kur = T(5.0e-03) + (T(5.0e-02) / (T(1.0) + crn_exp<T>((V - T(1.5e+01)) / T(-1.3e+01))));
    const T ua_inf = p_inv<T>(T(1.0) + crn_exp<T>((vs + T(2.03e+01)) / T(-9.6)));
    const T a_ui = p_inv<T>(T(2.1e+01) + (T(1.0) * crn_exp<T>((vs - T(1.95e+02)) / T(-2.8e+01))));
    const T b_ui = T(1.0) / crn_exp<T>((vs - T(1.68e+02)) / T(-1.6e+01));
    const T ui_inf = p_inv<T>(T(1.0) + crn_exp<T>((vs - T(1.0945e+02)) / T(2.748e+01)));

    // delayed rectifiers
    const T a_xr = fabs(V + T(1.41e+01)) < T(1.0e-10) ? T(1.5e-03)
                                                      : (T(3.0e-04) * (V + T(1.41e+01))) / (T(1.0) - crn_exp<T>((V + T(1.41e+01)) / T(-5.0)));
    const T b_xr = fabs(V - T(3.3328)) < T(1.0e-10) ? T(3.7836118e-04)
                                                    : (T(7.3898e-05) * (V - T(3.3328))) / (crn_exp<T>((V - T(3.3328)) / T(5.1237)) - T(1.0));
    const T xr_inf = p_inv<T>(T(1.0) + crn_exp<T>((V + T(1.41e+01)) / T(-6.5)));
    const bool at_xs = fabs(V - T(1.99e+01)) < T(1.0e-10);
    const T a_xs = at_xs ? T(6.8e-04) : (T(4.0e-05) * (V - T(1.99e+01))) / (T(1.0) - crn_exp<T>((V - T(1.99e+01)) / T(-1.7e+01)));
    const T b_xs = at_xs ? T(3.15e-04) : (T(3.5e-05) * (V - T(1.99e+01))) / (crn_exp<T>((V - T(1.99e+01)) / T(9.0)) - T(1.0));
    const T xs_inf = dpow<T>(T(1.0) + crn_exp<T>((V - T(1.99e+01)) / T(-1.27e+01)), T(-5.0e-01));

    // L-type Ca and its gates
    const T i_Ca_L = Cm * g_Ca_L * d * f * fCa * (V - T(6.5e+01));
    const T d_inf = p_inv<T>(T(1.0) + crn_exp<T>((V + T(1.0e+01)) / T(-8.0)));
    const T ed = crn_exp<T>((V + T(1.0e+01)) / T(-6.24));
    const T tau_d = fabs(V + T(1.0e+01)) < T(1.0e-10) ? T(4.579) / (T(1.0) + ed)
                                                      : (T(1.0) - ed) / (T(3.5e-02) * (V + T(1.0e+01)) * (T(1.0) + ed));
    const T ef = crn_exp<T>((-(V + T(2.8e+01))) / T(6.9));
    const T f_inf = ef / (T(1.0) + ef);
    const T tau_f = T(9.0) * p_inv<T>((T(1.97e-02) * crn_exp<T>((-p_sq<T>(T(3.37e-02))) * p_sq<T>(V + T(1.0e+01)))) + T(2.0e-02));
    const T fCa_inf = p_inv<T>(T(1.0) + (Ca_i / T(3.5e-04)));

    // pumps, exchanger, SR fluxes
    const T sigma = (T(1.0) / T(7.0)) * (crn_exp<T>(Na_o / T(6.73e+01)) - T(1.0));
    const T i_NaCa = (Cm * I_NaCa_max *
                      ((crn_exp<T>((gam * F * V) / (R * Tk)) * p_cube<T>(Na_i) * Ca_o) -
                       (crn_exp<T>(((gam - T(1.0)) * F * V) / (R * Tk)) * p_cube<T>(Na_o) * Ca_i))) /
                     ((p_cube<T>(K_mNa) + p_cube<T>(Na_o)) * (K_mCa + Ca_o) * (T(1.0) + (K_sat * crn_exp<T>(((gam - T(1.0)) * V * F) / (R * Tk)))));
    const T i_CaP = (Cm * i_CaP_max * Ca_i) / (T(5.0e-04) + Ca_i);
    const T i_rel = K_rel * p_sq<T>(u) * v * w * (Ca_rel - Ca_i);
    const T ew = crn_exp<T>((-(V - T(7.9))) / T(5.0));
    const T tau_w = fabs(V - T(7.9)) < T(1.0e-10) ? (T(6.0) * T(2.0e-01)) / T(1.3)
                                                  : (T(6.0) * (T(1.0) - ew)) / ((T(1.0) + (T(3.0e-01) * ew)) * T(1.0) * (V - T(7.9)));
    const T w_inf = T(1.0) - p_inv<T>(T(1.0) + crn_exp<T>((-(V - T(4.0e+01))) / T(1.7e+01)));
    const T i_tr = (Ca_up - Ca_rel) / tau_tr;
    const T i_up = I_up_max / (T(1.0) + (K_up / Ca_i));
    const T i_up_leak = (I_up_max * Ca_up) / Ca_up_max;
    const T V_i = V_cell * T(6.8e-01), V_rel = T(4.8e-03) * V_cell, V_up = T(5.52e-02) * V_cell;
    const T B2 = T(1.0) + ((TRPN_max * Km_TRPN) / p_sq<T>(Ca_i + Km_TRPN)) + ((CMDN_max * Km_CMDN) / p_sq<T>(Ca_i + Km_CMDN));

    // currents
    const T i_K1 = (Cm * g_K1 * (V - E_K)) / (T(1.0) + crn_exp<T>(T(7.000000000000001e-02) * (V + T(8.0e+01))));
    const T i_to = Cm * g_to * p_cube<T>(oa) * oi * (V - E_K);
    const T i_Kur = Cm * g_Kur * p_cube<T>(ua) * ui * (V - E_K);
    const T i_Kr = (Cm * g_Kr * xr * (V - E_K)) / (T(1.0) + crn_exp<T>((V + T(1.5e+01)) / T(2.24e+01)));
    const T i_Ks = Cm * g_Ks * p_sq<T>(xs) * (V - E_K);
    const T f_NaK = p_inv<T>(T(1.0) + (T(1.245e-01) * crn_exp<T>((T(-1.0e-01) * F * V) / (R * Tk))) +
                             (T(3.65e-02) * sigma * crn_exp<T>(((-F) * V) / (R * Tk))));
    const T i_B_Na = Cm * g_B_Na * (V - E_Na);
    const T i_B_Ca = Cm * g_B_Ca * (V - E_Ca);
    const T i_Na = Cm * g_Na * p_cube<T>(m) * h * j * (V - E_Na);
    const T i_NaK = (((Cm * i_NaK_max * f_NaK * T(1.0)) / (T(1.0) + dpow<T>(Km_Na_i / Na_i, T(1.5)))) * K_o) / (K_o + Km_K_o);
    const T Fn = T(1.0e+03) * ((T(1.0e-15) * V_rel * i_rel) - ((T(1.0e-15) / (T(2.0) * F)) * ((T(5.0e-01) * i_Ca_L) - (T(2.0e-01) * i_NaCa))));
    const T B1 = (((T(2.0) * i_NaCa) - (i_CaP + i_Ca_L + i_B_Ca)) / (T(2.0) * V_i * F)) + (((V_up * (i_up_leak - i_up)) + (i_rel * V_rel)) / V_i);
    const T eu = crn_exp<T>((-(Fn - T(3.4175e-13))) / T(1.367e-15));
    const T u_inf = p_inv<T>(T(1.0) + eu);
    const T tau_v = T(1.91) + (T(2.09) * p_inv<T>(T(1.0) + eu));
    const T v_inf = T(1.0) - p_inv<T>(T(1.0) + crn_exp<T>((-(Fn - T(6.834999999999999e-14))) / T(1.367e-15)));

    dy[0] = (-(i_Na + i_K1 + i_to + i_Kur + i_Kr + i_Ks + i_B_Na + i_B_Ca + i_NaK + i_CaP + i_NaCa + i_Ca_L + stim)) / Cm;
    dy[1] = ((a_m / (a_m + b_m)) - m) / (T(1.0) / (a_m + b_m));
    dy[2] = ((a_h / (a_h + b_h)) - h) / (T(1.0) / (a_h + b_h));
    dy[3] = ((a_j / (a_j + b_j)) - j) / (T(1.0) / (a_j + b_j));
    dy[4] = (oa_inf - oa) / (p_inv<T>(a_oa + b_oa) / K_Q10);
    dy[5] = (oi_inf - oi) / (p_inv<T>(a_oi + b_oi) / K_Q10);
    dy[6] = (ua_inf - ua) / (p_inv<T>(a_oa + b_oa) / K_Q10);
    dy[7] = (ui_inf - ui) / (p_inv<T>(a_ui + b_ui) / K_Q10);
    dy[8] = (xr_inf - xr) / p_inv<T>(a_xr + b_xr);
    dy[9] = (xs_inf - xs) / (T(5.0e-01) * p_inv<T>(a_xs + b_xs));
    dy[10] = (d_inf - d) / tau_d;
    dy[11] = (f_inf - f) / tau_f;
    dy[12] = (fCa_inf - fCa) / T(2.0);
    dy[13] = (u_inf - u) / T(8.0);
    dy[14] = (v_inf - v) / tau_v;
    dy[15] = (w_inf - w) / tau_w;
    dy[16] = ((T(-3.0) * i_NaK) - ((T(3.0) * i_NaCa) + i_B_Na + i_Na)) / (V_i * F);
    dy[17] = ((T(2.0) * i_NaK) - (i_K1 + i_to + i_Kur + i_Kr + i_Ks)) / (V_i * F);
    dy[18] = B1 / B2;
    dy[19] = i_up - (i_up_leak + ((i_tr * V_rel) / V_up));
    dy[20] = (i_tr - i_rel) * p_inv<T>(T(1.0) + ((CSQN_max * Km_CSQN) / p_sq<T>(Ca_rel + Km_CSQN)));
}

// One forward-Euler step of every state but V (solve_model_ode_cpu,
// default_solvers.c:48-59); returns the V increment dt * dV/dt.
template <typename T> __device__ __forceinline__ T crn1998_step(T *y, T stim, T dt) {
    T dy[kCRNNeq];
    crn1998_rhs<T>(y, dy, stim);
#pragma unroll
    for(int s = 1; s < kCRNNeq; s++) y[s] = dt * dy[s] + y[s];
    return dt * dy[0];
}

}  // namespace mono
