/*
 * Oracle: Courtemanche-Ramirez-Nattel 1998 human atrial model (21 states) and
 * the adaptive forward-Euler integrator of the reference's default solvers.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Model follows models_library/courtemanche_ramirez_nattel_1998/:
 *   initial state   courtemanche_ramirez_nattel_1998.c:31-51 (float literals),
 *                   INITIAL_V -81.18f (courtemanche_ramirez_nattel_1998.h:7)
 *   right-hand side courtemanche_ramirez_nattel_1998_common.inc.c:1-164
 * Parameters carry the reference's float rounding (an `f` literal there is
 * float-rounded even in the FP64 build) and every expression keeps its operand
 * order, so that, compiled without FMA contraction, results agree bit for bit.
 *
 * Integrators follow models_library/default_solvers.c:
 *   fixed step      solve_model_ode_cpu :48-59 (forward Euler on every state)
 *   adaptive step   solve_forward_euler_cpu_adpt :61-176 (Euler with an
 *                   embedded second RHS evaluation as error estimate, safety
 *                   0.8, dt clamped to [min_dt, max_dt], per-cell dt, time and
 *                   previous dt carried between calls)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#define CRN_NEQ 21

void orc_crn1998_init_cell(double *y) {
    static const float ic[CRN_NEQ] = {-8.118000e+01f, 2.908000e-03f, 9.649000e-01f, 9.775000e-01f, 3.043000e-02f,
                                      9.992000e-01f,  4.966000e-03f, 9.986000e-01f, 3.296000e-05f, 1.869000e-02f,
                                      1.367000e-04f,  9.996000e-01f, 7.755000e-01f, 0.0f,          1.000000e+00f,
                                      9.992000e-01f,  1.117000e+01f, 1.390000e+02f, 1.013000e-04f, 1.488000e+00f,
                                      1.488000e+00f};
    for(int s = 0; s < CRN_NEQ; s++) y[s] = (double)ic[s];
}

/* dy/dt of every state; stim in pA (the reference divides the total by Cm) */
void orc_crn1998_rhs(const double *y, double *dy, double stim) {
    const double V = y[0], m = y[1], h = y[2], j = y[3], oa = y[4], oi = y[5], ua = y[6], ui = y[7];
    const double xr = y[8], xs = y[9], d = y[10], f = y[11], fCa = y[12], u = y[13], v = y[14], w = y[15];
    const double Na_i = y[16], K_i = y[17], Ca_i = y[18], Ca_up = y[19], Ca_rel = y[20];

    /* parameters (:13-50), float-rounded */
    const double Cm = 1.0e+02f, g_Na = 7.8f, R = 8.3143f, T = 3.1e+02f, F = 9.64867e+01f;
    const double Na_o = 1.4e+02f, K_o = 5.4f, g_K1 = 9.0e-02f, g_to = 1.652e-01f, K_Q10 = 3.0f;
    const double g_Kr = 2.9411765e-02f, g_Ks = 1.2941176e-01f, g_Ca_L = 1.2375e-01f, i_NaK_max = 5.9933874e-01f;
    const double Km_Na_i = 1.0e+01f, Km_K_o = 1.5f, Ca_o = 1.8f, g_B_Na = 6.744375e-04f, g_B_Ca = 1.131e-03f;
    const double g_B_K = 0.0f, I_NaCa_max = 1.6e+03f, gamma = 3.5e-01f, K_mNa = 8.75e+01f, K_mCa = 1.38f;
    const double K_sat = 1.0e-01f, i_CaP_max = 2.75e-01f, K_rel = 3.0e+01f, tau_tr = 1.8e+02f;
    const double I_up_max = 5.0e-03f, K_up = 9.2e-04f, Ca_up_max = 1.5e+01f, CMDN_max = 5.0e-02f;
    const double Km_CMDN = 2.38e-03f, TRPN_max = 7.000000000000001e-02f, Km_TRPN = 5.0e-04f, CSQN_max = 1.0e+01f;
    const double Km_CSQN = 8.0e-01f, V_cell = 2.01e+04f;

    const double vs = V - (-1.0e+01); /* the reference's (V_old_-(-10)) */

    /* reversal potentials and gate rates, :52-99 */
    const double E_Na = ((R * T) / F) * log(Na_o / Na_i);
    double a_m;
    if(V == (-4.713e+01)) a_m = 3.2;
    else a_m = (3.2e-01 * (V + 4.713e+01)) / (1.0 - exp((-1.0e-01) * (V + 4.713e+01)));
    const double b_m = 8.0e-02 * exp((-V) / 1.1e+01);
    const double a_h = V < (-4.0e+01) ? 1.35e-01 * exp((V + 8.0e+01) / (-6.8)) : 0.0;
    const double b_h = V < (-4.0e+01) ? (3.56 * exp(7.9e-02 * V)) + (3.1e+05 * exp(3.5e-01 * V))
                                     : 1.0 / (1.3e-01 * (1.0 + exp((V + 1.066e+01) / (-1.11e+01))));
    const double a_j = V < (-4.0e+01)
                           ? ((((-1.2714e+05) * exp(2.444e-01 * V)) - (3.474e-05 * exp((-4.391e-02) * V))) * (V + 3.778e+01)) /
                                 (1.0 + exp(3.11e-01 * (V + 7.923e+01)))
                           : 0.0;
    const double b_j = V < (-4.0e+01) ? (1.212e-01 * exp((-1.052e-02) * V)) / (1.0 + exp((-1.378e-01) * (V + 4.014e+01)))
                                     : (3.0e-01 * exp((-2.535e-07) * V)) / (1.0 + exp((-1.0e-01) * (V + 3.2e+01)));
    const double E_K = ((R * T) / F) * log(K_o / K_i);
    const double a_oa = 6.5e-01 * pow(exp(vs / (-8.5)) + exp((vs - 4.0e+01) / (-5.9e+01)), -1.0);
    const double b_oa = 6.5e-01 * pow(2.5 + exp((vs + 7.2e+01) / 1.7e+01), -1.0);
    const double oa_inf = pow(1.0 + exp((vs + 1.047e+01) / (-1.754e+01)), -1.0);
    const double a_oi = pow(1.853e+01 + (1.0 * exp((vs + 1.037e+02) / 1.095e+01)), -1.0);
    const double b_oi = pow(3.556e+01 + (1.0 * exp((vs - 8.74) / (-7.44))), -1.0);
    const double oi_inf = pow(1.0 + exp((vs + 3.31e+01) / 5.3), -1.0);
    const double g_Kur = 5.0e-03 + (5.0e-02 / (1.0 + exp((V - 1.5e+01) / (-1.3e+01))));
    const double a_ua = 6.5e-01 * pow(exp(vs / (-8.5)) + exp((vs - 4.0e+01) / (-5.9e+01)), -1.0);
    const double b_ua = 6.5e-01 * pow(2.5 + exp((vs + 7.2e+01) / 1.7e+01), -1.0);
    const double ua_inf = pow(1.0 + exp((vs + 2.03e+01) / (-9.6)), -1.0);
    const double a_ui = pow(2.1e+01 + (1.0 * exp((vs - 1.95e+02) / (-2.8e+01))), -1.0);
    const double b_ui = 1.0 / exp((vs - 1.68e+02) / (-1.6e+01));
    const double ui_inf = pow(1.0 + exp((vs - 1.0945e+02) / 2.748e+01), -1.0);
    double a_xr, b_xr, a_xs, b_xs, tau_d, tau_w;
    if(fabs(V + 1.41e+01) < 1.0e-10) a_xr = 1.5e-03;
    else a_xr = (3.0e-04 * (V + 1.41e+01)) / (1.0 - exp((V + 1.41e+01) / (-5.0)));
    if(fabs(V - 3.3328) < 1.0e-10) b_xr = 3.7836118e-04;
    else b_xr = (7.3898e-05 * (V - 3.3328)) / (exp((V - 3.3328) / 5.1237) - 1.0);
    const double xr_inf = pow(1.0 + exp((V + 1.41e+01) / (-6.5)), -1.0);
    if(fabs(V - 1.99e+01) < 1.0e-10) a_xs = 6.8e-04;
    else a_xs = (4.0e-05 * (V - 1.99e+01)) / (1.0 - exp((V - 1.99e+01) / (-1.7e+01)));
    if(fabs(V - 1.99e+01) < 1.0e-10) b_xs = 3.15e-04;
    else b_xs = (3.5e-05 * (V - 1.99e+01)) / (exp((V - 1.99e+01) / 9.0) - 1.0);
    const double xs_inf = pow(1.0 + exp((V - 1.99e+01) / (-1.27e+01)), -5.0e-01);
    const double i_Ca_L = Cm * g_Ca_L * d * f * fCa * (V - 6.5e+01);
    const double d_inf = pow(1.0 + exp((V + 1.0e+01) / (-8.0)), -1.0);
    if(fabs(V + 1.0e+01) < 1.0e-10) tau_d = 4.579 / (1.0 + exp((V + 1.0e+01) / (-6.24)));
    else tau_d = (1.0 - exp((V + 1.0e+01) / (-6.24))) / (3.5e-02 * (V + 1.0e+01) * (1.0 + exp((V + 1.0e+01) / (-6.24))));
    const double f_inf = exp((-(V + 2.8e+01)) / 6.9) / (1.0 + exp((-(V + 2.8e+01)) / 6.9));
    const double tau_f = 9.0 * pow((1.97e-02 * exp((-pow(3.37e-02, 2.0)) * pow(V + 1.0e+01, 2.0))) + 2.0e-02, -1.0);
    const double fCa_inf = pow(1.0 + (Ca_i / 3.5e-04), -1.0);
    const double tau_fCa = 2.0;
    const double sigma = (1.0 / 7.0) * (exp(Na_o / 6.73e+01) - 1.0);
    const double E_Ca = ((R * T) / (2.0 * F)) * log(Ca_o / Ca_i);
    const double i_NaCa = (Cm * I_NaCa_max *
                           ((exp((gamma * F * V) / (R * T)) * pow(Na_i, 3.0) * Ca_o) -
                            (exp(((gamma - 1.0) * F * V) / (R * T)) * pow(Na_o, 3.0) * Ca_i))) /
                          ((pow(K_mNa, 3.0) + pow(Na_o, 3.0)) * (K_mCa + Ca_o) *
                           (1.0 + (K_sat * exp(((gamma - 1.0) * V * F) / (R * T)))));
    const double i_CaP = (Cm * i_CaP_max * Ca_i) / (5.0e-04 + Ca_i);
    const double i_rel = K_rel * pow(u, 2.0) * v * w * (Ca_rel - Ca_i);
    const double tau_u = 8.0;
    if(fabs(V - 7.9) < 1.0e-10) tau_w = (6.0 * 2.0e-01) / 1.3;
    else tau_w = (6.0 * (1.0 - exp((-(V - 7.9)) / 5.0))) / ((1.0 + (3.0e-01 * exp((-(V - 7.9)) / 5.0))) * 1.0 * (V - 7.9));
    const double w_inf = 1.0 - pow(1.0 + exp((-(V - 4.0e+01)) / 1.7e+01), -1.0);
    const double i_tr = (Ca_up - Ca_rel) / tau_tr;
    const double i_up = I_up_max / (1.0 + (K_up / Ca_i));
    const double i_up_leak = (I_up_max * Ca_up) / Ca_up_max;
    const double V_i = V_cell * 6.8e-01;
    const double V_rel = 4.8e-03 * V_cell;
    const double V_up = 5.52e-02 * V_cell;
    const double B2 = 1.0 + ((TRPN_max * Km_TRPN) / pow(Ca_i + Km_TRPN, 2.0)) + ((CMDN_max * Km_CMDN) / pow(Ca_i + Km_CMDN, 2.0));

    /* currents and derived rates, :100-130 */
    const double m_inf = a_m / (a_m + b_m), tau_m = 1.0 / (a_m + b_m);
    const double h_inf = a_h / (a_h + b_h), tau_h = 1.0 / (a_h + b_h);
    const double j_inf = a_j / (a_j + b_j), tau_j = 1.0 / (a_j + b_j);
    const double i_K1 = (Cm * g_K1 * (V - E_K)) / (1.0 + exp(7.000000000000001e-02 * (V + 8.0e+01)));
    const double i_to = Cm * g_to * pow(oa, 3.0) * oi * (V - E_K);
    const double tau_oa = pow(a_oa + b_oa, -1.0) / K_Q10;
    const double tau_oi = pow(a_oi + b_oi, -1.0) / K_Q10;
    const double i_Kur = Cm * g_Kur * pow(ua, 3.0) * ui * (V - E_K);
    const double tau_ua = pow(a_ua + b_ua, -1.0) / K_Q10;
    const double tau_ui = pow(a_ui + b_ui, -1.0) / K_Q10;
    const double i_Kr = (Cm * g_Kr * xr * (V - E_K)) / (1.0 + exp((V + 1.5e+01) / 2.24e+01));
    const double tau_xr = pow(a_xr + b_xr, -1.0);
    const double i_Ks = Cm * g_Ks * pow(xs, 2.0) * (V - E_K);
    const double tau_xs = 5.0e-01 * pow(a_xs + b_xs, -1.0);
    const double f_NaK = pow(1.0 + (1.245e-01 * exp(((-1.0e-01) * F * V) / (R * T))) + (3.65e-02 * sigma * exp(((-F) * V) / (R * T))), -1.0);
    const double i_B_Na = Cm * g_B_Na * (V - E_Na);
    const double i_B_Ca = Cm * g_B_Ca * (V - E_Ca);
    const double i_B_K = Cm * g_B_K * (V - E_K);
    const double i_Na = Cm * g_Na * pow(m, 3.0) * h * j * (V - E_Na);
    const double i_NaK = (((Cm * i_NaK_max * f_NaK * 1.0) / (1.0 + pow(Km_Na_i / Na_i, 1.5))) * K_o) / (K_o + Km_K_o);
    const double Fn = 1.0e+03 * ((1.0e-15 * V_rel * i_rel) - ((1.0e-15 / (2.0 * F)) * ((5.0e-01 * i_Ca_L) - (2.0e-01 * i_NaCa))));
    const double B1 = (((2.0 * i_NaCa) - (i_CaP + i_Ca_L + i_B_Ca)) / (2.0 * V_i * F)) + (((V_up * (i_up_leak - i_up)) + (i_rel * V_rel)) / V_i);
    const double u_inf = pow(1.0 + exp((-(Fn - 3.4175e-13)) / 1.367e-15), -1.0);
    const double tau_v = 1.91 + (2.09 * pow(1.0 + exp((-(Fn - 3.4175e-13)) / 1.367e-15), -1.0));
    const double v_inf = 1.0 - pow(1.0 + exp((-(Fn - 6.834999999999999e-14)) / 1.367e-15), -1.0);

    /* :132-152 */
    dy[0] = (-(i_Na + i_K1 + i_to + i_Kur + i_Kr + i_Ks + i_B_Na + i_B_Ca + i_NaK + i_CaP + i_NaCa + i_Ca_L + stim)) / Cm;
    dy[1] = (m_inf - m) / tau_m;
    dy[2] = (h_inf - h) / tau_h;
    dy[3] = (j_inf - j) / tau_j;
    dy[4] = (oa_inf - oa) / tau_oa;
    dy[5] = (oi_inf - oi) / tau_oi;
    dy[6] = (ua_inf - ua) / tau_ua;
    dy[7] = (ui_inf - ui) / tau_ui;
    dy[8] = (xr_inf - xr) / tau_xr;
    dy[9] = (xs_inf - xs) / tau_xs;
    dy[10] = (d_inf - d) / tau_d;
    dy[11] = (f_inf - f) / tau_f;
    dy[12] = (fCa_inf - fCa) / tau_fCa;
    dy[13] = (u_inf - u) / tau_u;
    dy[14] = (v_inf - v) / tau_v;
    dy[15] = (w_inf - w) / tau_w;
    dy[16] = (((-3.0) * i_NaK) - ((3.0 * i_NaCa) + i_B_Na + i_Na)) / (V_i * F);
    dy[17] = ((2.0 * i_NaK) - (i_K1 + i_to + i_Kur + i_Kr + i_Ks + i_B_K)) / (V_i * F);
    dy[18] = B1 / B2;
    dy[19] = i_up - (i_up_leak + ((i_tr * V_rel) / V_up));
    dy[20] = (i_tr - i_rel) * pow(1.0 + ((CSQN_max * Km_CSQN) / pow(Ca_rel + Km_CSQN, 2.0)), -1.0);
}

/* solve_model_ode_cpu: forward Euler on every state */
void orc_crn1998_step(double *y, double stim, double dt) {
    double dy[CRN_NEQ];
    orc_crn1998_rhs(y, dy, stim);
    for(int i = 0; i < CRN_NEQ; i++) y[i] = dt * dy[i] + y[i];
}

/* solve_forward_euler_cpu_adpt (default_solvers.c:61-176) for one cell: advance
 * y to final_time with the cell's own step size.  aux[0] = dt, aux[1] =
 * time_new, aux[2] = previous_dt (the reference's ode_dt, ode_time_new,
 * ode_previous_dt).  Returns the number of accepted steps with error >= 1
 * (the reference prints a line for each). */
typedef void (*orc_rhs_fn)(const double *y, double *dy, double stim);

int orc_adaptive_euler(orc_rhs_fn rhs, int neq, double *y, double stim, double final_time, double *aux, double min_dt,
                       double max_dt, double abs_tol, double rel_tol) {
    double k1[64], k2[64], y_old[64], y_new[64];
    double *dt = &aux[0], *time_new = &aux[1], *previous_dt = &aux[2];
    int forced = 0;
    *previous_dt = *dt; /* :72 */
    if(*time_new + *dt > final_time) *dt = final_time - *time_new;
    rhs(y, k1, stim);
    *time_new += *dt;
    const double tiny = pow(abs_tol, 2.0);
    for(;;) {
        for(int i = 0; i < neq; i++) {
            y_old[i] = y[i];
            y_new[i] = k1[i] * *dt + y_old[i];
            y[i] = y_new[i];
        }
        *time_new += *dt;
        rhs(y, k2, stim);
        *time_new -= *dt;
        double err = 0.0;
        for(int i = 0; i < neq; i++) {
            const double aux_tol = fabs(y_new[i]) * rel_tol;
            const double tol = (abs_tol > aux_tol) ? abs_tol : aux_tol;
            const double e = fabs(((*dt / 2.0) * (k1[i] - k2[i])) / tol);
            err = (e > err) ? e : err;
        }
        err += tiny;
        *previous_dt = *dt;
        *dt = 0.8 * (*dt) * sqrt(1.0f / err);
        if(*dt < min_dt) *dt = min_dt;
        else if(*dt > max_dt) *dt = max_dt;
        if(*time_new + *dt > final_time) *dt = final_time - *time_new;
        if(err >= 1.0f && *dt > min_dt) {
            for(int i = 0; i < neq; i++) y[i] = y_old[i];
        } else {
            if(err >= 1.0) forced++;
            for(int i = 0; i < neq; i++) {
                k1[i] = k2[i];
                y[i] = y_new[i];
            }
            if(*time_new + *previous_dt >= final_time) {
                if(final_time == *time_new) {
                    break;
                } else if(*time_new < final_time) {
                    *dt = *previous_dt = final_time - *time_new;
                    *time_new += *previous_dt;
                    break;
                }
            } else {
                *time_new += *previous_dt;
            }
        }
    }
    return forced;
}
