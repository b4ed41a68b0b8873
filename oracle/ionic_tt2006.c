/*
 * Oracle: ten Tusscher & Panfilov 2006 epicardial ventricular model, one
 * Rush-Larsen / forward-Euler step, FP64.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows models_library/ten_tusscher/ten_tusscher_2006_RS_CPU.c:
 *   initial state          :28-46   (float literals, INITIAL_V = -85.23f)
 *   step (RL gates 1..12,
 *   Euler for V, Ca, R', Na, K)   :100-134
 *   right-hand side        :137-351
 * Every expression keeps the reference's operand order and its float-rounded
 * literals (e.g. F = 96485.3415f) so that, compiled like the reference
 * (gcc, no FMA contraction), results agree bit for bit with it.
 */
#include <math.h>
#include <stdint.h>

#define TT_NEQ 19

void orc_tt2006_init_cell(double *y) {
    static const float ic[TT_NEQ] = {
        -85.23f, 0.00621f, 0.4712f, 0.0095f, 0.00172f, 0.7444f, 0.7045f, 3.373e-5f,
        0.7888f, 0.9755f, 0.9953f, 0.999998f, 2.42e-8f, 0.000126f, 3.64f, 0.00036f,
        0.9073f, 8.604f, 136.89f};
    for(int s = 0; s < TT_NEQ; s++) y[s] = (double)ic[s];
}

/* One step: y (in/out), stim in uA/uF (subtracted from dV/dt).  sKr, sCaL and
 * sKs multiply g_Kr, g_CaL and g_Ks (batch sweeps; no reference counterpart,
 * all 1 is the reference model: x * 1.0 is exact). */
void orc_tt2006_step_scaled(double *y, double stim, double dt, double sKr, double sCaL, double sKs) {
    /* state */
    const double V = y[0], Xr1 = y[1], Xr2 = y[2], Xs = y[3], m = y[4], h = y[5], j = y[6];
    const double d = y[7], f = y[8], f2 = y[9], fCass = y[10], s = y[11], r = y[12];
    const double Ca_i = y[13], Ca_SR = y[14], Ca_ss = y[15], Rp = y[16], Na_i = y[17], K_i = y[18];

    /* physical constants and conductances (RS_CPU.c:161-203) */
    const double R = 8314.472, T = 310.0, F = 96485.3415f, Cm = 0.185, V_c = 0.016404;
    const double Ko = 5.4, Nao = 140.0, Cao = 2.0;
    const double P_kna = 0.03, K_mk = 1.0, P_NaK = 2.724, K_mNa = 40.0, K_pCa = 0.0005;
    const double V_rel = 0.102, k1_prime = 0.15, max_sr = 2.5, min_sr = 1.0, EC = 1.5;
    const double Vmax_up = 0.006375;
    const double nx_alpha = 2.5, nx_gamma = 0.35, K_sat = 0.1, Km_Ca = 1.38, Km_Nai = 87.5;
    const double K_NaCa = 1000.0;
    const double g_to = 0.294, g_Kr = 0.153 * sKr, g_Ks = 0.098 * sKs, g_CaL = 3.98e-05 * sCaL, g_Na = 14.838;
    const double g_pK = 0.0146, g_bca = 0.000592, g_pCa = 0.1238, g_K1 = 5.405, g_bna = 0.00029;

    /* reversal potentials (:206-209) */
    const double EK = ((R * T) / F) * log(Ko / K_i);
    const double EKs = ((R * T) / F) * log((Ko + (P_kna * Nao)) / (K_i + (P_kna * Na_i)));
    const double ENa = ((R * T) / F) * log(Nao / Na_i);
    const double ECa = ((0.5f * R * T) / F) * log(Cao / Ca_i);

    /* potassium currents (:211-219) */
    const double bK1 = ((3.0f * exp(0.0002f * ((V - EK) + 100.0f))) + exp(0.1f * ((V - EK) - 10.0f))) /
                       (1.0f + exp((-0.5f) * (V - EK)));
    const double aK1 = 0.1f / (1.0f + exp(0.06f * ((V - EK) - 200.0f)));
    const double xK1 = aK1 / (aK1 + bK1);
    const double IK1 = g_K1 * xK1 * (V - EK);
    const double Ito = g_to * r * s * (V - EK);
    const double IKr = g_Kr * Xr1 * Xr2 * (V - EK) * sqrt(Ko / 5.4f);
    const double IKs = g_Ks * pow(Xs, 2.0f) * (V - EKs);
    const double IpK = (g_pK * (V - EK)) / (1.0f + exp((25.0f - V) / 5.98f));

    /* L-type calcium current with its removable singularity at V = 15 (:221) */
    double ICaL;
    if(V < 15.0f - 1.0e-5f || V > 15.0f + 1.0e-5f) {
        const double z = (2.0f * (V - 15.0f) * F) / (R * T);
        ICaL = (((g_CaL * d * f * f2 * fCass * 4.0f * (V - 15.0f) * pow(F, 2.0f)) / (R * T)) *
                ((0.25f * Ca_ss * exp(z)) - Cao)) / (expm1(z));
    } else {
        ICaL = g_CaL * d * f * f2 * fCass * 2.0f * F * (0.25f * Ca_ss - Cao);
    }
    const double IbCa = g_bca * (V - ECa);
    const double IpCa = (g_pCa * Ca_i) / (Ca_i + K_pCa);

    /* pumps and exchangers (:225-228) */
    const double INaK = ((((P_NaK * Ko) / (Ko + K_mk)) * Na_i) / (Na_i + K_mNa)) /
                        (1.0f + (0.1245f * exp(((-0.1f) * V * F) / (R * T))) + (0.0353f * exp(((-V) * F) / (R * T))));
    const double INa = g_Na * pow(m, 3.0f) * h * j * (V - ENa);
    const double IbNa = g_bna * (V - ENa);
    const double INaCa =
        (K_NaCa * ((exp((nx_gamma * V * F) / (R * T)) * pow(Na_i, 3.0f) * Cao) -
                   (exp(((nx_gamma - 1.0f) * V * F) / (R * T)) * pow(Nao, 3) * Ca_i * nx_alpha))) /
        ((pow(Km_Nai, 3.0f) + pow(Nao, 3.0f)) * (Km_Ca + Cao) * (1.0f + (K_sat * exp(((nx_gamma - 1.0f) * V * F) / (R * T)))));

    /* gate steady states and time constants (:233-282) */
    const double xr1_inf = 1.0f / (1.0f + exp(((-26.0f) - V) / 7.0f));
    const double tau_xr1 = 1.0f * (450.0f / (1.0f + exp(((-45.0f) - V) / 10.0f))) * (6.0f / (1.0f + exp((V + 30.0f) / 11.5f)));

    const double xr2_inf = 1.0f / (1.0f + exp((V + 88.0f) / 24.0f));
    const double tau_xr2 = 1.0f * (3.0f / (1.0f + exp(((-60.0f) - V) / 20.0f))) * (1.12f / (1.0f + exp((V - 60.0f) / 20.0f)));

    const double xs_inf = 1.0f / (1.0f + exp(((-5.0f) - V) / 14.0f));
    const double tau_xs = (1.0f * (1400.0f / sqrt(1.0f + exp((5.0f - V) / 6.0f))) * (1.0f / (1.0f + exp((V - 35.0f) / 15.0f)))) + 80.0f;

    const double m_inf = 1.0f / pow(1.0f + exp(((-56.86f) - V) / 9.03f), 2.0f);
    const double a_m = 1.0f / (1.0f + exp(((-60.0f) - V) / 5.0f));
    const double b_m = (0.1f / (1.0f + exp((V + 35.0f) / 5.0f))) + (0.1f / (1.0f + exp((V - 50.0f) / 200.0f)));
    const double tau_m = 1.0f * a_m * b_m;

    const double h_inf = 1.0f / pow(1.0f + exp((V + 71.55f) / 7.43f), 2.0f);
    const int low = V < (-40.0f);
    const double a_h = low ? (0.057f * exp((-(V + 80.0f)) / 6.8f)) : 0.0f;
    const double b_h = low ? ((2.7f * exp(0.079f * V)) + (310000.0f * exp(0.3485f * V)))
                           : (0.77f / (0.13f * (1.0f + exp((V + 10.66f) / (-11.1f)))));
    const double tau_h = 1.0f / (a_h + b_h);

    const double j_inf = 1.0f / pow(1.0f + exp((V + 71.55f) / 7.43f), 2.0f);
    const double a_j = low ? ((((((-25428.0f) * exp(0.2444f * V)) - (6.948e-06f * exp((-0.04391f) * V))) * (V + 37.78f)) / 1.0f) /
                              (1.0f + exp(0.311f * (V + 79.23f))))
                           : 0.0f;
    const double b_j = low ? ((0.02424f * exp((-0.01052f) * V)) / (1.0f + exp((-0.1378f) * (V + 40.14f))))
                           : ((0.6f * exp(0.057f * V)) / (1.0f + exp((-0.1f) * (V + 32.0f))));
    const double tau_j = 1.0f / (a_j + b_j);

    const double d_inf = 1.0f / (1.0f + exp(((-8.0f) - V) / 7.5f));
    const double a_d = (1.4f / (1.0f + exp(((-35.0f) - V) / 13.0f))) + 0.25f;
    const double b_d = 1.4f / (1.0f + exp((V + 5.0f) / 5.0f));
    const double g_d = 1.0f / (1.0f + exp((50.0f - V) / 20.0f));
    const double tau_d = (1.0f * a_d * b_d) + g_d;

    const double f_inf = 1.0f / (1.0f + exp((V + 20.0f) / 7.0f));
    const double tau_f = (1102.5f * exp((-pow(V + 27.0f, 2.0f)) / 225.0f)) + (200.0f / (1.0f + exp((13.0f - V) / 10.0f))) +
                         (180.0f / (1.0f + exp((V + 30.0f) / 10.0f))) + 20.0f;

    const double f2_inf = (0.67f / (1.0f + exp((V + 35.0f) / 7.0f))) + 0.33f;
    const double tau_f2 = (562.0f * exp((-pow(V + 27.0f, 2.0f)) / 240.0f)) + (31.0f / (1.0f + exp((25.0f - V) / 10.0f))) +
                          (80.0f / (1.0f + exp((V + 30.0f) / 10.0f)));

    const double fCass_inf = (0.6f / (1.0f + pow(Ca_ss / 0.05f, 2.0f))) + 0.4f;
    const double tau_fCass = (80.0f / (1.0f + pow(Ca_ss / 0.05f, 2.0f))) + 2.0f;

    const double s_inf = 1.0f / (1.0f + exp((V + 20.0f) / 5.0f));
    const double tau_s = (85.0f * exp((-pow(V + 45.0f, 2.0f)) / 320.0f)) + (5.0f / (1.0f + exp((V - 20.0f) / 5.0f))) + 3.0f;

    const double r_inf = 1.0f / (1.0f + exp((20.0f - V) / 6.0f));
    const double tau_r = (9.5f * exp((-pow(V + 40.0f, 2.0f)) / 1800.0f)) + 0.8f;

    /* calcium handling (:284-316) */
    const double kcasr = max_sr - ((max_sr - min_sr) / (1.0f + pow(EC / Ca_SR, 2.0f)));
    const double k1 = k1_prime / kcasr;
    const double k3 = 0.06;
    const double O = (k1 * pow(Ca_ss, 2.0f) * Rp) / (k3 + (k1 * pow(Ca_ss, 2.0f)));
    const double Irel = V_rel * O * (Ca_SR - Ca_ss);
    const double K_up = 0.00025;
    const double Iup = Vmax_up / (1.0f + (pow(K_up, 2.0f) / pow(Ca_i, 2.0f)));
    const double Ileak = (double)0.00036f * (Ca_SR - Ca_i);
    const double Ixfer = (double)0.0038f * (Ca_ss - Ca_i);
    const double k2 = (double)0.045f * kcasr;
    const double k4 = 0.005f;
    const double Buf_c = 0.2f, K_buf_c = 0.001f, K_buf_sr = 0.3f, Buf_sr = 10.0f;
    const double Buf_ss = 0.4f, K_buf_ss = 0.00025f, V_sr = 0.001094f, V_ss = 5.468e-05f;
    const double bufc = 1.0f / (1.0f + ((Buf_c * K_buf_c) / pow(Ca_i + K_buf_c, 2)));
    const double bufsr = 1.0f / (1.0f + ((Buf_sr * K_buf_sr) / pow(Ca_SR + K_buf_sr, 2)));
    const double bufss = 1.0f / (1.0f + ((Buf_ss * K_buf_ss) / pow(Ca_ss + K_buf_ss, 2.0f)));

    /* time derivatives (:318-325) */
    const double dV = -(IK1 + Ito + IKr + IKs + ICaL + INaK + INa + IbNa + INaCa + IbCa + IpK + IpCa + stim);
    const double dRp = ((-k2) * Ca_ss * Rp) + (k4 * (1.0f - Rp));
    const double dCai = bufc * (((((Ileak - Iup) * V_sr) / V_c) + Ixfer) - ((((IbCa + IpCa) - (2.0f * INaCa)) * Cm) / (2.0f * V_c * F)));
    const double dCaSR = bufsr * (Iup - (Irel + Ileak));
    const double dCass = bufss * (((((-ICaL) * Cm) / (2.0f * V_ss * F)) + ((Irel * V_sr) / V_ss)) - ((Ixfer * V_c) / V_ss));
    const double dNai = ((-(INa + IbNa + (3.0f * INaK) + (3.0f * INaCa))) / (V_c * F)) * Cm;
    const double dKi = ((-((IK1 + Ito + IKr + IKs + IpK + stim) - (2.0f * INaK))) / (V_c * F)) * Cm;

    /* Rush-Larsen for the gates (:337-348), forward Euler for the rest (:112,128-133) */
    y[1] = xr1_inf + (Xr1 - xr1_inf) * exp(-dt / tau_xr1);
    y[2] = xr2_inf + (Xr2 - xr2_inf) * exp(-dt / tau_xr2);
    y[3] = xs_inf + (Xs - xs_inf) * exp(-dt / tau_xs);
    y[4] = m_inf + (m - m_inf) * exp(-dt / tau_m);
    y[5] = h_inf + (h - h_inf) * exp(-dt / tau_h);
    y[6] = j_inf + (j - j_inf) * exp(-dt / tau_j);
    y[7] = d_inf + (d - d_inf) * exp(-dt / tau_d);
    y[8] = f_inf + (f - f_inf) * exp(-dt / tau_f);
    y[9] = f2_inf + (f2 - f2_inf) * exp(-dt / tau_f2);
    y[10] = fCass_inf + (fCass - fCass_inf) * exp(-dt / tau_fCass);
    y[11] = s_inf + (s - s_inf) * exp(-dt / tau_s);
    y[12] = r_inf + (r - r_inf) * exp(-dt / tau_r);

    y[0] = dt * dV + V;
    y[13] = dt * dCai + Ca_i;
    y[14] = dt * dCaSR + Ca_SR;
    y[15] = dt * dCass + Ca_ss;
    y[16] = dt * dRp + Rp;
    y[17] = dt * dNai + Na_i;
    y[18] = dt * dKi + K_i;
}

void orc_tt2006_step(double *y, double stim, double dt) { orc_tt2006_step_scaled(y, stim, dt, 1.0, 1.0, 1.0); }
