/*
 * Oracle: Mitchell-Schaeffer 2003 and the model dispatch.  TEST INFRASTRUCTURE
 * ONLY.
 *
 * Mitchell-Schaeffer follows models_library/mitchell_shaeffer/
 * mitchell_shaeffer_2003.c: initial state :24-25 (float literals), forward
 * Euler step :56-68, right-hand side :70-93 (the stimulus is ADDED to dV/dt,
 * :90, unlike TT2006/ORd which subtract it).
 * The per-cell loop follows solve_model_odes_cpu
 * (ten_tusscher_2006_RS_CPU.c:74-98): num_steps sub-steps with the same
 * stimulus value.
 */
#include <math.h>
#include <stdint.h>
#include "oracle.h"

void orc_tt2006_init_cell(double *y);
void orc_tt2006_step_scaled(double *y, double stim, double dt, double sKr, double sCaL, double sKs);
void orc_ord2011_init_cell(double *y);
void orc_ord2011_step(double *y, double stim, double dt);
void orc_crn1998_init_cell(double *y);
void orc_crn1998_step(double *y, double stim, double dt);
void orc_crn1998_rhs(const double *y, double *dy, double stim);
typedef void (*orc_rhs_fn)(const double *y, double *dy, double stim);
int orc_adaptive_euler(orc_rhs_fn rhs, int neq, double *y, double stim, double final_time, double *aux, double min_dt,
                       double max_dt, double abs_tol, double rel_tol);

static void ms_init_cell(double *y) {
    y[0] = (double)0.00000820413566106744f;
    y[1] = (double)0.8789655121804799f;
}

/* s_tau_close multiplies tau_close (no reference counterpart; 1 is the
 * reference model) */
static void ms_step(double *y, double stim, double dt, double s_tau_close) {
    const double V = y[0], h = y[1];
    const double tau_in = 0.3, tau_out = 6.0, V_gate = 0.13, tau_open = 120.0, tau_close = 150.0 * s_tau_close;
    const double J_in = (h * (pow(V, 2.00000) * (1.00000 - V))) / tau_in;
    const double J_out = -(V / tau_out);
    const double dV = J_out + J_in + stim;
    const double dh = (V < V_gate ? (1.00000 - h) / tau_open : -h / tau_close);
    y[0] = dt * dV + V;
    y[1] = dt * dh + h;
}

int orc_model_neq(int model) {
    switch(model) {
    case ORC_TT2006: return 19;
    case ORC_ORD2011: return 41;
    case ORC_MS2003: return 2;
    case ORC_CRN1998: return 21;
    default: return -1;
    }
}

/* INITIAL_V of each model header (ten_tusscher_2006.h:7, Ohara_Rudy_2011.h:8,
 * mitchell_shaeffer_2003.h:8) */
double orc_model_initial_v(int model) {
    switch(model) {
    case ORC_TT2006: return (double)(-85.23f);
    case ORC_ORD2011: return (double)(-85.f);
    case ORC_MS2003: return 0.00000820413566106744;
    case ORC_CRN1998: return (double)(-81.180000f); /* courtemanche_ramirez_nattel_1998.h:7 */
    default: return NAN;
    }
}

void orc_model_init(int model, double *sv, int64_t n) {
    const int neq = orc_model_neq(model);
    for(int64_t i = 0; i < n; i++) {
        double *y = sv + i * neq;
        if(model == ORC_TT2006) orc_tt2006_init_cell(y);
        else if(model == ORC_ORD2011) orc_ord2011_init_cell(y);
        else if(model == ORC_CRN1998) orc_crn1998_init_cell(y);
        else ms_init_cell(y);
    }
}

void orc_model_solve_scaled(int model, double *sv, int64_t n, const double *stim, const float *scale, double dt,
                            int num_steps) {
    const int neq = orc_model_neq(model);
#pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < n; i++) {
        double *y = sv + i * neq;
        const double st = stim ? stim[i] : 0.0;
        const double s0 = scale ? (double)scale[4 * i] : 1.0, s1 = scale ? (double)scale[4 * i + 1] : 1.0;
        const double s2 = scale ? (double)scale[4 * i + 2] : 1.0;
        for(int k = 0; k < num_steps; k++) {
            if(model == ORC_TT2006) orc_tt2006_step_scaled(y, st, dt, s0, s1, s2);
            else if(model == ORC_ORD2011) orc_ord2011_step(y, st, dt);
            else if(model == ORC_CRN1998) orc_crn1998_step(y, st, dt);
            else ms_step(y, st, dt, s0);
        }
    }
}

void orc_model_solve(int model, double *sv, int64_t n, const double *stim, double dt, int num_steps) {
    orc_model_solve_scaled(model, sv, n, stim, 0, dt, num_steps);
}

int64_t orc_model_solve_adaptive(int model, double *sv, int64_t n, const double *stim, double final_time, double *aux,
                                 double min_dt, double max_dt, double abs_tol, double rel_tol) {
    if(model != ORC_CRN1998) return -1;
    const int neq = orc_model_neq(model);
    int64_t forced = 0;
#pragma omp parallel for schedule(static) reduction(+ : forced)
    for(int64_t i = 0; i < n; i++)
        forced += orc_adaptive_euler(orc_crn1998_rhs, neq, sv + i * neq, stim ? stim[i] : 0.0, final_time, aux + 3 * i, min_dt,
                                     max_dt, abs_tol, rel_tol);
    return forced;
}
