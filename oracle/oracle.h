/*
 * CPU oracle for the MonoAlg3D operator-split monodomain step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and by nothing else.  It is a plain-C
 * restatement of the reference's CPU algorithm (FP64, glibc libm), each
 * function citing the reference file:line it follows.
 *
 * Layouts used here (all lexicographic, cell = x + nx*(y + ny*z)):
 *   - ionic state: array-of-structures, sv[cell*neq + s] (the reference CPU
 *     layout, ten_tusscher_2006_RS_CPU.c:20,94);
 *   - stencil operator: 27 coefficients per cell, coef[cell*27 + slot] with
 *     slot = (oz+1)*9 + (oy+1)*3 + (ox+1) for neighbour offset (ox,oy,oz);
 *     slot 13 is the diagonal.
 */
#ifndef MONO_ORACLE_H
#define MONO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum orc_model { ORC_TT2006 = 0, ORC_ORD2011 = 1, ORC_MS2003 = 2, ORC_CRN1998 = 3 };

int    orc_model_neq(int model);
double orc_model_initial_v(int model);
/* default initial conditions for n cells (AoS) */
void   orc_model_init(int model, double *sv, int64_t n);
/* num_steps explicit steps of size dt for n cells, per-cell stimulus (AoS) */
void   orc_model_solve(int model, double *sv, int64_t n, const double *stim,
                       double dt, int num_steps);
/* the same with per-cell model scalings, 4 floats per cell (TT2006: GKr, GCaL,
 * GKs; MS2003: tau_close; ignored for ORd); scale == NULL is all 1 */
void   orc_model_solve_scaled(int model, double *sv, int64_t n, const double *stim,
                              const float *scale, double dt, int num_steps);

/* the adaptive Euler solver (solve_forward_euler_cpu_adpt, default_solvers.c:
 * 61-176) for n cells to final_time; aux is (n, 3) AoS: dt, time_new,
 * previous_dt per cell.  Models with a plain RHS only (CRN 1998).  Returns the
 * number of steps accepted at min_dt with error >= 1, or -1 for another model. */
int64_t orc_model_solve_adaptive(int model, double *sv, int64_t n, const double *stim, double final_time, double *aux,
                                 double min_dt, double max_dt, double abs_tol, double rel_tol);

/* ALPHA(beta, cm, dt, dx, dy, dz) of monodomain/constants.h:10-11 */
double orc_alpha(double beta, double cm, double dt, double dx, double dy, double dz);

/* Isotropic finite-volume operator (homogeneous_sigma_assembly_matrix,
 * matrix_assembly.c:334-387 via assembly_common.c:4-76,911-1048).
 * sigma: per-cell (sx,sy,sz) triples; active: per-cell 0/1 (NULL = all). */
void orc_assemble_iso(int nx, int ny, int nz, double dx, double dy, double dz,
                      double alpha, const double *sigma3, const uint8_t *active,
                      double *coef27);

/* Anisotropic operator (anisotropic_sigma_assembly_matrix,
 * matrix_assembly.c:389-514 via assembly_common.c:115-830).
 * tensor6: per-cell (sxx, syy, szz, sxy, sxz, syz). */
void orc_assemble_aniso(int nx, int ny, int nz, double dx, double dy, double dz,
                        double alpha, const double *tensor6, const uint8_t *active,
                        double *coef27);

/* Conductivity tensor from a fibre direction (calc_tensor2,
 * assembly_common.c:1121-1139). out6 = (sxx, syy, szz, sxy, sxz, syz). */
void orc_fibre_tensor(const double f[3], double sigma_l, double sigma_t, double out6[6]);
/* In-place normalisation as read_fibers does (assembly_common.c:1141-1146). */
void orc_normalise(double f[3]);

/* y = A x on the 27-slot operator (inactive cells produce 0) */
void orc_stencil_apply(int nx, int ny, int nz, const double *coef27,
                       const uint8_t *active, const double *x, double *y);

/* Conjugate gradient with the reference's semantics
 * (cpu_conjugate_gradient, linear_system_solver.c:47-183): x is the warm start
 * and the result, stop when r.r <= tol (absolute, squared) or it >= max_its.
 * Returns the iteration count as the reference counts it; *err = last r.r. */
int orc_cg_stencil(int nx, int ny, int nz, const double *coef27, const uint8_t *active,
                   const double *b, double *x, double tol, int max_its, int jacobi,
                   double *err);

/* The same CG on a CSR matrix (rows in any order). */
int orc_cg_csr(int n, const int64_t *row_ptr, const int32_t *col, const double *val,
               const double *b, double *x, double tol, int max_its, int jacobi,
               double *err);

/* ---------------------------------------------------------------------
 * Full operator-split tissue run on a structured box
 * (solve_monodomain, monodomain_solver.c:616-925).
 * --------------------------------------------------------------------- */
struct orc_tissue {
    int nx, ny, nz;
    double dx, dy, dz;       /* um */
    int model;
    double dt;               /* dt_pde = dt_ode, ms */
    double beta, cm;
    double cg_tol;           /* absolute squared residual */
    int cg_max_its;
    int cg_jacobi;
    /* stimulus: cells with centre x < stim_x_limit get stim_current while
     * stim_start <= t <= stim_start + stim_duration (ode_solver.c:319-327) */
    double stim_start, stim_duration, stim_current, stim_x_limit, stim_period;
    /* state */
    const double *coef27;    /* operator, owned by caller */
    const uint8_t *active;   /* NULL = all active */
    double *sv;              /* AoS state, n*neq */
    double *v;               /* PDE potential (cell->v), n */
    double *b;               /* rhs, n */
    double cur_time;
    int64_t count;
    int64_t total_cg_its;
    int last_cg_its;
    double last_cg_err;
    /* optional (no reference counterpart; batch sweeps): per-cell 0/1 stimulated
     * cells replacing the x < stim_x_limit rule, and per-cell model scalings
     * (4 floats per cell, see orc_model_solve_scaled) */
    const uint8_t *stim_mask;
    const float *ionic_scale;
    /* adaptive Euler ODE solver (ode_solver.c:308-328): each step advances the
     * cells to cur_time + ode_min_dt with their own step sizes (aux: n x 3
     * dt, time_new, previous_dt, initialised by orc_tissue_init) */
    int adaptive;
    double ode_min_dt, ode_abs_tol, ode_rel_tol;
    double *aux;
};

/* set_initial_conditions_fvm (matrix_assembly.c:27-44) + model ICs */
void orc_tissue_init(struct orc_tissue *t);
/* One loop iteration of solve_monodomain (monodomain_solver.c:616-925):
 * Vm mirror (t>0), stimulus merge, ODE, update (b = alpha*Vm), CG. */
void orc_tissue_step(struct orc_tissue *t);

#ifdef __cplusplus
}
#endif

#endif
