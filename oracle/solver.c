/*
 * Oracle: conjugate gradient and the operator-split time loop, FP64.
 * TEST INFRASTRUCTURE ONLY.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* reference direction order used for the off-diagonal sum (cell.h:21-49) */
static const int kDirOff[26][3] = {
    {0, 0, 1},  {0, 0, -1}, {0, 1, 0},   {0, -1, 0},  {1, 0, 0},    {-1, 0, 0},   {0, 1, 1},
    {0, -1, 1}, {1, 0, 1},  {-1, 0, 1},  {1, 1, 1},   {-1, 1, 1},   {1, -1, 1},   {-1, -1, 1},
    {0, 1, -1}, {0, -1, -1}, {1, 0, -1}, {-1, 0, -1}, {1, 1, -1},   {-1, 1, -1},  {1, -1, -1},
    {-1, -1, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0},  {-1, -1, 0}};

/* Row product as the reference forms it: Ax = 0; Ax += value*x over the row's
 * elements, diagonal first (linear_system_solver.c:78-85). */
static inline double row_apply(int nx, int ny, int nz, const double *row, int64_t c, const double *x) {
    const int xi = (int)(c % nx), yi = (int)((c / nx) % ny), zi = (int)(c / ((int64_t)nx * ny));
    double acc = 0.0;
    acc += row[13] * x[c];
    for(int d = 0; d < 26; d++) {
        const int ox = kDirOff[d][0], oy = kDirOff[d][1], oz = kDirOff[d][2];
        const double a = row[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)];
        if(a == 0.0) continue;
        const int xx = xi + ox, yy = yi + oy, zz = zi + oz;
        if(xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
        acc += a * x[xx + (int64_t)nx * (yy + (int64_t)ny * zz)];
    }
    return acc;
}

void orc_stencil_apply(int nx, int ny, int nz, const double *coef27, const uint8_t *active, const double *x,
                       double *y) {
    const int64_t n = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for(int64_t c = 0; c < n; c++) y[c] = (active && !active[c]) ? 0.0 : row_apply(nx, ny, nz, coef27 + 27 * c, c, x);
}

/* Generic CG with the reference's control flow (cpu_conjugate_gradient,
 * linear_system_solver.c:47-183).  Only active rows take part. */
typedef void (*apply_fn)(const void *ctx, const double *x, double *y);
typedef double (*diag_fn)(const void *ctx, int64_t i);

static int cg_core(int64_t n, const uint8_t *active, apply_fn apply, diag_fn diag, const void *ctx, const double *b,
                   double *x, double tol, int max_its, int jacobi, double *err) {
    double *r = calloc(n, sizeof(double)), *p = calloc(n, sizeof(double));
    double *q = calloc(n, sizeof(double)), *z = calloc(n, sizeof(double));
    int its = 1;
    *err = 1.0;
    double rTr = 0.0, rTz = 0.0;
    apply(ctx, x, q);
    for(int64_t i = 0; i < n; i++) {
        if(active && !active[i]) continue;
        r[i] = b[i] - q[i];
        if(jacobi) {
            double d = diag(ctx, i);
            if(d == 0.0) d = 1.0;
            z[i] = (1.0 / d) * r[i];
            rTz += r[i] * z[i];
            p[i] = z[i];
        } else {
            p[i] = r[i];
        }
        rTr += r[i] * r[i];
    }
    *err = rTr;
    if(*err >= tol) {
        while(its < max_its) {
            apply(ctx, p, q);
            double pTAp = 0.0;
            for(int64_t i = 0; i < n; i++)
                if(!active || active[i]) pTAp += p[i] * q[i];
            const double alpha = jacobi ? rTz / pTAp : rTr / pTAp;
            double r1Tr1 = 0.0, r1Tz1 = 0.0;
            for(int64_t i = 0; i < n; i++) {
                if(active && !active[i]) continue;
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
                if(jacobi) {
                    double d = diag(ctx, i);
                    if(d == 0.0) d = 1.0;
                    z[i] = (1.0 / d) * r[i];
                    r1Tz1 += z[i] * r[i];
                }
                r1Tr1 += r[i] * r[i];
            }
            const double beta = jacobi ? r1Tz1 / rTz : r1Tr1 / rTr;
            *err = r1Tr1;
            its++;
            if(*err <= tol) break;
            for(int64_t i = 0; i < n; i++) {
                if(active && !active[i]) continue;
                p[i] = (jacobi ? z[i] : r[i]) + beta * p[i];
            }
            rTz = r1Tz1;
            rTr = r1Tr1;
        }
    }
    free(r);
    free(p);
    free(q);
    free(z);
    return its;
}

struct stencil_ctx {
    int nx, ny, nz;
    const double *coef27;
    const uint8_t *active;
};

static void stencil_apply_cb(const void *ctx, const double *x, double *y) {
    const struct stencil_ctx *s = ctx;
    orc_stencil_apply(s->nx, s->ny, s->nz, s->coef27, s->active, x, y);
}

static double stencil_diag_cb(const void *ctx, int64_t i) {
    const struct stencil_ctx *s = ctx;
    return s->coef27[27 * i + 13];
}

int orc_cg_stencil(int nx, int ny, int nz, const double *coef27, const uint8_t *active, const double *b, double *x,
                   double tol, int max_its, int jacobi, double *err) {
    struct stencil_ctx s = {nx, ny, nz, coef27, active};
    return cg_core((int64_t)nx * ny * nz, active, stencil_apply_cb, stencil_diag_cb, &s, b, x, tol, max_its, jacobi, err);
}

struct csr_ctx {
    int n;
    const int64_t *rp;
    const int32_t *col;
    const double *val;
};

static void csr_apply_cb(const void *ctx, const double *x, double *y) {
    const struct csr_ctx *s = ctx;
    for(int i = 0; i < s->n; i++) {
        double acc = 0.0;
        for(int64_t k = s->rp[i]; k < s->rp[i + 1]; k++) acc += s->val[k] * x[s->col[k]];
        y[i] = acc;
    }
}

static double csr_diag_cb(const void *ctx, int64_t i) {
    const struct csr_ctx *s = ctx;
    for(int64_t k = s->rp[i]; k < s->rp[i + 1]; k++)
        if(s->col[k] == i) return s->val[k];
    return 0.0;
}

int orc_cg_csr(int n, const int64_t *row_ptr, const int32_t *col, const double *val, const double *b, double *x,
               double tol, int max_its, int jacobi, double *err) {
    struct csr_ctx s = {n, row_ptr, col, val};
    return cg_core(n, NULL, csr_apply_cb, csr_diag_cb, &s, b, x, tol, max_its, jacobi, err);
}

/* ------------------------------------------------------------------ */
/* time loop                                                          */
/* ------------------------------------------------------------------ */

void orc_tissue_init(struct orc_tissue *t) {
    const int64_t n = (int64_t)t->nx * t->ny * t->nz;
    const double alpha = orc_alpha(t->beta, t->cm, t->dt, t->dx, t->dy, t->dz);
    const double v0 = orc_model_initial_v(t->model);
    orc_model_init(t->model, t->sv, n);
    /* set_initial_conditions_fvm (matrix_assembly.c:27-44) */
    for(int64_t i = 0; i < n; i++) {
        t->v[i] = v0;
        t->b[i] = v0 * alpha;
    }
    /* adaptive solver rows (the models' set_model_initial_conditions_cpu, e.g.
     * courtemanche_ramirez_nattel_1998.c:12-24): dt = min_dt, the rest 0 */
    if(t->adaptive && t->aux)
        for(int64_t i = 0; i < n; i++) {
            t->aux[3 * i] = t->ode_min_dt;
            t->aux[3 * i + 1] = 0.0;
            t->aux[3 * i + 2] = 0.0;
        }
    t->cur_time = 0.0;
    t->count = 0;
    t->total_cg_its = 0;
}

void orc_tissue_step(struct orc_tissue *t) {
    const int64_t n = (int64_t)t->nx * t->ny * t->nz;
    const int neq = orc_model_neq(t->model);
    const double alpha = orc_alpha(t->beta, t->cm, t->dt, t->dx, t->dy, t->dz);
    /* Vm mirror from the previous solve, skipped at t = 0 (monodomain_solver.c:660-669,
     * 1012-1105) */
    if(t->cur_time > 0.0) {
        for(int64_t i = 0; i < n; i++)
            if(!t->active || t->active[i]) t->sv[i * neq] = t->v[i];
    }
    /* stimulus merge (ode_solver.c:288-341): one sub-step, window inclusive on the
     * accumulated double time, periodic start advanced after the step */
    double *stim = calloc(n, sizeof(double));
    const double time = t->cur_time;
    if((time >= t->stim_start) && (time <= t->stim_start + t->stim_duration)) {
        for(int64_t i = 0; i < n; i++) {
            const int xi = (int)(i % t->nx);
            const double cx = xi * t->dx + t->dx / 2.0;
            if(t->stim_mask ? t->stim_mask[i] != 0 : cx < t->stim_x_limit) stim[i] += t->stim_current;
        }
    }
    /* the shifted start is written back into the configuration as "%lf" text
     * and parsed again (ode_solver.c:224-230), so it keeps six decimals */
    const double sub_dt = t->adaptive ? t->ode_min_dt : t->dt;
    if(t->stim_period > 0.0 && time + sub_dt >= t->stim_start + t->stim_period) {
        char buf[64];
        snprintf(buf, sizeof(buf), "%lf", t->stim_start + t->stim_period);
        t->stim_start = strtod(buf, NULL);
    }
    /* ODE for the active cells */
    if(t->adaptive) {
        for(int64_t i = 0; i < n; i++)
            if(!t->active || t->active[i])
                orc_model_solve_adaptive(t->model, t->sv + i * neq, 1, stim + i, time + t->ode_min_dt, t->aux + 3 * i,
                                         t->ode_min_dt, t->dt, t->ode_abs_tol, t->ode_rel_tol);
    } else if(!t->active) {
        orc_model_solve_scaled(t->model, t->sv, n, stim, t->ionic_scale, t->dt, 1);
    } else {
        for(int64_t i = 0; i < n; i++)
            if(t->active[i])
                orc_model_solve_scaled(t->model, t->sv + i * neq, 1, stim + i,
                                       t->ionic_scale ? t->ionic_scale + 4 * i : 0, t->dt, 1);
    }
    free(stim);
    /* update_monodomain_default (update_monodomain.c:19-56) */
    for(int64_t i = 0; i < n; i++)
        if(!t->active || t->active[i]) t->b[i] = t->sv[i * neq] * alpha;
    /* diffusion solve, warm start from the previous v */
    t->last_cg_its = orc_cg_stencil(t->nx, t->ny, t->nz, t->coef27, t->active, t->b, t->v, t->cg_tol, t->cg_max_its,
                                    t->cg_jacobi, &t->last_cg_err);
    t->total_cg_its += t->last_cg_its;
    t->count++;
    t->cur_time += t->dt;
}
