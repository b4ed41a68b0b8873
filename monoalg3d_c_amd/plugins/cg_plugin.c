/*
 * Linear-solver plugin for the reference's SOLVE_LINEAR_SYSTEM ABI, backed by
 * libmonoalg_hip.so (CSR conjugate gradient on the GPU).  The reference selects
 * it with
 *     [linear_system_solver]  library_file=<path>/libmono_hip_cg.so
 *                             main_function=mono_conjugate_gradient
 *                             tolerance=...  max_iterations=...  use_preconditioner=...
 * and binds init_mono_conjugate_gradient / end_mono_conjugate_gradient by name
 * (config_common.c:53-63).  Same parameters and semantics as the reference's
 * conjugate_gradient (linear_system_solver.c:38-200): x0 = cell->v, stop when
 * r.r <= tolerance or the count (from 1) reaches max_iterations, optional
 * Jacobi preconditioner, *error = final r.r, result written to cell->v.
 *
 * Rows are the active cells in the core's order; columns come from each
 * element's neighbour (element.cell->grid_position), as grid_to_csr does
 * (grid.c:573-649).  The CSR is built once per grid size.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "3dparty/stb_ds.h"
#include "alg/grid/grid.h"
#include "config/linear_system_solver_config.h"
#include "monoalg_hip.h"

static real_cpu tol = 1e-16;
static int max_its = 200;
static bool use_preconditioner = false;
static uint32_t nrows = 0;
static int64_t *row_ptr = NULL;
static int32_t *col = NULL;
static double *val = NULL, *bvec = NULL, *xvec = NULL;

static void release(void) {
    free(row_ptr);
    free(col);
    free(val);
    free(bvec);
    free(xvec);
    row_ptr = NULL;
    col = NULL;
    val = bvec = xvec = NULL;
    nrows = 0;
}

/* value of a [linear_system_solver] key, or NULL; config_data is the core's
 * string hash (an stb_ds array of key/value pairs, common_types.h:138-141) */
static const char *param(const struct config *config, const char *key) {
    const struct string_hash_entry *h = config->config_data;
    for(ptrdiff_t i = 0; h && i < shlen(h); i++)
        if(h[i].key && strcmp(h[i].key, key) == 0) return h[i].value;
    return NULL;
}

INIT_LINEAR_SYSTEM(init_mono_conjugate_gradient) {
    (void)the_grid;
    (void)is_purkinje;
    /* the reference reads the same keys (linear_system_solver.c:38-42); plain
     * numbers are accepted here, booleans as its IS_TRUE does */
    const char *v;
    if((v = param(config, "tolerance"))) tol = strtod(v, NULL);
    if((v = param(config, "max_iterations"))) max_its = (int)strtod(v, NULL);
    if((v = param(config, "use_preconditioner")))
        use_preconditioner = strcmp(v, "true") == 0 || strcmp(v, "yes") == 0 || strcmp(v, "1") == 0;
    release();
}

END_LINEAR_SYSTEM(end_mono_conjugate_gradient) {
    (void)config;
    release();
}

static void build_csr(uint32_t n, struct cell_node **cells) {
    release();
    int64_t nnz = 0;
    for(uint32_t i = 0; i < n; i++) nnz += (int64_t)arrlen(cells[i]->elements);
    row_ptr = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    val = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    bvec = (double *)malloc(sizeof(double) * (n ? n : 1));
    xvec = (double *)malloc(sizeof(double) * (n ? n : 1));
    if(!row_ptr || !col || !val || !bvec || !xvec) {
        fprintf(stderr, "libmono_hip_cg: out of host memory\n");
        exit(EXIT_FAILURE);
    }
    int64_t k = 0;
    for(uint32_t i = 0; i < n; i++) {
        row_ptr[i] = k;
        const struct element *e = cells[i]->elements;
        const size_t ne = arrlen(e);
        for(size_t j = 0; j < ne; j++) {
            col[k] = (int32_t)e[j].cell->grid_position;
            val[k] = e[j].value;
            k++;
        }
    }
    row_ptr[n] = k;
    nrows = n;
}

SOLVE_LINEAR_SYSTEM(mono_conjugate_gradient) {
    (void)time_info;
    (void)config;
    (void)the_grid;
    if(!row_ptr || nrows != num_active_cells) build_csr(num_active_cells, active_cells);
    for(uint32_t i = 0; i < num_active_cells; i++) {
        bvec[i] = active_cells[i]->b;
        xvec[i] = active_cells[i]->v;
    }
    int its = 0;
    double err = 0.0;
    if(mono_cg_csr((int)num_active_cells, row_ptr, col, val, bvec, xvec, tol, max_its, use_preconditioner ? 1 : 0, &its, &err)) {
        fprintf(stderr, "libmono_hip_cg: mono_cg_csr failed: %s\n", mono_last_error());
        exit(EXIT_FAILURE);
    }
    for(uint32_t i = 0; i < num_active_cells; i++) active_cells[i]->v = xvec[i];
    *number_of_iterations = (uint32_t)its;
    *error = err;
}
