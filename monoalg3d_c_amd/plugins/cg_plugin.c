/*
 * Linear-solver plugin for the reference's SOLVE_LINEAR_SYSTEM ABI, backed by
 * libmonoalg_hip.so (CSR conjugate gradient on the GPU).  The reference selects
 * it with
 *     [linear_system_solver]  library_file=<path>/libmono_hip_monodomain.so
 *                             main_function=mono_conjugate_gradient
 *                             tolerance=...  max_iterations=...  use_preconditioner=...
 * and binds init_mono_conjugate_gradient / end_mono_conjugate_gradient by name
 * (config_common.c:53-63).  Same parameters and semantics as the reference's
 * conjugate_gradient (linear_system_solver.c:38-200): x0 = cell->v, stop when
 * r.r <= tolerance or the count (from 1) reaches max_iterations, optional
 * Jacobi preconditioner, *error = final r.r, result written to cell->v.
 *
 * INIT_LINEAR_SYSTEM builds the CSR of the active cells once (rows in the
 * core's order, columns from each element's neighbour, as grid_to_csr does,
 * grid.c:573-649) and uploads it to the device, keeping the handle in
 * config->persistent_data as init_gpu_conjugate_gradient does
 * (gpu_solvers_cublas_12.c:46-119); a solve moves only b and x.  When the
 * [update_monodomain] section names this same library, its update leaves b on
 * the device (mono_update_plugin.c) and only x crosses PCIe.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "3dparty/stb_ds.h"
#include "alg/grid/grid.h"
#include "config/linear_system_solver_config.h"
#include "monoalg_hip.h"
#include "mono_plugin_common.h"

struct mono_cg_data mono_cg_state = {0};

static void release(struct mono_cg_data *s) {
    if(s->A && mono_csr_destroy(s->A)) fprintf(stderr, "libmono_hip: mono_csr_destroy failed: %s\n", mono_last_error());
    free(s->bvec);
    free(s->xvec);
    memset(s, 0, sizeof(*s));
}

/* value of a key of the plugin's section, or NULL; config_data is the core's
 * string hash (an stb_ds array of key/value pairs, common_types.h:138-141) */
const char *mono_plugin_param(const struct config *config, const char *key) {
    const struct string_hash_entry *h = config->config_data;
    for(ptrdiff_t i = 0; h && i < shlen(h); i++)
        if(h[i].key && strcmp(h[i].key, key) == 0) return h[i].value;
    return NULL;
}

static void build(struct mono_cg_data *s, uint32_t n, struct cell_node **cells) {
    int64_t nnz = 0;
    for(uint32_t i = 0; i < n; i++) nnz += (int64_t)arrlen(cells[i]->elements);
    int64_t *row_ptr = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    double *val = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    s->bvec = (double *)malloc(sizeof(double) * (n ? n : 1));
    s->xvec = (double *)malloc(sizeof(double) * (n ? n : 1));
    if(!row_ptr || !col || !val || !s->bvec || !s->xvec) mono_plugin_die("out of host memory");
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
    if(n > 0 && mono_csr_create((int)n, row_ptr, col, val, &s->A)) mono_plugin_fail("mono_csr_create");
    free(row_ptr);
    free(col);
    free(val);
    s->n = n;
    s->cells = cells;
}

INIT_LINEAR_SYSTEM(init_mono_conjugate_gradient) {
    if(is_purkinje) mono_plugin_die("the Purkinje network is not supported");
    if(the_grid->adaptive) mono_plugin_die("mesh adaptivity is not supported (the CSR is uploaded once)");
    struct mono_cg_data *s = &mono_cg_state;
    release(s);
    /* the reference reads the same keys with the same defaults
     * (linear_system_solver.c:11-15,38-42); booleans as its IS_TRUE does */
    s->tol = 1e-16;
    s->max_its = 200;
    s->jacobi = false;
    const char *v;
    if((v = mono_plugin_param(config, "tolerance"))) s->tol = strtod(v, NULL);
    if((v = mono_plugin_param(config, "max_iterations"))) s->max_its = (int)strtod(v, NULL);
    if((v = mono_plugin_param(config, "use_preconditioner"))) s->jacobi = mono_plugin_is_true(v);
    build(s, the_grid->num_active_cells, the_grid->active_cells);
    s->initialised = true;
    config->persistent_data = s;
}

END_LINEAR_SYSTEM(end_mono_conjugate_gradient) {
    if(config->persistent_data == &mono_cg_state) config->persistent_data = NULL;
    release(&mono_cg_state);
}

SOLVE_LINEAR_SYSTEM(mono_conjugate_gradient) {
    (void)time_info;
    (void)the_grid;
    struct mono_cg_data *s = (struct mono_cg_data *)config->persistent_data;
    if(!s || !s->initialised)
        mono_plugin_die("mono_conjugate_gradient needs init_function=init_mono_conjugate_gradient in [linear_system_solver]");
    if(num_active_cells != s->n || active_cells != s->cells) {
        /* another set of cells than the one uploaded (the core re-runs
         * INIT_LINEAR_SYSTEM after a remesh, monodomain_solver.c:838-839) */
        const double tol = s->tol;
        const int max_its = s->max_its;
        const bool jacobi = s->jacobi;
        release(s);
        s->tol = tol;
        s->max_its = max_its;
        s->jacobi = jacobi;
        build(s, num_active_cells, active_cells);
        s->initialised = true;
    }
    if(num_active_cells == 0) {
        *number_of_iterations = 0;
        *error = 0.0;
        return;
    }
    /* b: left on the device by this library's UPDATE_MONODOMAIN, or cell->b */
    const double *b = NULL;
    if(!s->b_on_device) {
        for(uint32_t i = 0; i < num_active_cells; i++) s->bvec[i] = active_cells[i]->b;
        b = s->bvec;
    }
    s->b_on_device = false;
    for(uint32_t i = 0; i < num_active_cells; i++) s->xvec[i] = active_cells[i]->v;
    int its = 0;
    double err = 0.0;
    if(mono_csr_solve(s->A, b, s->xvec, s->tol, s->max_its, s->jacobi ? 1 : 0, &its, &err)) mono_plugin_fail("mono_csr_solve");
    for(uint32_t i = 0; i < num_active_cells; i++) active_cells[i]->v = s->xvec[i];
    *number_of_iterations = (uint32_t)its;
    *error = err;
}
