/*
 * UPDATE_MONODOMAIN plugin over libmonoalg_hip.so, for
 *     [update_monodomain]  library_file=<path>/libmono_hip_monodomain.so
 *                          main_function=mono_update_monodomain
 * Same result as update_monodomain_default (update_monodomain.c:19-56):
 * b_i = Vm_i * ALPHA(beta, cm, dt_pde, dx, dy, dz) of cell i (constants.h:10-11),
 * Vm read from the ODE state (row sv_position of the CPU solver's
 * array-of-structures sv).  The product is formed on the device.  When the
 * [linear_system_solver] section names this same library (its CG is
 * initialised on these cells), b stays on the device for the next solve;
 * otherwise it is written to cell->b for whichever solver runs.
 */
#include <stdio.h>
#include <stdlib.h>

#include "alg/grid/grid.h"
#include "config/update_monodomain_config.h"
#include "monoalg_hip.h"
#include "mono_plugin_common.h"
#include "ode_solver/ode_solver.h"

static double *vm_buf = NULL, *alpha_buf = NULL, *b_buf = NULL;
static uint32_t buf_n = 0;

static void ensure(uint32_t n) {
    if(buf_n >= n && vm_buf) return;
    free(vm_buf);
    free(alpha_buf);
    free(b_buf);
    vm_buf = (double *)malloc(sizeof(double) * (n ? n : 1));
    alpha_buf = (double *)malloc(sizeof(double) * (n ? n : 1));
    b_buf = (double *)malloc(sizeof(double) * (n ? n : 1));
    if(!vm_buf || !alpha_buf || !b_buf) mono_plugin_die("out of host memory");
    buf_n = n;
}

UPDATE_MONODOMAIN(mono_update_monodomain) {
    (void)time_info;
    (void)config;
    (void)the_grid;
    (void)initial_number_of_cells;
    if(the_ode_solver->gpu)
        mono_plugin_die("mono_update_monodomain reads the CPU ODE solver's states (the reference core here runs use_gpu=no)");
    const real_cpu beta = the_solver->beta, cm = the_solver->cm, dt_pde = the_solver->dt;
    const int neq = the_ode_solver->model_data.number_of_ode_equations;
    const real *sv = the_ode_solver->sv;
    ensure(num_active_cells);
    for(uint32_t i = 0; i < num_active_cells; i++) vm_buf[i] = sv[(size_t)active_cells[i]->sv_position * neq];

    struct mono_cg_data *cg = &mono_cg_state;
    const bool linked = cg->initialised && cg->A && cg->n == num_active_cells && cg->cells == active_cells;
    if(linked) {
        if(!cg->alpha_set) {
            for(uint32_t i = 0; i < num_active_cells; i++)
                alpha_buf[i] = ALPHA(beta, cm, dt_pde, active_cells[i]->discretization.x, active_cells[i]->discretization.y,
                                     active_cells[i]->discretization.z);
            if(mono_csr_set_alpha(cg->A, alpha_buf)) mono_plugin_fail("mono_csr_set_alpha");
            cg->alpha_set = true;
        }
        if(mono_csr_update_rhs(cg->A, vm_buf)) mono_plugin_fail("mono_csr_update_rhs");
        cg->b_on_device = true;
        return;
    }
    for(uint32_t i = 0; i < num_active_cells; i++)
        alpha_buf[i] = ALPHA(beta, cm, dt_pde, active_cells[i]->discretization.x, active_cells[i]->discretization.y,
                             active_cells[i]->discretization.z);
    if(num_active_cells && mono_update_rhs((int64_t)num_active_cells, alpha_buf, vm_buf, b_buf)) mono_plugin_fail("mono_update_rhs");
    for(uint32_t i = 0; i < num_active_cells; i++) active_cells[i]->b = b_buf[i];
}
