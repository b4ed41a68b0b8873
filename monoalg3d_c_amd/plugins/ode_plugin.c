/*
 * Ionic-model plugin for the reference's ODE-solver ABI, backed by
 * libmonoalg_hip.so.  Built four times (MONO_PLUGIN_MODEL = 0, 1, 2, 3) into
 * libmono_hip_tt2006.so, libmono_hip_ord2011.so, libmono_hip_ms2003.so and
 * libmono_hip_crn1998.so; the reference selects one with
 *     [ode_solver]  library_file=<path>/libmono_hip_<model>.so  use_gpu=no
 * and dlsym's the three symbols below (ode_solver.c:179-233).  CRN 1998 also
 * takes [ode_solver] adaptive=true: the per-cell step-size arrays are
 * allocated here as the reference's default-solver models do
 * (courtemanche_ramirez_nattel_1998.c:12-24) and passed to the adaptive kernel.
 *
 * Compiled against the reference's own headers (struct ode_solver and the
 * GET_CELL_MODEL_DATA / SET_ODE_INITIAL_CONDITIONS_CPU / SOLVE_MODEL_ODES
 * macros, ode_solver.h:33-87) with CELL_MODEL_REAL_DOUBLE, so `real` is double
 * and sv is the reference's array-of-structures layout sv[cell*neq + s].
 * With the CPU core the states live on the host; each call uploads them, runs
 * the HIP kernel and downloads them (the device-resident whole step is
 * mono_tissue_step).  The GPU pair (set_model_initial_conditions_gpu /
 * solve_model_odes_gpu, the core's use_gpu path) keeps the states on the
 * device between calls as the reference's GPU models do: SoA rows of `pitch`
 * bytes, row 0 = Vm.
 */
#include <stdio.h>
#include <stdlib.h>

#include "ode_solver/ode_solver.h"
#include "monoalg_hip.h"

#ifndef MONO_PLUGIN_MODEL
#error "define MONO_PLUGIN_MODEL (0 tt2006, 1 ord2011, 2 ms2003, 3 crn1998)"
#endif

static void fail(const char *what) {
    fprintf(stderr, "libmono_hip: %s failed: %s\n", what, mono_last_error());
    exit(EXIT_FAILURE); /* the reference's plugins exit on errors */
}

/* GET_CELL_MODEL_DATA (ode_solver.h:33) */
GET_CELL_MODEL_DATA(init_cell_model_data) {
    if(get_initial_v) cell_model->initial_v = (real)mono_model_initial_v(MONO_PLUGIN_MODEL);
    if(get_neq) cell_model->number_of_ode_equations = mono_model_neq(MONO_PLUGIN_MODEL);
}

/* SET_ODE_INITIAL_CONDITIONS_CPU (ode_solver.h:42): the plugin allocates sv
 * (the core frees it, ode_solver.c:149-158) */
SET_ODE_INITIAL_CONDITIONS_CPU(set_model_initial_conditions_cpu) {
    (void)ode_extra_config;
    const int neq = mono_model_neq(MONO_PLUGIN_MODEL);
    double ic[64];
    if(mono_model_initial_state(MONO_PLUGIN_MODEL, ic)) fail("mono_model_initial_state");
    const uint32_t n = solver->original_num_cells;
    solver->sv = (real *)malloc(sizeof(real) * (size_t)neq * n);
    if(!solver->sv) {
        fprintf(stderr, "libmono_hip: out of host memory\n");
        exit(EXIT_FAILURE);
    }
    for(uint32_t i = 0; i < n; i++)
        for(int s = 0; s < neq; s++) solver->sv[(size_t)i * neq + s] = ic[s];
    if(solver->adaptive) {
        if(MONO_PLUGIN_MODEL != MONO_MODEL_CRN1998) {
            fprintf(stderr, "libmono_hip: adaptive=true needs a default-solver model (libmono_hip_crn1998.so)\n");
            exit(EXIT_FAILURE);
        }
        solver->ode_dt = (real *)malloc(sizeof(real) * n);
        solver->ode_previous_dt = (real *)calloc(n, sizeof(real));
        solver->ode_time_new = (real *)calloc(n, sizeof(real));
        if(!solver->ode_dt || !solver->ode_previous_dt || !solver->ode_time_new) {
            fprintf(stderr, "libmono_hip: out of host memory\n");
            exit(EXIT_FAILURE);
        }
        for(uint32_t i = 0; i < n; i++) solver->ode_dt[i] = solver->min_dt;
    }
}

/* SOLVE_MODEL_ODES (ode_solver.h:37): num_steps steps of min_dt for every cell,
 * stim_currents in cell order (cells_to_solve is only set by adaptive runs) */
SOLVE_MODEL_ODES(solve_model_odes_cpu) {
    (void)ode_extra_config;
    if(ode_solver->cells_to_solve) {
        fprintf(stderr, "libmono_hip: adaptive cell lists are not supported\n");
        exit(EXIT_FAILURE);
    }
    if(ode_solver->adaptive) {
        /* one adaptive call to current_t + min_dt (default_solvers.c:9-46) */
        if(mono_ionic_run_adaptive(MONO_PLUGIN_MODEL, MONO_FP64, ode_solver->sv, (int64_t)ode_solver->num_cells_to_solve,
                                   stim_currents, current_t, ode_solver->ode_dt, ode_solver->ode_time_new,
                                   ode_solver->ode_previous_dt, ode_solver->min_dt, ode_solver->max_dt, ode_solver->abs_tol,
                                   ode_solver->rel_tol))
            fail("mono_ionic_run_adaptive");
        return;
    }
    if(mono_ionic_run(MONO_PLUGIN_MODEL, MONO_FP64, ode_solver->sv, (int64_t)ode_solver->num_cells_to_solve, stim_currents,
                      ode_solver->min_dt, (int)ode_solver->num_steps))
        fail("mono_ionic_run");
}

/* SET_ODE_INITIAL_CONDITIONS_GPU (ode_solver.h:46; e.g. ten_tusscher_2006_RS_GPU.cu:8-40):
 * device sv, every cell at the default initial state; returns the row pitch
 * in bytes (the core stores it in solver->pitch and frees sv at the end) */
SET_ODE_INITIAL_CONDITIONS_GPU(set_model_initial_conditions_gpu) {
    (void)ode_extra_config;
    if(solver->adaptive) {
        fprintf(stderr, "libmono_hip: the GPU ODE path runs fixed steps (adaptive=true runs through the CPU symbols)\n");
        exit(EXIT_FAILURE);
    }
    double *sv = NULL;
    size_t pitch = 0;
    if(mono_ode_gpu_alloc(MONO_PLUGIN_MODEL, (int64_t)solver->original_num_cells, &sv, &pitch)) fail("mono_ode_gpu_alloc");
    solver->sv = sv;
    return pitch;
}

/* SOLVE_MODEL_ODES on the device (solve_model_odes_gpu, e.g.
 * ten_tusscher_2006_RS_GPU.cu:42-156): num_steps steps of min_dt for every
 * cell, stim_currents a host array in cell order */
SOLVE_MODEL_ODES(solve_model_odes_gpu) {
    (void)ode_extra_config;
    (void)current_t;
    if(ode_solver->cells_to_solve) {
        fprintf(stderr, "libmono_hip: adaptive cell lists are not supported\n");
        exit(EXIT_FAILURE);
    }
    if(mono_ode_gpu_solve(MONO_PLUGIN_MODEL, ode_solver->sv, ode_solver->pitch, (int64_t)ode_solver->num_cells_to_solve,
                          stim_currents, ode_solver->min_dt, (int)ode_solver->num_steps))
        fail("mono_ode_gpu_solve");
}
