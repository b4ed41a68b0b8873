/*
 * monoalg_hip.h - C ABI of the MI355X monodomain timestep library
 * (libmonoalg_hip.so).  Plain pointers and sizes only.
 *
 * The library is the device side of MonoAlg3D's operator-split monodomain
 * step.  Each entry point names the reference interface it replaces:
 *
 *   mono_tissue_create      solve_monodomain set-up (monodomain_solver.c:39-614):
 *                           SET_ODE_INITIAL_CONDITIONS_GPU (ode_solver.h:46,
 *                           ode_solver.c:235-286), INIT_ASSEMBLY_MATRIX
 *                           set_initial_conditions_fvm (matrix_assembly.c:27-44),
 *                           ASSEMBLY_MATRIX (assembly_matrix_config.h:15;
 *                           matrix_assembly.c:334-514), SET_SPATIAL_STIM
 *                           stim_if_x_less_than (stimuli.c:56-91) and
 *                           INIT_LINEAR_SYSTEM (linear_system_solver_config.h:18).
 *   mono_tissue_step        the body of the time loop (monodomain_solver.c:616-925):
 *                           SOLVE_MODEL_ODES (ode_solver.h:37), UPDATE_MONODOMAIN
 *                           (update_monodomain_config.h:15-19) and
 *                           SOLVE_LINEAR_SYSTEM (linear_system_solver_config.h:12-16),
 *                           all device-resident.
 *   mono_tissue_get_v /     cell->v read-back as done by the save plugins
 *   mono_tissue_get_state   (save_mesh.c:248-420) and the SV dump of
 *                           save_state.c:16-155.
 *   mono_tissue_destroy     END_LINEAR_SYSTEM + the SV free of ode_solver.c:149-158.
 *   mono_tissue_observe     the activation-time / APD bookkeeping of
 *                           save_with_activation_times (save_mesh.c:730-913).
 *   mono_ionic_run          SET_ODE_INITIAL_CONDITIONS_GPU + SOLVE_MODEL_ODES on
 *                           caller-given states and per-cell stimuli
 *                           (ode_solver.h:37-47; the GPU plugins'
 *                           solve_model_odes_gpu take a host stim array).
 *   mono_ionic_run_adaptive the same with the adaptive Euler solver
 *                           (solve_forward_euler_cpu_adpt, default_solvers.c:61-176).
 *   mono_assemble           ASSEMBLY_MATRIX rows (matrix_assembly.c:334-514) as a
 *                           dense 27-point stencil per cell.
 *   mono_cg_stencil /       SOLVE_LINEAR_SYSTEM conjugate_gradient
 *   mono_cg_csr             (linear_system_solver.c:185-200; CSR as built by
 *                           grid_to_csr, grid.c:573-649).
 *
 * Errors: every int-returning function returns 0 on success and a negative
 * value on failure; mono_last_error() then describes it (the reference's
 * plugins exit() instead, e.g. log_error_and_exit).
 *
 * Cell order is lexicographic on the structured box: cell = x + nx*(y + ny*z),
 * cell centre = ((x+0.5)dx, (y+0.5)dy, (z+0.5)dz).
 */
#ifndef MONOALG_HIP_H
#define MONOALG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ionic models (library_file of [ode_solver]) */
enum {
    MONO_MODEL_TT2006 = 0,  /* libten_tusscher_2006.so, 19 states */
    MONO_MODEL_ORD2011 = 1, /* libohara_rudy_endo_2011.so, 41 states */
    MONO_MODEL_MS2003 = 2,  /* libmitchell_shaeffer_2003.so, 2 states */
    MONO_MODEL_CRN1998 = 3  /* libcourtemanche_ramirez_nattel_1998.so, 21 states, forward or adaptive Euler */
};

/* arithmetic of the ionic step and of the diffusion solve */
enum {
    MONO_FP64 = 0,        /* FP64 states, FP64 CG (the reference's CELL_MODEL_REAL_DOUBLE) */
    MONO_FP32_REFINE = 1  /* FP32 states, FP32 CG inside FP64 defect correction */
};

/* how the conductivity tensor is given (anisotropic_sigma_assembly_matrix) */
enum {
    MONO_TENSOR_UNIFORM = 0,   /* 6 doubles */
    MONO_TENSOR_PER_LAYER = 1, /* nz x 6 doubles (fibres varying with depth) */
    MONO_TENSOR_PER_CELL = 2   /* n x 6 doubles */
};

typedef struct mono_tissue_desc {
    /* structured box of cells (initialize_grid_with_cuboid_mesh) */
    int nx, ny, nz;
    double dx, dy, dz;              /* um */
    int model;                      /* MONO_MODEL_* */
    int precision;                  /* MONO_FP64 | MONO_FP32_REFINE */
    double dt_pde, dt_ode;          /* ms; ODE sub-steps = (int)(dt_pde/dt_ode) */
    double beta, cm;                /* 1/um, uF/cm^2 ([main] defaults 0.14, 1.0) */
    /* operator: aniso = 0 -> homogeneous_sigma_assembly_matrix with the diagonal
     * of the tensor as (sigma_x, sigma_y, sigma_z); aniso = 1 ->
     * anisotropic_sigma_assembly_matrix.  Tensor components are
     * (sxx, syy, szz, sxy, sxz, syz) in mS/um. */
    int aniso;
    int tensor_source;              /* MONO_TENSOR_* */
    const double *tensor;           /* host */
    const double *g_cell;           /* optional host per-cell conductivity factor */
    const uint8_t *active;          /* optional host per-cell 0/1 (inactive cells) */
    /* linear solver ([linear_system_solver]) */
    double cg_tol;                  /* absolute r.r threshold (reference `tolerance`) */
    double cg_rel_tol;              /* > 0: stop on ||r|| <= rel * ||b|| instead */
    int cg_max_its;                 /* `max_iterations` */
    int cg_jacobi;                  /* `use_preconditioner` */
    double inner_rel;               /* FP32_REFINE: inner relative tolerance (0 = 1e-4) */
    /* stimulus ([stim_*] stim_if_x_less_than) */
    double stim_start, stim_duration, stim_current, stim_x_limit, stim_period;
    /* optional host per-cell 0/1: the cells the stimulus reaches (replaces the
     * x < stim_x_limit rule; e.g. one region per instance of a batch) */
    const uint8_t *stim_mask;
    /* optional host per-cell model scalings, 4 floats per cell: TT2006 GKr,
     * GCaL, GKs (4th unused); MS2003 tau_close (2nd-4th unused); all 1 is the
     * reference model.  No counterpart in the reference (batch sweeps). */
    const float *ionic_scale;
    /* [ode_solver] adaptive=true (ode_solver.c:308-328; CRN 1998 only): the
     * adaptive Euler solver with min dt = dt_ode and max dt = dt_pde; each step
     * advances every cell to t + dt_ode with its own step size */
    int ode_adaptive;
    double ode_abs_tol, ode_rel_tol;  /* [ode_solver] abstol / reltol (reference defaults 1e-5) */
    /* > 0: the dt of ALPHA when it differs from the loop's dt_pde.  The
     * reference raises the loop's dt_pde to dt_ode when dt_ode is larger
     * (monodomain_solver.c:512-519) while ALPHA keeps the configured dt_pde
     * (the_monodomain_solver->dt, monodomain_solver.c:1326); 0 = dt_pde. */
    double alpha_dt;
    /* optional per-instance model scalings (batches of many instances; no
     * reference counterpart): the instance of each cell (n uint16, host) and a
     * table of 4 floats per instance (n_instances x 4, host; the factors of
     * ionic_scale), staged in LDS by the ionic kernel.  Replaces ionic_scale;
     * at most 4096 instances. */
    const uint16_t *instance;
    const float *instance_scale;
    int n_instances;
} mono_tissue_desc;

typedef struct mono_stats {
    int64_t steps;                  /* loop iterations done */
    double time;                    /* current simulation time (ms) */
    int64_t cg_iterations_total;    /* sum of per-step iteration counts */
    int last_cg_iterations;         /* reference count (1 + CG iterations) */
    double last_cg_error;           /* last r.r */
    int last_outer;                 /* FP32_REFINE: outer sweeps of the last step */
    int64_t inner_iterations_total; /* CG iterations actually run (all sweeps) */
    int64_t outer_sweeps_total;
    int nan_detected;               /* the reference stops on a NaN solver error */
    int class_operator;             /* operator storage: 0 per-cell rows, 1 per-layer class table,
                                       2 neighbour-code table (one byte per cell), 3 rows formed
                                       in the solves from a 4-byte neighbour mask per cell (FP32
                                       and FP64; stored FP64 rows kept for Jacobi FP64 solves) */
    int stencil_points;             /* stored non-zeros per row */
    double ionic_ms_total;          /* device time in the ionic kernel (timing on) */
    double solve_ms_total;          /* device time in the diffusion solve (timing on) */
    int ionic_block;                /* threads per block of the ionic step kernel picked by the
                                       launch-shape tuner (0: not tuned yet, or tuning off) */
} mono_stats;

typedef struct mono_tissue mono_tissue;

const char *mono_last_error(void);
int mono_device_count(void);

int mono_tissue_create(const mono_tissue_desc *desc, mono_tissue **out);
int mono_tissue_destroy(mono_tissue *t);
/* advance nsteps loop iterations (ODE + update + solve); returns 1 if a NaN
 * solver error stopped the run (monodomain_solver.c:746-754) */
int mono_tissue_step(mono_tissue *t, int nsteps);
int mono_tissue_sync(mono_tissue *t);
int64_t mono_tissue_num_cells(const mono_tissue *t);
int mono_tissue_neq(const mono_tissue *t);
int mono_tissue_get_v(mono_tissue *t, double *host_v);               /* n doubles */
int mono_tissue_set_v(mono_tissue *t, const double *host_v);
int mono_tissue_get_state(mono_tissue *t, double *host_sv);          /* neq x n, row-major by state */
int mono_tissue_set_state(mono_tissue *t, const double *host_sv);
int mono_tissue_get_stats(const mono_tissue *t, mono_stats *out);
/* 1 = time the ionic and solve phases with HIP events (adds syncs) */
int mono_tissue_set_timing(mono_tissue *t, int on);
/* No reference counterpart (diagnostic): how often the solver replayed a
 * recorded hipGraph of the launches it enqueues ahead of its first status poll,
 * and how many such graphs it recorded.  Recording is opt-in: MONO_GRAPH=1 in
 * the environment when the tissue is created. */
int mono_tissue_graph_counts(const mono_tissue *t, int64_t *launches, int64_t *captures);
/* device pointers for zero-copy consumers: v (FP64, n), sv (neq rows of
 * `pitch` elements, FP64 or FP32 by precision), b (FP64, n) */
int mono_tissue_device_ptrs(mono_tissue *t, void **v, void **sv, size_t *pitch, void **b);

/* Checkpoint and restart: SAVE_STATE save_simulation_state
 * (save_state_config.h, save_state.c:16-155) and RESTORE_STATE
 * restore_simulation_state (restore_state_config.h, restore_state.c:45-256).
 * save writes grid_checkpoint.dat, monodomain_solver_checkpoint.dat and
 * ode_solver_checkpoint.dat in the reference's layout (states of the active
 * cells, gpu = 1 state-major, real = double) into the existing directory dir,
 * plus hip_solver_checkpoint.dat (exact stimulus start, solver schedule,
 * statistics, activation observer) with which a restart repeats the
 * uninterrupted run bit for bit.  restore reads the reference's files
 * (matching cells by centre, either state layout, double or float states) and
 * the fourth file when present; the tissue must have the same shape, spacing,
 * model, dt_pde, active cells and ODE solver kind.  Without the fourth file the
 * stimulus restarts from the configured start, as the reference's restart
 * does.  Adaptive-mesh checkpoints (a cells-to-solve list) are refused. */
int mono_tissue_save_state(mono_tissue *t, const char *dir);
int mono_tissue_restore_state(mono_tissue *t, const char *dir);

/* Activation times and APDs (save_with_activation_times, save_mesh.c:730-913),
 * kept on the device.  Call mono_tissue_observe where the reference calls its
 * save function: every print_rate iterations, before mono_tissue_step, the
 * first time at t = 0.  Thresholds in mV and ms (reference defaults -30, -83,
 * 10); up to max_events activation times and APDs are kept per cell.
 * vm_double = 0 keeps the previous Vm in float as the reference does;
 * 1 keeps it in FP64 (a sample rounding to the threshold then cannot hide an
 * activation; used to compare runs of different precision). */
int mono_tissue_observe_setup(mono_tissue *t, float activation_threshold, float apd_threshold, float time_threshold,
                              int max_events, int vm_double);
int mono_tissue_observe(mono_tissue *t);
/* n_act, n_apd: n counts; act, apd: max_events x n floats (row k = k-th event,
 * NaN where absent).  Any pointer may be NULL. */
int mono_tissue_get_activations(mono_tissue *t, int32_t *n_act, float *act, int32_t *n_apd, float *apd);
/* max_events of the observer (set up, or restored with a checkpoint); 0 when none */
int mono_tissue_observe_events(const mono_tissue *t);

/* First-action-potential tracker for precision studies (no reference
 * counterpart; save_mesh.c:754-913 records APD at a fixed threshold instead).
 * setup records every cell's current Vm as its resting potential; each
 * mono_tissue_apd_sample (after mono_tissue_step) takes one sample of Vm at
 * the current time.  Per cell: lat = first upward crossing of
 * activation_threshold, linearly interpolated between samples; vmax = the
 * peak after it; apd90 = time from lat to the downward crossing of
 * V90 = vmax - 0.9 (vmax - vrest), interpolated.  NaN where absent. */
int mono_tissue_apd_setup(mono_tissue *t, double activation_threshold);
int mono_tissue_apd_sample(mono_tissue *t);
int mono_tissue_apd_get(mono_tissue *t, double *lat, double *vmax, double *apd90);

/* Standalone ionic solve: n cells, states AoS (n x neq, FP64, in/out), per-cell
 * stimulus (n doubles, may be NULL), nsteps ODE steps of dt.  precision selects
 * the FP64 or FP32 kernel. */
int mono_ionic_run(int model, int precision, double *states, int64_t n, const double *stim, double dt, int nsteps);
/* Sensor loss and its gradient (no counterpart in the reference; for fitting
 * conductivity and ionic-parameter fields to recordings).  Runs the tissue of
 * desc for nsteps steps and records V at the nsens sensor cells every
 * record_every steps from step 0 (nrec = nsteps/record_every + 1 records,
 * traces: nrec x nsens, may be NULL).  With target (nrec x nsens) it also
 * returns loss = sum (traces - target)^2 and, when grad_g / grad_scale are
 * non-NULL, dloss/dg_cell (n doubles; g_cell = 1 where desc.g_cell is NULL)
 * and dloss/d ionic_scale[4*i] (n doubles; MS2003 tau_close factor) by the
 * discrete adjoint of the step: FP64 adjoint solves with the same operator
 * (symmetric), the forward states recomputed from checkpoints kept every
 * checkpoint_every steps (<= 0: every step's states are kept, 16 bytes per
 * cell and step, and nothing is recomputed).  Mitchell-Schaeffer, isotropic
 * uniform tensor, dt_ode = dt_pde; the forward may be MONO_FP32_REFINE. */
int mono_sensor_gradient(const mono_tissue_desc *desc, int nsteps, const int64_t *sensors, int nsens, int record_every,
                         const double *target, int checkpoint_every, double *traces, double *loss, double *grad_g,
                         double *grad_scale);

/* Standalone adaptive-Euler solve (solve_model_odes_cpu with ode_solver->adaptive,
 * default_solvers.c:9-46,61-176): advances n cells (states AoS n x neq, in/out)
 * to current_t + min_dt, each from its own time; ode_dt, ode_time_new and
 * ode_previous_dt are the reference's per-cell arrays (n doubles each, in/out;
 * start with dt = min_dt, time_new = previous_dt = 0).  CRN 1998 only. */
int mono_ionic_run_adaptive(int model, int precision, double *states, int64_t n, const double *stim, double current_t,
                            double *ode_dt, double *ode_time_new, double *ode_previous_dt, double min_dt, double max_dt,
                            double abs_tol, double rel_tol);
/* GPU side of the ODE-solver plugin ABI (the reference core's use_gpu path):
 *   mono_ode_gpu_alloc  SET_ODE_INITIAL_CONDITIONS_GPU set_model_initial_conditions_gpu
 *                       (ode_solver.h:46; e.g. ten_tusscher_2006_RS_GPU.cu:8-40):
 *                       allocates the FP64 state of n cells on the device as
 *                       neq rows of *pitch_bytes (row 0 = Vm), every cell at
 *                       the model's default initial state.
 *   mono_ode_gpu_solve  SOLVE_MODEL_ODES solve_model_odes_gpu (ode_solver.h:37;
 *                       e.g. ten_tusscher_2006_RS_GPU.cu:42-156): num_steps steps
 *                       of dt for every cell, V read from and written back to
 *                       row 0, stim a host array of n (NULL: none).
 *   mono_device_free    the core's free of a GPU sv (ode_solver.c:149-158). */
int mono_ode_gpu_alloc(int model, int64_t n, double **sv, size_t *pitch_bytes);
int mono_ode_gpu_solve(int model, double *sv, size_t pitch_bytes, int64_t n, const double *stim, double dt, int num_steps);
int mono_device_free(void *p);

int mono_model_neq(int model);
/* default initial state (neq doubles) and INITIAL_V of a model */
int mono_model_initial_state(int model, double *out);
double mono_model_initial_v(int model);

/* Assemble the operator of a desc into coef27 (27 x n doubles, slot-major,
 * slot = (oz+1)*9 + (oy+1)*3 + (ox+1), slot 13 = diagonal); alpha_out gets
 * ALPHA(beta, cm, dt_pde, dx, dy, dz). */
int mono_assemble(const mono_tissue_desc *desc, double *coef27, double *alpha_out);

/* CG on a 27-slot stencil operator (host arrays, slot-major as above):
 * solves A x = b from the warm start x.  precision: MONO_FP64 or
 * MONO_FP32_REFINE.  iterations/error follow the reference's counting. */
int mono_cg_stencil(int nx, int ny, int nz, const double *coef27, const uint8_t *active, const double *b,
                    double *x, double tol, double rel_tol, int max_its, int jacobi, int precision,
                    int *iterations, double *error);
/* CG on a CSR matrix (host arrays), FP64: create + solve + destroy below. */
int mono_cg_csr(int n, const int64_t *row_ptr, const int32_t *col, const double *val, const double *b,
                double *x, double tol, int max_its, int jacobi, int *iterations, double *error);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the solve of
 * mono_cg_csr on operands already on the device (row_ptr, col, val, b and x
 * are device pointers; x is the warm start and receives the solution) on the
 * given HIP stream (NULL: the null stream).  Every buffer of the solve lives in
 * a fixed device area that the library reserves per device when it is loaded
 * (about 40 MiB, held until the process ends), so n is limited to
 * MONO_CG_CSR_NOALLOC_MAX_ROWS; larger n is refused.  The operands are checked
 * on the device (row_ptr starts at 0 and never decreases, columns in [0, n))
 * before the solve.  The call allocates nothing (no hipMalloc, hipFree, pinned
 * or host allocation).  Calls that overlap on one device share the area and
 * race; calls on different devices do not.  Same results as mono_cg_csr, bit
 * for bit. */
#define MONO_CG_CSR_NOALLOC_MAX_ROWS (1 << 20)
int mono_cg_csr_device_noalloc(int n, const int64_t *row_ptr, const int32_t *col, const double *val, const double *b,
                               double *x, double tol, int max_its, int jacobi, int *iterations, double *error,
                               void *stream);

/* A CSR operator resident on the device between solves.
 *   mono_csr_create   INIT_LINEAR_SYSTEM (linear_system_solver_config.h:18) as
 *                     init_gpu_conjugate_gradient does it (gpu_solvers_cublas_12.c:46-119):
 *                     the matrix (rows as grid_to_csr builds them, grid.c:573-649)
 *                     is uploaded once; the Jacobi diagonal is taken on the device.
 *   mono_csr_solve    SOLVE_LINEAR_SYSTEM conjugate_gradient (linear_system_solver.c:47-200):
 *                     x (host, n) is the warm start and the result; b (host, n), or
 *                     NULL for the right-hand side mono_csr_update_rhs left on the
 *                     device.  Only b and x cross PCIe.
 *   mono_csr_set_alpha / mono_csr_update_rhs
 *                     UPDATE_MONODOMAIN update_monodomain_default
 *                     (update_monodomain_config.h:15-19, update_monodomain.c:19-56):
 *                     b = alpha * Vm formed on the device, for the next mono_csr_solve.
 *   mono_csr_destroy  END_LINEAR_SYSTEM (linear_system_solver_config.h:21). */
typedef struct mono_csr mono_csr;
int mono_csr_create(int n, const int64_t *row_ptr, const int32_t *col, const double *val, mono_csr **out);
int mono_csr_num_rows(const mono_csr *A);
int mono_csr_set_alpha(mono_csr *A, const double *alpha);
int mono_csr_update_rhs(mono_csr *A, const double *vm);
int mono_csr_get_rhs(mono_csr *A, double *b);
int mono_csr_solve(mono_csr *A, const double *b, double *x, double tol, int max_its, int jacobi, int *iterations,
                   double *error);
int mono_csr_destroy(mono_csr *A);

/* UPDATE_MONODOMAIN update_monodomain_default (update_monodomain.c:19-56) on host
 * arrays: b[i] = vm[i] * alpha[i], alpha = ALPHA(beta, cm, dt_pde, dx, dy, dz)
 * of each cell (constants.h:10-11). */
int mono_update_rhs(int64_t n, const double *alpha, const double *vm, double *b);

#ifdef __cplusplus
}
#endif

#endif
