/*
 * Drives an ionic-model library through the GPU half of the reference's
 * ODE-solver ABI as the core's use_gpu path calls it (ode_solver.c:179-286,
 * 288-353): init_cell_model_data, set_model_initial_conditions_gpu (device sv,
 * the returned pitch stored in solver->pitch), then one solve_model_odes_gpu
 * per step with a host stimulus array, the state read back row by row.  The
 * reference's core here is built without a GPU, so this is how a GPU plugin is
 * exercised through its exported symbols.
 *
 *   gpu_abi_driver <model.so> <n> <nsteps> <dt> <stim_in> <out> [record_every]
 *
 * stim_in: nsteps*n doubles (step-major).  out: the n x neq state (cell-major,
 * as the CPU path's sv) after every record_every steps.
 */
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ode_solver/ode_solver.h"

static void die(const char *msg) {
    fprintf(stderr, "gpu_abi_driver: %s\n", msg);
    exit(2);
}

static void *must_sym(void *h, const char *name) {
    void *p = dlsym(h, name);
    if(!p) {
        fprintf(stderr, "gpu_abi_driver: missing symbol %s\n", name);
        exit(2);
    }
    return p;
}

int main(int argc, char **argv) {
    if(argc < 7) die("usage: gpu_abi_driver <model.so> <n> <nsteps> <dt> <stim_in> <out> [record_every]");
    const uint32_t n = (uint32_t)atol(argv[2]);
    const long nsteps = atol(argv[3]);
    const double dt = atof(argv[4]);
    const long rec = argc > 7 ? atol(argv[7]) : nsteps;
    if(n == 0 || nsteps < 1 || rec < 1) die("bad sizes");
    void *h = dlopen(argv[1], RTLD_NOW);
    if(!h) die(dlerror());
    get_cell_model_data_fn *get_data = must_sym(h, "init_cell_model_data");
    set_ode_initial_conditions_gpu_fn *set_ic = must_sym(h, "set_model_initial_conditions_gpu");
    solve_model_ode_gpu_fn *solve = must_sym(h, "solve_model_odes_gpu");

    struct ode_solver *s = calloc(1, sizeof(struct ode_solver));
    if(!s) die("out of memory");
    get_data(&s->model_data, true, true);
    const int neq = s->model_data.number_of_ode_equations;
    s->gpu = true;
    s->original_num_cells = n;
    s->num_cells_to_solve = n;
    s->min_dt = dt;
    s->num_steps = 1;
    s->pitch = set_ic(s, NULL);
    if(!s->sv || s->pitch < n * sizeof(double)) die("set_model_initial_conditions_gpu gave no state");

    const size_t nstim = (size_t)nsteps * n;
    double *stim = malloc(sizeof(double) * nstim);
    double *rows = malloc(sizeof(double) * (size_t)neq * n);
    double *aos = malloc(sizeof(double) * (size_t)neq * n);
    if(!stim || !rows || !aos) die("out of memory");
    FILE *fi = fopen(argv[5], "rb");
    if(!fi || fread(stim, sizeof(double), nstim, fi) != nstim) die("cannot read the stimulus");
    fclose(fi);
    FILE *fo = fopen(argv[6], "wb");
    if(!fo) die("cannot open the output");
    double t = 0.0;
    for(long k = 0; k < nsteps; k++) {
        solve(s, NULL, (real)t, stim + (size_t)k * n);
        t += dt;
        if((k + 1) % rec) continue;
        if(hipMemcpy2D(rows, n * sizeof(double), s->sv, s->pitch, n * sizeof(double), (size_t)neq, hipMemcpyDeviceToHost) !=
           hipSuccess)
            die("hipMemcpy2D failed");
        for(uint32_t i = 0; i < n; i++)
            for(int q = 0; q < neq; q++) aos[(size_t)i * neq + q] = rows[(size_t)q * n + i];
        fwrite(aos, sizeof(double), (size_t)neq * n, fo);
    }
    fclose(fo);
    /* the core frees a GPU sv with the device free (ode_solver.c:149-158) */
    if(hipFree(s->sv) != hipSuccess) die("hipFree failed");
    printf("neq=%d pitch=%zu\n", neq, s->pitch);
    return 0;
}
