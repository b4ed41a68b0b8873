/*
 * Golden-vector harness over the reference's own compiled CPU plugins.
 * TEST INFRASTRUCTURE ONLY: built by oracle/ref.mk into oracle/_ref/bin from
 * this file plus the reference's headers and core objects; it drives the
 * reference plugins through their dlopen ABI exactly as the reference core
 * does, and writes raw little-endian binary records that
 * tests/golden/make_golden.py turns into committed fixtures.
 *
 * Modes
 *   ionic <model.so> <n> <nsteps> <dt> <states_in|-> <stim_in> <out> [record_every]
 *       set_model_initial_conditions_cpu (or load n*neq doubles from states_in),
 *       then nsteps calls of solve_model_odes_cpu with the per-cell stimulus
 *       from stim_in (nsteps*n doubles, step-major).  Writes the n*neq state
 *       after every record_every steps (default: only the last).
 *       Driver shape: ode_solver.c:235-353 (num_steps = 1, cells_to_solve NULL).
 *   ionic_adpt <model.so> <n> <nsteps> <dt_pde> <min_dt> <abs_tol> <rel_tol> <states_in|-> <stim_in> <out> [record_every]
 *       The same with the adaptive Euler solver (ode_solver->adaptive, max_dt =
 *       dt_pde, ode_solver.c:308-328; num_steps = 1, monodomain_solver.c:510-520):
 *       each call advances every cell to t + min_dt.  Records the n*neq state
 *       followed by the per-cell dt, time_new and previous_dt (3*n doubles).
 *   assemble <iso|aniso> <nx> <ny> <nz> <h> <cells_in> <out> [sx sy sz | sl st sn]
 *       Cuboid mesh through set_cuboid_domain_mesh (domain_helpers.c:19-73),
 *       optional deactivation and per-cell fibres from cells_in (per
 *       lexicographic cell: uint8 active, double f[3]); then the assembly plugin
 *       (homogeneous_sigma_assembly_matrix or anisotropic_sigma_assembly_matrix
 *       with fibers_in_mesh) and a dump of every active row as
 *       (int32 row_lex, int32 nelem, nelem x (int32 col_lex, double value)).
 *   dirmap <out>
 *       Lexicographic offsets of the 26 neighbour directions of an interior
 *       cell (cell.h:21-49 enum order), as 26 x int32[3].
 */
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "alg/grid/grid.h"
#include "config/assembly_matrix_config.h"
#include "config/config_common.h"
#include "monodomain/monodomain_solver.h"
#include "ode_solver/ode_solver.h"
#include "3dparty/stb_ds.h"

typedef int (*cuboid_fn)(struct grid *, real_cpu, real_cpu, real_cpu, real_cpu, real_cpu, real_cpu);

static void die(const char *msg) {
    fprintf(stderr, "ref_harness: %s\n", msg);
    exit(2);
}

static void *must_sym(void *h, const char *name) {
    void *p = dlsym(h, name);
    if(!p) {
        fprintf(stderr, "ref_harness: missing symbol %s\n", name);
        exit(2);
    }
    return p;
}

static double *read_doubles(const char *path, size_t n) {
    double *p = malloc(n * sizeof(double));
    FILE *f = fopen(path, "rb");
    if(!f) die("cannot open input");
    if(fread(p, sizeof(double), n, f) != n) die("short input");
    fclose(f);
    return p;
}

static int mode_ionic(int argc, char **argv) {
    if(argc < 9) die("ionic <model.so> <n> <nsteps> <dt> <states_in|-> <stim_in> <out> [record_every]");
    const char *lib = argv[2];
    uint32_t n = (uint32_t)atol(argv[3]);
    long nsteps = atol(argv[4]);
    double dt = atof(argv[5]);
    const char *states_in = argv[6];
    const char *stim_in = argv[7];
    const char *out = argv[8];
    long rec = argc > 9 ? atol(argv[9]) : nsteps;

    void *h = dlopen(lib, RTLD_LAZY);
    if(!h) die(dlerror());
    get_cell_model_data_fn *get_data = must_sym(h, "init_cell_model_data");
    set_ode_initial_conditions_cpu_fn *set_ic = must_sym(h, "set_model_initial_conditions_cpu");
    solve_model_ode_cpu_fn *solve = must_sym(h, "solve_model_odes_cpu");

    struct ode_solver *s = calloc(1, sizeof(struct ode_solver));
    get_data(&s->model_data, true, true);
    int neq = s->model_data.number_of_ode_equations;
    s->original_num_cells = n;
    s->num_cells_to_solve = n;
    s->cells_to_solve = NULL;
    s->min_dt = dt;
    s->num_steps = 1;
    set_ic(s, NULL);
    if(strcmp(states_in, "-") != 0) {
        double *st = read_doubles(states_in, (size_t)n * neq);
        memcpy(s->sv, st, (size_t)n * neq * sizeof(double));
        free(st);
    }
    double *stim = read_doubles(stim_in, (size_t)nsteps * n);
    FILE *fo = fopen(out, "wb");
    if(!fo) die("cannot open output");
    double t = 0.0;
    for(long k = 0; k < nsteps; k++) {
        solve(s, NULL, (real)t, stim + (size_t)k * n);
        t += dt;
        if((k + 1) % rec == 0) fwrite(s->sv, sizeof(double), (size_t)n * neq, fo);
    }
    fclose(fo);
    printf("neq=%d initial_v=%.17g\n", neq, (double)s->model_data.initial_v);
    return 0;
}

static int mode_ionic_adpt(int argc, char **argv) {
    if(argc < 12) die("ionic_adpt <model.so> <n> <nsteps> <dt_pde> <min_dt> <abs_tol> <rel_tol> <states_in|-> <stim_in> <out> [rec]");
    const char *lib = argv[2];
    uint32_t n = (uint32_t)atol(argv[3]);
    long nsteps = atol(argv[4]);
    double dt_pde = atof(argv[5]);
    const char *states_in = argv[9];
    const char *stim_in = argv[10];
    const char *out = argv[11];
    long rec = argc > 12 ? atol(argv[12]) : nsteps;

    void *h = dlopen(lib, RTLD_LAZY);
    if(!h) die(dlerror());
    get_cell_model_data_fn *get_data = must_sym(h, "init_cell_model_data");
    set_ode_initial_conditions_cpu_fn *set_ic = must_sym(h, "set_model_initial_conditions_cpu");
    solve_model_ode_cpu_fn *solve = must_sym(h, "solve_model_odes_cpu");

    struct ode_solver *s = calloc(1, sizeof(struct ode_solver));
    get_data(&s->model_data, true, true);
    int neq = s->model_data.number_of_ode_equations;
    s->original_num_cells = n;
    s->num_cells_to_solve = n;
    s->cells_to_solve = NULL;
    s->adaptive = true;
    s->max_dt = (real)dt_pde;
    s->min_dt = (real)atof(argv[6]);
    s->abs_tol = (real)atof(argv[7]);
    s->rel_tol = (real)atof(argv[8]);
    s->num_steps = 1;
    set_ic(s, NULL);
    if(strcmp(states_in, "-") != 0) {
        double *st = read_doubles(states_in, (size_t)n * (neq + 3));
        memcpy(s->sv, st, (size_t)n * neq * sizeof(double));
        memcpy(s->ode_dt, st + (size_t)n * neq, n * sizeof(double));
        memcpy(s->ode_time_new, st + (size_t)n * (neq + 1), n * sizeof(double));
        memcpy(s->ode_previous_dt, st + (size_t)n * (neq + 2), n * sizeof(double));
        free(st);
    }
    double *stim = read_doubles(stim_in, (size_t)nsteps * n);
    FILE *fo = fopen(out, "wb");
    if(!fo) die("cannot open output");
    double t = 0.0;
    for(long k = 0; k < nsteps; k++) {
        solve(s, NULL, (real)t, stim + (size_t)k * n);
        t += dt_pde;
        if((k + 1) % rec == 0) {
            fwrite(s->sv, sizeof(double), (size_t)n * neq, fo);
            fwrite(s->ode_dt, sizeof(double), n, fo);
            fwrite(s->ode_time_new, sizeof(double), n, fo);
            fwrite(s->ode_previous_dt, sizeof(double), n, fo);
        }
    }
    fclose(fo);
    printf("neq=%d initial_v=%.17g\n", neq, (double)s->model_data.initial_v);
    return 0;
}

static int lex_of(const struct cell_node *c, double h, int nx, int ny) {
    int ix = (int)floor(c->center.x / h), iy = (int)floor(c->center.y / h), iz = (int)floor(c->center.z / h);
    return ix + nx * (iy + ny * iz);
}

static struct grid *make_cuboid(const char *domains_lib, int nx, int ny, int nz, double h) {
    void *hd = dlopen(domains_lib, RTLD_LAZY);
    if(!hd) die(dlerror());
    cuboid_fn cub = must_sym(hd, "set_cuboid_domain_mesh");
    struct grid *g = new_grid();
    if(!cub(g, h, h, h, nx * h, ny * h, nz * h)) die("cuboid mesh failed");
    return g;
}

static int mode_assemble(int argc, char **argv) {
    if(argc < 9) die("assemble <iso|aniso> <nx> <ny> <nz> <h> <cells_in|-> <out> [s1 s2 s3]");
    int aniso = strcmp(argv[2], "aniso") == 0;
    int nx = atoi(argv[3]), ny = atoi(argv[4]), nz = atoi(argv[5]);
    double h = atof(argv[6]);
    const char *cells_in = argv[7];
    const char *out = argv[8];
    const char *s1 = argc > 9 ? argv[9] : "0.00013";
    const char *s2 = argc > 10 ? argv[10] : s1;
    const char *s3 = argc > 11 ? argv[11] : s2;

    struct grid *g = make_cuboid("./shared_libs/libdefault_domains.so", nx, ny, nz, h);
    int ncell = nx * ny * nz;
    if(strcmp(cells_in, "-") != 0) {
        FILE *f = fopen(cells_in, "rb");
        if(!f) die("cannot open cells_in");
        unsigned char *act = malloc(ncell);
        double *fib = malloc(sizeof(double) * 3 * ncell);
        for(int i = 0; i < ncell; i++) {
            if(fread(act + i, 1, 1, f) != 1 || fread(fib + 3 * i, sizeof(double), 3, f) != 3) die("short cells_in");
        }
        fclose(f);
        FOR_EACH_CELL(g) {
            if(cell->active) {
                int l = lex_of(cell, h, nx, ny);
                if(!act[l]) cell->active = false;
                cell->sigma.fibers.f[0] = fib[3 * l];
                cell->sigma.fibers.f[1] = fib[3 * l + 1];
                cell->sigma.fibers.f[2] = fib[3 * l + 2];
            }
        }
        free(act);
        free(fib);
    }
    order_grid_cells(g);

    struct monodomain_solver *ms = calloc(1, sizeof(struct monodomain_solver));
    ms->beta = 0.14;
    ms->cm = 1.0;
    ms->dt = 0.02;

    struct config *cfg = alloc_and_init_config_data();
    if(aniso) {
        cfg->main_function_name = "anisotropic_sigma_assembly_matrix";
        shput(cfg->config_data, "sigma_l", strdup(s1));
        shput(cfg->config_data, "sigma_t", strdup(s2));
        shput(cfg->config_data, "sigma_n", strdup(s3));
        shput(cfg->config_data, "fibers_in_mesh", "true");
    } else {
        cfg->main_function_name = "homogeneous_sigma_assembly_matrix";
        shput(cfg->config_data, "sigma_x", strdup(s1));
        shput(cfg->config_data, "sigma_y", strdup(s2));
        shput(cfg->config_data, "sigma_z", strdup(s3));
    }
    init_config_functions(cfg, "./shared_libs/libdefault_matrix_assembly.so", "assembly_matrix");
    ((assembly_matrix_fn *)cfg->main_function)(cfg, ms, g);

    FILE *fo = fopen(out, "wb");
    if(!fo) die("cannot open output");
    for(uint32_t i = 0; i < g->num_active_cells; i++) {
        struct cell_node *c = g->active_cells[i];
        int32_t row = lex_of(c, h, nx, ny);
        int32_t ne = (int32_t)arrlen(c->elements);
        fwrite(&row, 4, 1, fo);
        fwrite(&ne, 4, 1, fo);
        for(int e = 0; e < ne; e++) {
            int32_t col = lex_of(c->elements[e].cell, h, nx, ny);
            double v = c->elements[e].value;
            fwrite(&col, 4, 1, fo);
            fwrite(&v, 8, 1, fo);
        }
    }
    fclose(fo);
    printf("active=%u\n", g->num_active_cells);
    return 0;
}

static int mode_dirmap(int argc, char **argv) {
    if(argc < 3) die("dirmap <out>");
    const double h = 100.0;
    struct grid *g = make_cuboid("./shared_libs/libdefault_domains.so", 4, 4, 4, h);
    order_grid_cells(g);
    struct cell_node *c = NULL;
    for(uint32_t i = 0; i < g->num_active_cells; i++) {
        struct cell_node *k = g->active_cells[i];
        if(lex_of(k, h, 4, 4) == 1 + 4 * (1 + 4 * 1)) c = k;
    }
    if(!c) die("centre cell not found");
    FILE *fo = fopen(argv[2], "wb");
    for(int d = 0; d < NUM_DIRECTIONS; d++) {
        struct cell_node *nb = get_cell_neighbour_with_same_refinement_level(c, d);
        int32_t off[3] = {99, 99, 99};
        if(nb) {
            off[0] = (int32_t)lround((nb->center.x - c->center.x) / h);
            off[1] = (int32_t)lround((nb->center.y - c->center.y) / h);
            off[2] = (int32_t)lround((nb->center.z - c->center.z) / h);
        }
        fwrite(off, 4, 3, fo);
    }
    fclose(fo);
    return 0;
}

int main(int argc, char **argv) {
    if(argc < 2) die("mode: ionic | assemble | dirmap");
    if(strcmp(argv[1], "ionic") == 0) return mode_ionic(argc, argv);
    if(strcmp(argv[1], "ionic_adpt") == 0) return mode_ionic_adpt(argc, argv);
    if(strcmp(argv[1], "assemble") == 0) return mode_assemble(argc, argv);
    if(strcmp(argv[1], "dirmap") == 0) return mode_dirmap(argc, argv);
    die("unknown mode");
    return 2;
}
