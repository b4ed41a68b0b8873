/*
 * ASSEMBLY_MATRIX plugin over libmonoalg_hip.so, for
 *     [assembly_matrix]  library_file=<path>/libmono_hip_monodomain.so
 *                        init_function=mono_set_initial_conditions_fvm
 *                        main_function=mono_homogeneous_sigma_assembly_matrix
 *                                   or mono_anisotropic_sigma_assembly_matrix
 * with the reference's keys (sigma_x/y/z; sigma_l/t/n with fibers_file,
 * fibers_in_mesh or f/s/n).  The rows are assembled on the device
 * (mono_assemble, the finite-volume rules of matrix_assembly.c:334-514 and
 * assembly_common.c:4-1188) over the structured box that holds the active
 * cells, and written back as each cell's element list:
 *   - homogeneous: diagonal first, then the face neighbours in the order the
 *     reference's loop inserts them when it runs on one thread (cells in the
 *     core's order, faces BACK, FRONT, TOP, DOWN, RIGHT, LEFT, each pair
 *     entered into both cells' lists, fill_discretization_matrix_elements,
 *     assembly_common.c:911-1048), the diagonal accumulated in that order;
 *   - anisotropic: diagonal first, then the 26 directions in enum order
 *     (cell.h:21-49), non-zero values only (fill_elements_aniso,
 *     assembly_common.c:437-830).
 * The grid must be uniform (one discretization, no refinement): the reference's
 * anisotropic assembly refuses adaptive meshes as well.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "3dparty/stb_ds.h"
#include "alg/grid/grid.h"
#include "config/assembly_matrix_config.h"
#include "monoalg_hip.h"
#include "mono_plugin_common.h"
#include "monodomain/monodomain_solver.h"
#include "ode_solver/ode_solver.h"

/* neighbour offsets (x, y, z) of the reference's 26 directions (cell.h:21-49):
 * FRONT/BACK = +z/-z, TOP/DOWN = +y/-y, RIGHT/LEFT = +x/-x, compounds summed */
static const int kDir[26][3] = {
    {0, 0, 1},  {0, 0, -1},  {0, 1, 0},   {0, -1, 0}, {1, 0, 0},   {-1, 0, 0},  {0, 1, 1},   {0, -1, 1},   {1, 0, 1},
    {-1, 0, 1}, {1, 1, 1},   {-1, 1, 1},  {1, -1, 1}, {-1, -1, 1}, {0, 1, -1},  {0, -1, -1}, {1, 0, -1},   {-1, 0, -1},
    {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}, {1, 1, 0}, {-1, 1, 0},  {1, -1, 0},  {-1, -1, 0}};
enum { D_FRONT = 0, D_BACK = 1, D_TOP = 2, D_DOWN = 3, D_RIGHT = 4, D_LEFT = 5 };

static int slot_of(int ox, int oy, int oz) { return (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1); }

/* The structured box holding the active cells: lexicographic index of each
 * active cell and the cell at each box position (NULL where none). */
struct box {
    int nx, ny, nz;
    double dx, dy, dz;
    int64_t n;
    int64_t *lex;             /* per active cell */
    struct cell_node **at;    /* per box cell */
};

static void box_free(struct box *b) {
    free(b->lex);
    free(b->at);
    memset(b, 0, sizeof(*b));
}

static void box_of(struct grid *g, struct box *b) {
    memset(b, 0, sizeof(*b));
    const uint32_t na = g->num_active_cells;
    struct cell_node **ac = g->active_cells;
    if(na == 0) mono_plugin_die("no active cells to assemble");
    if(g->adaptive) mono_plugin_die("mesh adaptivity is not supported by the device assembly");
    b->dx = ac[0]->discretization.x;
    b->dy = ac[0]->discretization.y;
    b->dz = ac[0]->discretization.z;
    double lo[3] = {ac[0]->center.x, ac[0]->center.y, ac[0]->center.z}, hi[3] = {lo[0], lo[1], lo[2]};
    for(uint32_t i = 0; i < na; i++) {
        const struct cell_node *c = ac[i];
        if(c->discretization.x != b->dx || c->discretization.y != b->dy || c->discretization.z != b->dz)
            mono_plugin_die("the device assembly needs one discretization for every active cell");
        const double p[3] = {c->center.x, c->center.y, c->center.z};
        for(int a = 0; a < 3; a++) {
            if(p[a] < lo[a]) lo[a] = p[a];
            if(p[a] > hi[a]) hi[a] = p[a];
        }
    }
    const double h[3] = {b->dx, b->dy, b->dz};
    int L[3];
    for(int a = 0; a < 3; a++) L[a] = (int)llround((hi[a] - lo[a]) / h[a]) + 1;
    b->nx = L[0];
    b->ny = L[1];
    b->nz = L[2];
    b->n = (int64_t)L[0] * L[1] * L[2];
    if(b->n > ((int64_t)1 << 31)) mono_plugin_die("the box of the active cells is too large");
    b->lex = (int64_t *)malloc(sizeof(int64_t) * na);
    b->at = (struct cell_node **)calloc((size_t)b->n, sizeof(struct cell_node *));
    if(!b->lex || !b->at) mono_plugin_die("out of host memory");
    for(uint32_t i = 0; i < na; i++) {
        const struct cell_node *c = ac[i];
        const double p[3] = {c->center.x, c->center.y, c->center.z};
        int q[3];
        for(int a = 0; a < 3; a++) {
            const double f = (p[a] - lo[a]) / h[a];
            q[a] = (int)llround(f);
            if(fabs(f - q[a]) > 1e-6) mono_plugin_die("active cell centres are not on one lattice");
        }
        const int64_t k = q[0] + (int64_t)b->nx * (q[1] + (int64_t)b->ny * q[2]);
        if(b->at[k]) mono_plugin_die("two active cells share a centre");
        b->at[k] = ac[i];
        b->lex[i] = k;
    }
}

/* mono_assemble over the box: 27-slot rows (slot-major, 27 x n) */
static double *assemble_box(const struct box *b, struct monodomain_solver *solver, int aniso, int tensor_source,
                            const double *tensor) {
    uint8_t *act = (uint8_t *)malloc((size_t)b->n);
    double *coef = (double *)malloc(sizeof(double) * 27 * (size_t)b->n);
    if(!act || !coef) mono_plugin_die("out of host memory");
    for(int64_t k = 0; k < b->n; k++) act[k] = b->at[k] ? 1 : 0;
    mono_tissue_desc d;
    memset(&d, 0, sizeof(d));
    d.nx = b->nx;
    d.ny = b->ny;
    d.nz = b->nz;
    d.dx = b->dx;
    d.dy = b->dy;
    d.dz = b->dz;
    d.model = MONO_MODEL_TT2006; /* not used by the assembly */
    d.precision = MONO_FP64;
    d.dt_pde = solver->dt;
    d.beta = solver->beta;
    d.cm = solver->cm;
    d.aniso = aniso;
    d.tensor_source = tensor_source;
    d.tensor = tensor;
    d.active = act;
    double alpha = 0.0;
    if(mono_assemble(&d, coef, &alpha)) mono_plugin_fail("mono_assemble");
    free(act);
    return coef;
}

static struct element make_element(struct cell_node *nb, double value) {
    struct element e;
    memset(&e, 0, sizeof(e));
    e.value = value;
    e.value_ecg = 0.0;
    e.column = nb->grid_position;
    e.cell = nb;
    return e;
}

/* initialize_diagonal_elements (assembly_common.c:4-39): a fresh list holding
 * the diagonal element with value `diag` */
static void reset_elements(struct cell_node *c, double diag) {
    if(c->elements) arrfree(c->elements);
    c->elements = NULL;
    arrsetcap(c->elements, 7);
    arrput(c->elements, make_element(c, diag));
}

static bool has_column(const struct cell_node *c, uint32_t col) {
    const size_t ne = arrlen(c->elements);
    for(size_t i = 1; i < ne; i++)
        if(c->elements[i].column == col) return true;
    return false;
}

INIT_ASSEMBLY_MATRIX(mono_set_initial_conditions_fvm) {
    /* set_initial_conditions_fvm (matrix_assembly.c:27-44) */
    (void)config;
    (void)the_ode_solver;
    (void)purkinje_initial_v;
    const uint32_t na = the_grid->num_active_cells;
    struct cell_node **ac = the_grid->active_cells;
    for(uint32_t i = 0; i < na; i++) {
        const real_cpu alpha = ALPHA(the_solver->beta, the_solver->cm, the_solver->dt, ac[i]->discretization.x,
                                     ac[i]->discretization.y, ac[i]->discretization.z);
        ac[i]->v = initial_v;
        ac[i]->b = initial_v * alpha;
    }
}

static double param_num(struct config *config, const char *key, bool required, double dflt) {
    const char *v = mono_plugin_param(config, key);
    if(!v) {
        if(required) {
            fprintf(stderr, "libmono_hip: [assembly_matrix] needs %s\n", key);
            exit(EXIT_FAILURE);
        }
        return dflt;
    }
    return strtod(v, NULL);
}

ASSEMBLY_MATRIX(mono_homogeneous_sigma_assembly_matrix) {
    /* the reference reads the sigmas as `real` (matrix_assembly.c:347-354) */
    const real sx = (real)param_num(config, "sigma_x", true, 0.0);
    const real sy = (real)param_num(config, "sigma_y", true, 0.0);
    const real sz = (real)param_num(config, "sigma_z", true, 0.0);
    const uint32_t na = the_grid->num_active_cells;
    struct cell_node **ac = the_grid->active_cells;
    for(uint32_t i = 0; i < na; i++) {
        ac[i]->sigma.x = sx;
        ac[i]->sigma.y = sy;
        ac[i]->sigma.z = sz;
    }
    struct box b;
    box_of(the_grid, &b);
    const double tensor[6] = {sx, sy, sz, 0.0, 0.0, 0.0};
    double *coef = assemble_box(&b, the_solver, 0, MONO_TENSOR_UNIFORM, tensor);
    const int64_t n = b.n;
    for(uint32_t i = 0; i < na; i++)
        reset_elements(ac[i], ALPHA(the_solver->beta, the_solver->cm, the_solver->dt, ac[i]->discretization.x,
                                    ac[i]->discretization.y, ac[i]->discretization.z));
    /* the reference's loop on one thread: each cell enters each face pair into
     * both lists unless already there, adding sigma*area/dist to each diagonal */
    static const int faces[6] = {D_BACK, D_FRONT, D_TOP, D_DOWN, D_RIGHT, D_LEFT};
    for(uint32_t i = 0; i < na; i++) {
        struct cell_node *c = ac[i];
        const int64_t k = b.lex[i];
        const int x = (int)(k % b.nx), y = (int)((k / b.nx) % b.ny), z = (int)(k / ((int64_t)b.nx * b.ny));
        for(int f = 0; f < 6; f++) {
            const int *o = kDir[faces[f]];
            const int xx = x + o[0], yy = y + o[1], zz = z + o[2];
            if(xx < 0 || yy < 0 || zz < 0 || xx >= b.nx || yy >= b.ny || zz >= b.nz) continue;
            const int64_t kn = xx + (int64_t)b.nx * (yy + (int64_t)b.ny * zz);
            struct cell_node *nb = b.at[kn];
            if(!nb) continue;
            /* the off-diagonal of the pair (its two rows hold the same value:
             * the harmonic mean of equal conductivities) */
            const double v = coef[(int64_t)slot_of(o[0], o[1], o[2]) * n + k];
            if(!has_column(c, nb->grid_position)) {
                c->elements[0].value += -v;
                arrput(c->elements, make_element(nb, v));
            }
            if(!has_column(nb, c->grid_position)) {
                nb->elements[0].value += -v;
                arrput(nb->elements, make_element(c, v));
            }
        }
    }
    free(coef);
    box_free(&b);
}

/* the fibre file as read_fibers reads it (assembly_common.c:1148-1188): nine
 * numbers per line (f, s, n), each vector normalised */
static double *read_fibre_file(const char *path, size_t *count) {
    FILE *fh = fopen(path, "r");
    if(!fh) {
        fprintf(stderr, "libmono_hip: cannot open fibre file %s\n", path);
        exit(EXIT_FAILURE);
    }
    double *out = NULL;
    size_t n = 0, cap = 0;
    char *line = NULL;
    size_t len = 0;
    while(getline(&line, &len, fh) != -1) {
        double v[9] = {0};
        char *p = line;
        for(int j = 0; j < 9; j++) {
            char *e = NULL;
            v[j] = strtod(p, &e);
            p = e;
        }
        for(int j = 0; j < 3; j++) {
            double *u = v + 3 * j;
            const double m = sqrt((u[0] * u[0]) + (u[1] * u[1]) + (u[2] * u[2]));
            u[0] = u[0] / m;
            u[1] = u[1] / m;
            u[2] = u[2] / m;
        }
        if(n == cap) {
            cap = cap ? 2 * cap : 1024;
            out = (double *)realloc(out, sizeof(double) * 9 * cap);
            if(!out) mono_plugin_die("out of host memory");
        }
        memcpy(out + 9 * n, v, sizeof(v));
        n++;
    }
    free(line);
    fclose(fh);
    *count = n;
    return out;
}

/* calc_tensor2 (sigma_t = sigma_n) / calc_tensor (assembly_common.c:1102-1139) */
static void fibre_tensor(double D[3][3], const double *f, const double *s, const double *nv, double sl, double st, double sn) {
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++) {
            if(st == sn) D[i][j] = (f[i] * f[j]) * (sl - st) + (i == j ? st : 0.0);
            else D[i][j] = ((f[i] * f[j]) * sl + (s[i] * s[j]) * st) + (nv[i] * nv[j]) * sn;
        }
}

static void param_vec(struct config *config, const char *key, double out[3], const double dflt[3]) {
    const char *v = mono_plugin_param(config, key);
    memcpy(out, dflt, sizeof(double) * 3);
    if(!v) return;
    /* "[a, b, c]" or "a, b, c" */
    const char *p = v;
    for(int j = 0; j < 3; j++) {
        while(*p && (*p == '[' || *p == ',' || *p == ' ')) p++;
        char *e = NULL;
        out[j] = strtod(p, &e);
        p = e;
    }
}

ASSEMBLY_MATRIX(mono_anisotropic_sigma_assembly_matrix) {
    if(the_grid->adaptive) mono_plugin_die("anisotropic assembly does not support mesh adaptivity");
    const double sl = param_num(config, "sigma_l", true, 0.0);
    const double st = param_num(config, "sigma_t", true, 0.0);
    const double sn = param_num(config, "sigma_n", true, 0.0);
    const char *fibre_path = mono_plugin_param(config, "fibers_file");
    const char *fim = mono_plugin_param(config, "fibers_in_mesh");
    const bool fibres_in_mesh = fim && mono_plugin_is_true(fim);
    double *fibres = NULL;
    size_t nfib = 0;
    double f0[3], s0[3], n0[3];
    static const double df[3] = {1, 0, 0}, ds[3] = {0, 1, 0}, dn[3] = {0, 0, 1};
    if(fibre_path) fibres = read_fibre_file(fibre_path, &nfib);
    else if(!fibres_in_mesh) {
        param_vec(config, "f", f0, df);
        param_vec(config, "s", s0, ds);
        param_vec(config, "n", n0, dn);
    }
    const uint32_t na = the_grid->num_active_cells;
    struct cell_node **ac = the_grid->active_cells;
    struct box b;
    box_of(the_grid, &b);
    double *tensor = (double *)calloc(6 * (size_t)b.n, sizeof(double));
    if(!tensor) mono_plugin_die("out of host memory");
    for(uint32_t i = 0; i < na; i++) {
        struct cell_node *c = ac[i];
        double D[3][3];
        if(fibres) {
            const uint32_t idx = c->original_position_in_file;
            if(idx == (uint32_t)-1 || idx >= nfib) mono_plugin_die("a cell has no line in the fibre file");
            const double *r = fibres + 9 * (size_t)idx;
            fibre_tensor(D, r, r + 3, r + 6, sl, st, sn);
            memcpy(c->sigma.fibers.f, r, sizeof(double) * 3);
            memcpy(c->sigma.fibers.s, r + 3, sizeof(double) * 3);
            memcpy(c->sigma.fibers.n, r + 6, sizeof(double) * 3);
        } else if(fibres_in_mesh) {
            fibre_tensor(D, c->sigma.fibers.f, c->sigma.fibers.s, c->sigma.fibers.n, sl, st, sn);
        } else {
            fibre_tensor(D, f0, s0, n0, sl, st, sn);
        }
        c->sigma.x = D[0][0];
        c->sigma.y = D[1][1];
        c->sigma.z = D[2][2];
        c->sigma.xy = D[0][1];
        c->sigma.xz = D[0][2];
        c->sigma.yz = D[1][2];
        double *t = tensor + 6 * (size_t)b.lex[i];
        t[0] = D[0][0];
        t[1] = D[1][1];
        t[2] = D[2][2];
        t[3] = D[0][1];
        t[4] = D[0][2];
        t[5] = D[1][2];
    }
    free(fibres);
    double *coef = assemble_box(&b, the_solver, 1, MONO_TENSOR_PER_CELL, tensor);
    free(tensor);
    const int64_t n = b.n;
    for(uint32_t i = 0; i < na; i++) {
        struct cell_node *c = ac[i];
        const int64_t k = b.lex[i];
        reset_elements(c, coef[(int64_t)13 * n + k]);
        const int x = (int)(k % b.nx), y = (int)((k / b.nx) % b.ny), z = (int)(k / ((int64_t)b.nx * b.ny));
        for(int d = 0; d < 26; d++) {
            const int *o = kDir[d];
            const int xx = x + o[0], yy = y + o[1], zz = z + o[2];
            if(xx < 0 || yy < 0 || zz < 0 || xx >= b.nx || yy >= b.ny || zz >= b.nz) continue;
            struct cell_node *nb = b.at[xx + (int64_t)b.nx * (yy + (int64_t)b.ny * zz)];
            if(!nb) continue;
            const double v = coef[(int64_t)slot_of(o[0], o[1], o[2]) * n + k];
            if(v != 0) arrput(c->elements, make_element(nb, v));
        }
    }
    free(coef);
    box_free(&b);
}
