// Device-resident operator-split tissue run and the C ABI (monoalg_hip.h).
//
// One loop iteration follows solve_monodomain (monodomain_solver.c:616-925):
//   stimulus window test on the accumulated double time (ode_solver.c:319-336)
//   -> ionic step with the fused RHS update (ionic.hip)
//   -> diffusion solve with warm start from the previous potential (cg.hip)
//   -> count++, time += dt_pde.
// Nothing crosses PCIe inside the loop except the solver's small status record.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "../../include/monoalg_hip.h"
#include "cg.h"
#include "ionic.h"
#include "mono_common.h"
#include "observer.h"
#include "operator.h"
#include "tissue_impl.h"

// cells of a box row per wave of the per-instance ||b||^2 kernel
static constexpr int kInstCols = 256;

static thread_local char g_err[1024];

void mono_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

using namespace mono;


static double alpha_of(const mono_tissue_desc &d) {
    // ALPHA (monodomain/constants.h:10-11); the 1e-8 factor is a float literal there
    const double dt = d.alpha_dt > 0 ? d.alpha_dt : d.dt_pde;
    return (((d.beta * d.cm) / dt) * 0.00000001f) * ((d.dx * d.dy) * d.dz);
}

static int validate(const mono_tissue_desc *d) {
    if(!d) {
        mono_set_error("null descriptor");
        return -1;
    }
    if(d->nx <= 0 || d->ny <= 0 || d->nz <= 0) {
        mono_set_error("box must have at least one cell per axis (got %d x %d x %d)", d->nx, d->ny, d->nz);
        return -1;
    }
    if((int64_t)d->nx * d->ny * d->nz > (int64_t)1 << 31) {
        mono_set_error("too many cells for one tissue: %lld", (long long)d->nx * d->ny * d->nz);
        return -1;
    }
    if(!(d->dx > 0) || !(d->dy > 0) || !(d->dz > 0) || !(d->dt_pde > 0)) {
        mono_set_error("dx, dy, dz and dt_pde must be positive");
        return -1;
    }
    if(d->model < 0 || d->model >= kNumModels) {
        mono_set_error("unknown ionic model %d", d->model);
        return -1;
    }
    if(d->ode_adaptive) {
        if(d->model != kCRN1998) {
            mono_set_error("the adaptive ODE solver runs CRN 1998 (model %d given)", d->model);
            return -1;
        }
        if(d->dt_ode < 0 || d->dt_ode > d->dt_pde || d->ode_abs_tol < 0 || d->ode_rel_tol < 0) {
            mono_set_error("adaptive ODE solver: need 0 <= dt_ode <= dt_pde and non-negative tolerances");
            return -1;
        }
    }
    if(d->precision != MONO_FP64 && d->precision != MONO_FP32_REFINE) {
        mono_set_error("unknown precision %d", d->precision);
        return -1;
    }
    if(!d->tensor) {
        mono_set_error("conductivity tensor missing");
        return -1;
    }
    if(d->tensor_source < 0 || d->tensor_source > 2) {
        mono_set_error("unknown tensor source %d", d->tensor_source);
        return -1;
    }
    if(d->instance || d->instance_scale) {
        if(!d->instance || !d->instance_scale || d->n_instances < 1 || d->n_instances > kMaxInstances) {
            mono_set_error("per-instance scalings need instance, instance_scale and 1 <= n_instances <= %d", kMaxInstances);
            return -1;
        }
        if(d->ionic_scale) {
            mono_set_error("give per-cell (ionic_scale) or per-instance (instance, instance_scale) scalings, not both");
            return -1;
        }
        const int64_t n = (int64_t)d->nx * d->ny * d->nz;
        for(int64_t i = 0; i < n; i++)
            if(d->instance[i] >= d->n_instances) {
                mono_set_error("cell %lld names instance %d of %d", (long long)i, (int)d->instance[i], d->n_instances);
                return -1;
            }
    }
    return 0;
}

static size_t tensor_count(const mono_tissue_desc &d, int64_t n) {
    return d.tensor_source == MONO_TENSOR_UNIFORM ? 6 : (d.tensor_source == MONO_TENSOR_PER_LAYER ? 6 * (size_t)d.nz : 6 * (size_t)n);
}

// Upload the optional per-cell inputs and assemble; coef27 is left allocated on return.
static int assemble_into(const mono_tissue_desc &d, int64_t n, double **tensor_dev, double **g_dev, uint8_t **act_dev,
                         double **coef27, AssemblyParams *ap, hipStream_t s) {
    const size_t nt = tensor_count(d, n);
    MONO_CHECK(hipMalloc(tensor_dev, nt * sizeof(double)));
    MONO_CHECK(hipMemcpyAsync(*tensor_dev, d.tensor, nt * sizeof(double), hipMemcpyHostToDevice, s));
    if(d.g_cell) {
        MONO_CHECK(hipMalloc(g_dev, n * sizeof(double)));
        MONO_CHECK(hipMemcpyAsync(*g_dev, d.g_cell, n * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if(d.active) {
        MONO_CHECK(hipMalloc(act_dev, n));
        MONO_CHECK(hipMemcpyAsync(*act_dev, d.active, n, hipMemcpyHostToDevice, s));
    }
    MONO_CHECK(hipMalloc(coef27, 27 * n * sizeof(double)));
    ap->nx = d.nx;
    ap->ny = d.ny;
    ap->nz = d.nz;
    ap->dx = d.dx;
    ap->dy = d.dy;
    ap->dz = d.dz;
    ap->alpha = alpha_of(d);
    ap->aniso = d.aniso;
    ap->tensor_source = d.tensor_source;
    ap->tensor = *tensor_dev;
    ap->g = *g_dev;
    ap->active = *act_dev;
    MONO_CHECK(launch_assemble(*ap, *coef27, s));
    return 0;
}

// adaptive call of this step: every cell to cur_time + min dt
// (solve_model_odes_cpu, default_solvers.c:9-46; max dt = dt_pde, ode_solver.c:311-328)
template <typename T> static void set_adaptive(const mono_tissue *t, IonicLaunch<T> &a) {
    a.aux = t->aux;
    a.final_time = t->cur_time + t->dt_ode;
    a.min_dt = t->dt_ode;
    a.max_dt = t->d.dt_pde;
    a.abs_tol = t->d.ode_abs_tol;
    a.rel_tol = t->d.ode_rel_tol;
}

extern "C" {

const char *mono_last_error(void) { return g_err; }

int mono_device_count(void) {
    int c = 0;
    if(hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int mono_model_neq(int model) { return (model >= 0 && model < kNumModels) ? model_neq(model) : -1; }

int mono_model_initial_state(int model, double *out) {
    if(model < 0 || model >= kNumModels || !out) {
        mono_set_error("bad model or output");
        return -1;
    }
    default_initial_state(model, out);
    return 0;
}

double mono_model_initial_v(int model) { return (model >= 0 && model < kNumModels) ? model_initial_v(model) : NAN; }

int mono_tissue_create(const mono_tissue_desc *desc, mono_tissue **out) {
    if(validate(desc)) return -1;
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    mono_tissue *t = new mono_tissue();
    t->d = *desc;
    t->d.tensor = nullptr;
    t->d.g_cell = nullptr;
    t->d.active = nullptr;
    t->d.stim_mask = nullptr;
    t->d.ionic_scale = nullptr;
    t->d.instance = nullptr;
    t->d.instance_scale = nullptr;
    const mono_tissue_desc &d = *desc;
    t->n = (int64_t)d.nx * d.ny * d.nz;
    t->neq = model_neq(d.model);
    t->pitch = (size_t)((t->n + 63) / 64 * 64);
    t->alpha = alpha_of(d);
    // [ode_solver] dt left at 0 (configure_ode_solver_from_options,
    // ode_solver.c:311-336): 0.01 ms for fixed steps, min dt 1e-10 adaptive
    t->dt_ode = d.dt_ode > 0 ? d.dt_ode : (d.ode_adaptive ? 1e-10 : 0.01);
    // sub-steps per PDE step (monodomain_solver.c:510-520); one adaptive call
    t->num_ode_steps = d.ode_adaptive ? 1 : (d.dt_pde >= t->dt_ode ? (int)(d.dt_pde / t->dt_ode) : 1);
    if(d.ode_adaptive) {
        // reference defaults of [ode_solver] abs_tol / rel_tol (config_parser.c:307-311)
        if(t->d.ode_abs_tol == 0) t->d.ode_abs_tol = 1e-5;
        if(t->d.ode_rel_tol == 0) t->d.ode_rel_tol = 1e-5;
    }
    t->stim_start = d.stim_start;
    t->cgcfg.tol = d.cg_tol;
    t->cgcfg.rel_tol = d.cg_rel_tol;
    t->cgcfg.max_its = d.cg_max_its > 0 ? d.cg_max_its : 200;
    t->cgcfg.jacobi = d.cg_jacobi;
    t->cgcfg.mixed = d.precision == MONO_FP32_REFINE;
    t->cgcfg.inner_rel = d.inner_rel > 0 ? d.inner_rel : 1e-4;
    if(t->cgcfg.mixed && t->cgcfg.jacobi) {
        mono_set_error("FP32_REFINE does not take the Jacobi preconditioner");
        delete t;
        return -1;
    }
    hipStream_t s;
    if(hipStreamCreate(&s) != hipSuccess) {
        mono_set_error("hipStreamCreate failed");
        delete t;
        return -1;
    }
    t->stream = s;
    for(int k = 0; k < 3; k++)
        if(hipEventCreate(&t->ev[k]) != hipSuccess) {
            mono_set_error("hipEventCreate failed");
            mono_tissue_destroy(t);
            return -1;
        }
    {
        const char *e = getenv("MONO_TUNE");
        if(!desc->ode_adaptive && !(e && atoi(e) == 0)) {
            if(hipEventCreate(&t->tune.e0) != hipSuccess || hipEventCreate(&t->tune.e1) != hipSuccess) {
                mono_set_error("hipEventCreate failed");
                mono_tissue_destroy(t);
                return -1;
            }
            t->tune.trial = 0;
        }
    }
    auto fail = [&](void) {
        mono_tissue_destroy(t);
        return -1;
    };

    // adaptive ODE solver rows: dt = min dt, time_new = previous_dt = 0 (as the
    // models' set_model_initial_conditions_cpu, courtemanche_ramirez_nattel_1998.c:12-24)
    if(d.ode_adaptive) {
        std::vector<double> haux(3 * t->pitch, 0.0);
        for(int64_t i = 0; i < t->n; i++) haux[i] = t->dt_ode;
        if(hipMalloc(&t->aux, sizeof(double) * 3 * t->pitch) != hipSuccess ||
           hipMemcpy(t->aux, haux.data(), sizeof(double) * 3 * t->pitch, hipMemcpyHostToDevice) != hipSuccess) {
            mono_set_error("adaptive solver state allocation failed");
            return fail();
        }
    }

    // per-cell stimulus mask and model scalings
    if(d.stim_mask) {
        if(hipMalloc(&t->stim_mask, t->n) != hipSuccess ||
           hipMemcpyAsync(t->stim_mask, d.stim_mask, t->n, hipMemcpyHostToDevice, s) != hipSuccess) {
            mono_set_error("stimulus mask upload failed");
            return fail();
        }
    }
    if(d.ionic_scale) {
        if(hipMalloc(&t->ionic_scale, sizeof(float) * 4 * t->n) != hipSuccess ||
           hipMemcpyAsync(t->ionic_scale, d.ionic_scale, sizeof(float) * 4 * t->n, hipMemcpyHostToDevice, s) != hipSuccess) {
            mono_set_error("ionic scale upload failed");
            return fail();
        }
    }
    if(d.instance) {
        const size_t tb = sizeof(float) * 4 * (size_t)d.n_instances;
        if(hipMalloc(&t->inst, sizeof(uint16_t) * t->n) != hipSuccess || hipMalloc(&t->inst_scale, tb) != hipSuccess ||
           hipMemcpyAsync(t->inst, d.instance, sizeof(uint16_t) * t->n, hipMemcpyHostToDevice, s) != hipSuccess ||
           hipMemcpyAsync(t->inst_scale, d.instance_scale, tb, hipMemcpyHostToDevice, s) != hipSuccess) {
            mono_set_error("instance table upload failed");
            return fail();
        }
        if(d.cg_rel_tol > 0.0) {
            // the relative test per instance: ||r_i|| <= rel ||b_i|| for every
            // instance i follows from r.r <= rel^2 * min_i ||b_i||^2
            const int ni = d.n_instances;
            std::vector<int> box((size_t)ni * 6);
            for(int i = 0; i < ni; i++) {
                box[(size_t)i * 6 + 0] = box[(size_t)i * 6 + 2] = box[(size_t)i * 6 + 4] = INT_MAX;
                box[(size_t)i * 6 + 1] = box[(size_t)i * 6 + 3] = box[(size_t)i * 6 + 5] = -1;
            }
            for(int64_t c = 0; c < t->n; c++) {
                if(d.active && !d.active[c]) continue;
                int *bx = &box[(size_t)d.instance[c] * 6];
                const int x = (int)(c % d.nx), y = (int)((c / d.nx) % d.ny), z = (int)(c / ((int64_t)d.nx * d.ny));
                bx[0] = std::min(bx[0], x);
                bx[1] = std::max(bx[1], x);
                bx[2] = std::min(bx[2], y);
                bx[3] = std::max(bx[3], y);
                bx[4] = std::min(bx[4], z);
                bx[5] = std::max(bx[5], z);
            }
            // work items: (instance, box row, chunk of kInstCols columns), those
            // of one instance consecutive from first[i]
            std::vector<int4> items;
            std::vector<int> first((size_t)ni + 1, 0);
            for(int i = 0; i < ni; i++) {
                first[i] = (int)items.size();
                if(box[(size_t)i * 6 + 1] < 0) continue;
                const int w = box[(size_t)i * 6 + 1] - box[(size_t)i * 6 + 0] + 1;
                const int rows = (box[(size_t)i * 6 + 3] - box[(size_t)i * 6 + 2] + 1) * (box[(size_t)i * 6 + 5] - box[(size_t)i * 6 + 4] + 1);
                for(int r = 0; r < rows; r++)
                    for(int x = 0; x < w; x += kInstCols) items.push_back(make_int4(i, r, x, 0));
            }
            first[ni] = (int)items.size();
            t->inst_nitems = (int)items.size();
            if(t->inst_nitems < 1) {
                mono_set_error("a batch with a relative CG tolerance has no active instance cells");
                return fail();
            }
            if(hipMalloc(&t->inst_box, sizeof(int) * box.size()) != hipSuccess ||
               hipMalloc(&t->inst_items, sizeof(int4) * items.size()) != hipSuccess ||
               hipMalloc(&t->inst_first, sizeof(int) * first.size()) != hipSuccess ||
               hipMalloc(&t->inst_part, sizeof(double) * items.size()) != hipSuccess ||
               hipMalloc(&t->inst_bb, sizeof(double) * ni) != hipSuccess ||
               hipMalloc(&t->tol_dev, sizeof(double)) != hipSuccess ||
               hipMemcpyAsync(t->inst_box, box.data(), sizeof(int) * box.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
               hipMemcpyAsync(t->inst_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
               hipMemcpyAsync(t->inst_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
               hipStreamSynchronize(s) != hipSuccess) {
                mono_set_error("instance box upload failed");
                return fail();
            }
            t->cgcfg.tol_dev = t->tol_dev;
        }
    }

    // operator
    double *coef27 = nullptr;
    AssemblyParams ap{};
    if(assemble_into(d, t->n, &t->tensor, &t->g, &t->active, &coef27, &ap, s)) return fail();
    if(build_operator(ap, coef27, &t->op, s)) return fail();
    if(hipFree(coef27) != hipSuccess) {
        mono_set_error("hipFree failed");
        return fail();
    }
    t->st.class_operator = t->op.class_mode ? 1 : (t->op.code_mode ? 2 : (t->op.mask32 ? 3 : 0));
    t->st.stencil_points = t->op.nslots;

    // state, potential, rhs
    const size_t wsv = d.precision == MONO_FP64 ? sizeof(double) : sizeof(float);
    if(hipMalloc(&t->sv, wsv * t->pitch * t->neq) != hipSuccess || hipMalloc(&t->v, sizeof(double) * t->n) != hipSuccess ||
       hipMalloc(&t->b, sizeof(double) * t->n) != hipSuccess) {
        mono_set_error("out of device memory for %lld cells", (long long)t->n);
        return fail();
    }
    if(hipMemsetAsync(t->v, 0, sizeof(double) * t->n, s) != hipSuccess) return fail();
    double ic[64];
    default_initial_state(d.model, ic);
    hipError_t e;
    if(d.precision == MONO_FP64) {
        double *dic = nullptr;
        if(hipMalloc(&dic, sizeof(double) * t->neq) != hipSuccess ||
           hipMemcpy(dic, ic, sizeof(double) * t->neq, hipMemcpyHostToDevice) != hipSuccess)
            return fail();
        e = launch_ionic_init(d.model, (double *)t->sv, t->pitch, nullptr, t->n, dic, s);
        if(e == hipSuccess) e = hipStreamSynchronize(s);
        if(hipFree(dic) != hipSuccess) return fail();
    } else {
        float icf[64];
        for(int k = 0; k < t->neq; k++) icf[k] = (float)ic[k];
        float *dic = nullptr;
        if(hipMalloc(&dic, sizeof(float) * t->neq) != hipSuccess ||
           hipMemcpy(dic, icf, sizeof(float) * t->neq, hipMemcpyHostToDevice) != hipSuccess)
            return fail();
        e = launch_ionic_init(d.model, (float *)t->sv, t->pitch, nullptr, t->n, dic, s);
        if(e == hipSuccess) e = hipStreamSynchronize(s);
        if(hipFree(dic) != hipSuccess) return fail();
    }
    if(e != hipSuccess) {
        mono_set_error("initial-state kernel: %s", hipGetErrorString(e));
        return fail();
    }
    // set_initial_conditions_fvm: v = INITIAL_V on active cells (0 elsewhere)
    {
        std::vector<double> hv(t->n, 0.0);
        const double v0 = model_initial_v(d.model);
        for(int64_t i = 0; i < t->n; i++) hv[i] = (!desc->active || desc->active[i]) ? v0 : 0.0;
        if(hipMemcpy(t->v, hv.data(), sizeof(double) * t->n, hipMemcpyHostToDevice) != hipSuccess) return fail();
        for(int64_t i = 0; i < t->n; i++) hv[i] *= t->alpha;
        if(hipMemcpy(t->b, hv.data(), sizeof(double) * t->n, hipMemcpyHostToDevice) != hipSuccess) return fail();
    }
    if(t->cg.init(t->n, t->pitch)) return fail();
    *out = t;
    return 0;
}

int mono_tissue_destroy(mono_tissue *t) {
    if(!t) return 0;
    int rc = 0;
    if(t->stream && hipStreamSynchronize(t->stream) != hipSuccess) rc = -1;
    if(t->cg.release()) rc = -1;
    if(t->op.release()) rc = -1;
    if(t->obs.release()) rc = -1;
    if(t->apd.release()) rc = -1;
    void *bufs[] = {t->sv, t->v, t->b, t->active, t->g, t->tensor, t->stim_mask, t->ionic_scale, t->aux, t->inst, t->inst_scale,
                    t->inst_box, t->inst_items, t->inst_first, t->inst_part, t->inst_bb, t->tol_dev};
    for(void *p : bufs)
        if(p && hipFree(p) != hipSuccess) rc = -1;
    for(int k = 0; k < 3; k++)
        if(t->ev[k] && hipEventDestroy(t->ev[k]) != hipSuccess) rc = -1;
    for(hipEvent_t e : {t->tune.e0, t->tune.e1})
        if(e && hipEventDestroy(e) != hipSuccess) rc = -1;
    if(t->stream && hipStreamDestroy(t->stream) != hipSuccess) rc = -1;
    delete t;
    return rc;
}

int mono_tissue_observe_setup(mono_tissue *t, float activation_threshold, float apd_threshold, float time_threshold,
                              int max_events, int vm_double) {
    if(!t || max_events < 1 || max_events > 1024) {
        mono_set_error("mono_tissue_observe_setup: bad arguments");
        return -1;
    }
    return t->obs.init(t->n, activation_threshold, apd_threshold, time_threshold, max_events, vm_double, t->stream);
}

int mono_tissue_observe(mono_tissue *t) {
    if(!t) {
        mono_set_error("null tissue");
        return -1;
    }
    return t->obs.observe(t->v, t->active, t->cur_time, t->stream);
}

int mono_tissue_get_activations(mono_tissue *t, int32_t *n_act, float *act, int32_t *n_apd, float *apd) {
    if(!t || !t->obs.last_v) {
        mono_set_error("activation observer not set up");
        return -1;
    }
    const size_t n = (size_t)t->n, k = (size_t)t->obs.max_events;
    if(n_act) MONO_CHECK(hipMemcpyAsync(n_act, t->obs.n_act, sizeof(int32_t) * n, hipMemcpyDeviceToHost, t->stream));
    if(act) MONO_CHECK(hipMemcpyAsync(act, t->obs.act, sizeof(float) * n * k, hipMemcpyDeviceToHost, t->stream));
    if(n_apd) MONO_CHECK(hipMemcpyAsync(n_apd, t->obs.n_apd, sizeof(int32_t) * n, hipMemcpyDeviceToHost, t->stream));
    if(apd) MONO_CHECK(hipMemcpyAsync(apd, t->obs.apd, sizeof(float) * n * k, hipMemcpyDeviceToHost, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

int mono_tissue_apd_setup(mono_tissue *t, double activation_threshold) {
    if(!t) {
        mono_set_error("null tissue");
        return -1;
    }
    return t->apd.init(t->n, activation_threshold, t->v, t->cur_time, t->stream);
}

int mono_tissue_apd_sample(mono_tissue *t) {
    if(!t) {
        mono_set_error("null tissue");
        return -1;
    }
    return t->apd.sample(t->v, t->active, t->cur_time, t->stream);
}

int mono_tissue_apd_get(mono_tissue *t, double *lat, double *vmax, double *apd90) {
    if(!t || !t->apd.prev) {
        mono_set_error("APD90 tracker not set up");
        return -1;
    }
    const size_t b = sizeof(double) * (size_t)t->n;
    if(lat) MONO_CHECK(hipMemcpyAsync(lat, t->apd.lat, b, hipMemcpyDeviceToHost, t->stream));
    if(vmax) MONO_CHECK(hipMemcpyAsync(vmax, t->apd.vmax, b, hipMemcpyDeviceToHost, t->stream));
    if(apd90) MONO_CHECK(hipMemcpyAsync(apd90, t->apd.apd90, b, hipMemcpyDeviceToHost, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

int mono_tissue_observe_events(const mono_tissue *t) { return (t && t->obs.last_v) ? t->obs.max_events : 0; }

int64_t mono_tissue_num_cells(const mono_tissue *t) { return t ? t->n : -1; }
int mono_tissue_neq(const mono_tissue *t) { return t ? t->neq : -1; }

int mono_tissue_set_timing(mono_tissue *t, int on) {
    if(!t) return -1;
    t->timing = on;
    return 0;
}

int mono_tissue_graph_counts(const mono_tissue *t, int64_t *launches, int64_t *captures) {
    if(!t) return -1;
    if(launches) *launches = t->cg.graph_launches;
    if(captures) *captures = t->cg.graph_captures;
    return 0;
}

static int launch_ionic(mono_tissue *t, double stim_amp, int block) {
    const mono_tissue_desc &d = t->d;
    if(d.precision == MONO_FP64) {
        IonicLaunch<double> a{};
        a.sv = (double *)t->sv;
        a.pitch = t->pitch;
        a.v = t->v;
        a.b = t->b;
        a.n = t->n;
        a.active = t->active;
        a.stim_mask = t->stim_mask;
        a.scale = t->ionic_scale;
        a.inst = t->inst;
        a.inst_scale = t->inst_scale;
        a.ninst = d.n_instances;
        a.nx = d.nx;
        a.dx = d.dx;
        a.stim_x_limit = d.stim_x_limit;
        a.stim_amp = stim_amp;
        a.alpha = t->alpha;
        a.dt = t->dt_ode;
        a.num_steps = t->num_ode_steps;
        a.first_step = t->st.steps == 0;
        a.block = block;
        if(d.ode_adaptive) {
            set_adaptive(t, a);
            MONO_CHECK(launch_ionic_adaptive(d.model, a, t->stream));
        } else {
            MONO_CHECK(launch_ionic_step(d.model, a, t->stream));
        }
    } else {
        IonicLaunch<float> a{};
        a.sv = (float *)t->sv;
        a.pitch = t->pitch;
        a.v = t->v;
        a.b = t->b;
        a.n = t->n;
        a.active = t->active;
        a.stim_mask = t->stim_mask;
        a.scale = t->ionic_scale;
        a.inst = t->inst;
        a.inst_scale = t->inst_scale;
        a.ninst = d.n_instances;
        a.nx = d.nx;
        a.dx = d.dx;
        a.stim_x_limit = d.stim_x_limit;
        a.stim_amp = stim_amp;
        a.alpha = t->alpha;
        a.dt = (float)t->dt_ode;
        a.num_steps = t->num_ode_steps;
        a.first_step = t->st.steps == 0;
        a.block = block;
        if(d.ode_adaptive) {
            set_adaptive(t, a);
            MONO_CHECK(launch_ionic_adaptive(d.model, a, t->stream));
        } else {
            MONO_CHECK(launch_ionic_step(d.model, a, t->stream));
        }
    }
    return 0;
}

// The block size for this step's ionic launch from the tuner: the previous
// trial's time is read back (its launch finished within the previous step's
// solve, which waited on the device), then the next candidate, or the fastest
// once every candidate has run twice after the first step.
// Partial ||b_i||^2 of one work item (one wave): up to kInstCols cells of one
// row of instance i's bounding box (a row: one (y, z) line of the box), all
// loads issued at once, summed in a fixed order into part[item]
__global__ __launch_bounds__(64) void k_instance_part(const double *b, const uint16_t *inst, const uint8_t *active, const int *box,
                                                      const int4 *items, int nx, int ny, double *part) {
    const int4 it = items[blockIdx.x];  // instance, box row, first column of the chunk
    const int i = it.x;
    const int *bx = box + 6 * i;
    const int w = bx[1] - bx[0] + 1, h = bx[3] - bx[2] + 1;
    const int y = bx[2] + it.y % h, z = bx[4] + it.y / h;
    const int64_t c0 = bx[0] + (int64_t)nx * (y + (int64_t)ny * z);
    double v[kInstCols / 64];
#pragma unroll
    for(int u = 0; u < kInstCols / 64; u++) {
        const int x = it.z + u * 64 + (int)threadIdx.x;
        const int64_t c = c0 + (x < w ? x : 0);
        const bool in = x < w && inst[c] == i && (!active || active[c]);
        v[u] = in ? b[c] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for(int u = 0; u < kInstCols / 64; u++) s += v[u] * v[u];
    for(int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// ||b_i||^2 of instance i (one wave): its items' partials, lane-strided, then
// a fixed-order wave reduction; -1 for an instance without cells
__global__ __launch_bounds__(64) void k_instance_sum(const double *part, const int *first, double *bb) {
    const int i = blockIdx.x;
    double s = 0.0;
    for(int k = first[i] + (int)threadIdx.x; k < first[i + 1]; k += 64) s += part[k];
    for(int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if(threadIdx.x == 0) bb[i] = first[i + 1] > first[i] ? s : -1.0;
}

// threshold rel^2 * min_i ||b_i||^2 over the instances that have cells
__global__ __launch_bounds__(256) void k_instance_tol(const double *bb, int ni, double rel2, double *tol) {
    double m = -1.0;
    for(int i = threadIdx.x; i < ni; i += 256) {
        const double s = bb[i];
        if(s >= 0.0 && (m < 0.0 || s < m)) m = s;
    }
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for(int o = 128; o > 0; o >>= 1) {
        if(threadIdx.x < o) {
            const double a = red[threadIdx.x], b = red[threadIdx.x + o];
            red[threadIdx.x] = a < 0.0 ? b : (b < 0.0 ? a : (b < a ? b : a));
        }
        __syncthreads();
    }
    if(threadIdx.x == 0) *tol = rel2 * (red[0] < 0.0 ? 0.0 : red[0]);
}

static int launch_instance_tolerance(mono_tissue *t) {
    const mono_tissue_desc &d = t->d;
    k_instance_part<<<t->inst_nitems, 64, 0, t->stream>>>(t->b, t->inst, t->active, t->inst_box, t->inst_items, d.nx, d.ny,
                                                           t->inst_part);
    k_instance_sum<<<d.n_instances, 64, 0, t->stream>>>(t->inst_part, t->inst_first, t->inst_bb);
    k_instance_tol<<<1, 256, 0, t->stream>>>(t->inst_bb, d.n_instances, d.cg_rel_tol * d.cg_rel_tol, t->tol_dev);
    MONO_CHECK(hipGetLastError());
    return 0;
}

static int tuned_block(mono_tissue *t, bool &timed) {
    auto &u = t->tune;
    timed = false;
    if(u.trial < 0) return u.chosen;
    if(u.pending >= 0) {
        float ms = 0.f;
        MONO_CHECK(hipEventSynchronize(u.e1));
        MONO_CHECK(hipEventElapsedTime(&ms, u.e0, u.e1));
        const int c = u.pending;
        u.best_ms[c] = (u.best_ms[c] == 0.f || ms < u.best_ms[c]) ? ms : u.best_ms[c];
        u.pending = -1;
    }
    if(t->st.steps == 0) return u.chosen;  // the first step reads V from the state rows: not timed
    if(u.trial >= 6) {
        int best = 0;
        for(int c = 1; c < 3; c++)
            if(u.best_ms[c] < u.best_ms[best]) best = c;
        u.chosen = u.cand[best];
        u.trial = -1;
        t->st.ionic_block = u.chosen;
        return u.chosen;
    }
    u.pending = u.trial % 3;
    u.trial++;
    timed = true;
    return u.cand[u.pending];
}

int mono_tissue_step(mono_tissue *t, int nsteps) {
    if(!t) {
        mono_set_error("null tissue");
        return -1;
    }
    const mono_tissue_desc &d = t->d;
    for(int k = 0; k < nsteps; k++) {
        // merged stimulus (ode_solver.c:288-341): inclusive window on the
        // accumulated time, summed over the ODE sub-steps
        double amp = 0.0;
        double tt = t->cur_time;
        for(int j = 0; j < t->num_ode_steps; j++) {
            if((tt >= t->stim_start) && (tt <= t->stim_start + d.stim_duration)) amp += d.stim_current;
            tt += t->dt_ode;
        }
        // the shifted start goes back into the configuration as "%lf" text and
        // is parsed again (ode_solver.c:224-230): it keeps six decimals
        if(d.stim_period > 0.0 && tt >= t->stim_start + d.stim_period) {
            char buf[64];
            snprintf(buf, sizeof(buf), "%lf", t->stim_start + d.stim_period);
            t->stim_start = strtod(buf, nullptr);
        }

        bool tuned = false;
        const int block = tuned_block(t, tuned);
        if(block < 0) return -1;
        if(tuned) MONO_CHECK(hipEventRecord(t->tune.e0, t->stream));
        if(t->timing) MONO_CHECK(hipEventRecord(t->ev[0], t->stream));
        if(launch_ionic(t, amp, block)) return -1;
        if(t->timing) MONO_CHECK(hipEventRecord(t->ev[1], t->stream));
        if(tuned) MONO_CHECK(hipEventRecord(t->tune.e1, t->stream));
        if(t->tol_dev && launch_instance_tolerance(t)) return -1;
        CgStats cs{};
        int rc;
        // the Jacobi update reads each row's diagonal from the stored rows, so a
        // preconditioned FP64 solve keeps them (mask-mode views store none)
        if(d.precision == MONO_FP64)
            rc = t->cg.solve_fp64(t->cgcfg.jacobi ? t->op.view64() : t->op.view64_solve(), t->b, t->v, t->cgcfg, &cs, t->stream);
        else rc = t->cg.solve_mixed(t->op.view64_solve(), t->op.view32(), t->b, t->v, t->cgcfg, &cs, t->stream);
        if(rc) return -1;
        if(t->timing) {
            MONO_CHECK(hipEventRecord(t->ev[2], t->stream));
            MONO_CHECK(hipEventSynchronize(t->ev[2]));
            float a = 0, b = 0;
            MONO_CHECK(hipEventElapsedTime(&a, t->ev[0], t->ev[1]));
            MONO_CHECK(hipEventElapsedTime(&b, t->ev[1], t->ev[2]));
            t->st.ionic_ms_total += a;
            t->st.solve_ms_total += b;
        }
        t->st.steps++;
        t->st.cg_iterations_total += cs.iterations;
        t->st.last_cg_iterations = cs.iterations;
        t->st.last_cg_error = cs.error;
        t->st.last_outer = cs.outer;
        t->st.inner_iterations_total += cs.inner_total;
        t->st.outer_sweeps_total += cs.outer;
        t->cur_time += d.dt_pde;
        t->st.time = t->cur_time;
        if(isnan(cs.error)) {
            t->st.nan_detected = 1;
            return 1;
        }
    }
    return 0;
}

int mono_tissue_sync(mono_tissue *t) {
    if(!t) return -1;
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

int mono_tissue_get_v(mono_tissue *t, double *host_v) {
    if(!t || !host_v) return -1;
    MONO_CHECK(hipMemcpyAsync(host_v, t->v, sizeof(double) * t->n, hipMemcpyDeviceToHost, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

int mono_tissue_set_v(mono_tissue *t, const double *host_v) {
    if(!t || !host_v) return -1;
    MONO_CHECK(hipMemcpyAsync(t->v, host_v, sizeof(double) * t->n, hipMemcpyHostToDevice, t->stream));
    // the first step reads V from state row 0 (the ODE solver's own initial
    // state, as the reference's first step has no Vm mirror): a potential set
    // before it goes there too, in the state's precision
    if(t->st.steps == 0) {
        if(t->d.precision == MONO_FP64) {
            MONO_CHECK(hipMemcpyAsync(t->sv, host_v, sizeof(double) * t->n, hipMemcpyHostToDevice, t->stream));
        } else {
            std::vector<float> f(t->n);
            for(int64_t i = 0; i < t->n; i++) f[i] = (float)host_v[i];
            MONO_CHECK(hipMemcpyAsync(t->sv, f.data(), sizeof(float) * t->n, hipMemcpyHostToDevice, t->stream));
            MONO_CHECK(hipStreamSynchronize(t->stream));
        }
    }
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

int mono_tissue_get_state(mono_tissue *t, double *host_sv) {
    if(!t || !host_sv) return -1;
    const size_t w = t->d.precision == MONO_FP64 ? sizeof(double) : sizeof(float);
    std::vector<char> tmp(w * t->pitch * t->neq);
    MONO_CHECK(hipMemcpyAsync(tmp.data(), t->sv, tmp.size(), hipMemcpyDeviceToHost, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    for(int s = 0; s < t->neq; s++)
        for(int64_t i = 0; i < t->n; i++) {
            const size_t k = (size_t)s * t->pitch + i;
            host_sv[(size_t)s * t->n + i] = w == 8 ? ((double *)tmp.data())[k] : (double)((float *)tmp.data())[k];
        }
    return 0;
}

int mono_tissue_set_state(mono_tissue *t, const double *host_sv) {
    if(!t || !host_sv) return -1;
    const size_t w = t->d.precision == MONO_FP64 ? sizeof(double) : sizeof(float);
    std::vector<char> tmp(w * t->pitch * t->neq, 0);
    for(int s = 0; s < t->neq; s++)
        for(int64_t i = 0; i < t->n; i++) {
            const size_t k = (size_t)s * t->pitch + i;
            if(w == 8) ((double *)tmp.data())[k] = host_sv[(size_t)s * t->n + i];
            else ((float *)tmp.data())[k] = (float)host_sv[(size_t)s * t->n + i];
        }
    MONO_CHECK(hipMemcpyAsync(t->sv, tmp.data(), tmp.size(), hipMemcpyHostToDevice, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

int mono_tissue_get_stats(const mono_tissue *t, mono_stats *out) {
    if(!t || !out) return -1;
    *out = t->st;
    return 0;
}

int mono_tissue_device_ptrs(mono_tissue *t, void **v, void **sv, size_t *pitch, void **b) {
    if(!t) return -1;
    if(v) *v = t->v;
    if(sv) *sv = t->sv;
    if(pitch) *pitch = t->pitch;
    if(b) *b = t->b;
    return 0;
}

// ---------------------------------------------------------------------------
// standalone entry points
// ---------------------------------------------------------------------------

int mono_ionic_run(int model, int precision, double *states, int64_t n, const double *stim, double dt, int nsteps) {
    if(model < 0 || model >= kNumModels || !states || n < 0 || nsteps < 0) {
        mono_set_error("mono_ionic_run: bad arguments");
        return -1;
    }
    if(precision != MONO_FP64 && precision != MONO_FP32_REFINE) {
        mono_set_error("mono_ionic_run: unknown precision %d", precision);
        return -1;
    }
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    if(n == 0) return 0;
    const int neq = model_neq(model);
    const size_t pitch = (size_t)((n + 63) / 64 * 64);
    const size_t w = precision == MONO_FP64 ? 8 : 4;
    std::vector<char> h(w * pitch * neq, 0);
    for(int s = 0; s < neq; s++)
        for(int64_t i = 0; i < n; i++) {
            const double val = states[i * neq + s];
            if(w == 8) ((double *)h.data())[s * pitch + i] = val;
            else ((float *)h.data())[s * pitch + i] = (float)val;
        }
    void *dsv = nullptr;
    double *dv = nullptr, *dstim = nullptr;
    MONO_CHECK(hipMalloc(&dsv, h.size()));
    MONO_CHECK(hipMalloc(&dv, sizeof(double) * n));
    MONO_CHECK(hipMemcpy(dsv, h.data(), h.size(), hipMemcpyHostToDevice));
    if(stim) {
        MONO_CHECK(hipMalloc(&dstim, sizeof(double) * n));
        MONO_CHECK(hipMemcpy(dstim, stim, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    // V travels in FP64 between calls through v, as in the tissue path
    std::vector<double> hv(n);
    for(int64_t i = 0; i < n; i++) hv[i] = states[i * neq];
    MONO_CHECK(hipMemcpy(dv, hv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    for(int k = 0; k < nsteps; k++) {
        hipError_t e;
        if(precision == MONO_FP64) {
            IonicLaunch<double> a{};
            a.sv = (double *)dsv;
            a.pitch = pitch;
            a.v = dv;
            a.b = dv;  // b = alpha * V* with alpha = 1 keeps V* in v for the next step
            a.alpha = 1.0;
            a.n = n;
            a.stim_cell = dstim;
            a.nx = 1;
            a.dx = 1.0;
            a.dt = dt;
            a.num_steps = 1;
            a.write_vm_row = 1;
            e = launch_ionic_step(model, a, nullptr);
        } else {
            IonicLaunch<float> a{};
            a.sv = (float *)dsv;
            a.pitch = pitch;
            a.v = dv;
            a.b = dv;
            a.alpha = 1.0;
            a.n = n;
            a.stim_cell = dstim;
            a.nx = 1;
            a.dx = 1.0;
            a.dt = (float)dt;
            a.num_steps = 1;
            a.write_vm_row = 1;
            e = launch_ionic_step(model, a, nullptr);
        }
        if(e != hipSuccess) {
            mono_set_error("ionic kernel launch: %s", hipGetErrorString(e));
            return -1;
        }
    }
    MONO_CHECK(hipDeviceSynchronize());
    MONO_CHECK(hipMemcpy(h.data(), dsv, h.size(), hipMemcpyDeviceToHost));
    MONO_CHECK(hipMemcpy(hv.data(), dv, sizeof(double) * n, hipMemcpyDeviceToHost));
    for(int s = 0; s < neq; s++)
        for(int64_t i = 0; i < n; i++)
            states[i * neq + s] = w == 8 ? ((double *)h.data())[s * pitch + i] : (double)((float *)h.data())[s * pitch + i];
    for(int64_t i = 0; i < n; i++) states[i * neq] = hv[i];
    MONO_CHECK(hipFree(dsv));
    MONO_CHECK(hipFree(dv));
    if(dstim) MONO_CHECK(hipFree(dstim));
    return 0;
}

// GPU side of the ODE-solver plugin ABI: the state lives on the device
// between calls, row 0 of it is Vm (the core mirrors cell->v into that row and
// update_monodomain reads it, monodomain_solver.c:1023-1046,
// update_monodomain.c:31-48).
int mono_ode_gpu_alloc(int model, int64_t n, double **sv, size_t *pitch_bytes) {
    if(model < 0 || model >= kNumModels || n <= 0 || !sv || !pitch_bytes) {
        mono_set_error("mono_ode_gpu_alloc: bad arguments");
        return -1;
    }
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    const int neq = model_neq(model);
    const size_t pitch = (size_t)((n + 63) / 64 * 64);
    double ic[64];
    default_initial_state(model, ic);
    double *d = nullptr, *dic = nullptr;
    MONO_CHECK(hipMalloc(&d, sizeof(double) * pitch * neq));
    hipError_t e = hipMalloc(&dic, sizeof(double) * neq);
    if(e == hipSuccess) e = hipMemcpy(dic, ic, sizeof(double) * neq, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = launch_ionic_init(model, d, pitch, nullptr, n, dic, nullptr);
    if(e == hipSuccess) e = hipDeviceSynchronize();
    if(dic) (void)hipFree(dic);
    if(e != hipSuccess) {
        (void)hipFree(d);
        mono_set_error("mono_ode_gpu_alloc: %s", hipGetErrorString(e));
        return -1;
    }
    *sv = d;
    *pitch_bytes = sizeof(double) * pitch;
    return 0;
}

int mono_ode_gpu_solve(int model, double *sv, size_t pitch_bytes, int64_t n, const double *stim, double dt, int num_steps) {
    if(model < 0 || model >= kNumModels || !sv || n < 0 || num_steps < 1 || pitch_bytes % sizeof(double) ||
       pitch_bytes / sizeof(double) < (size_t)n) {
        mono_set_error("mono_ode_gpu_solve: bad arguments");
        return -1;
    }
    if(n == 0) return 0;
    // the per-cell stimulus of the call (a host array, as the reference's
    // solve_model_odes_gpu receives it), kept in a buffer that grows as needed
    static double *dstim = nullptr;
    static int64_t cap = 0;
    if(stim) {
        if(cap < n) {
            if(dstim) MONO_CHECK(hipFree(dstim));
            dstim = nullptr;
            cap = 0;
            MONO_CHECK(hipMalloc(&dstim, sizeof(double) * n));
            cap = n;
        }
        MONO_CHECK(hipMemcpy(dstim, stim, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    IonicLaunch<double> a{};
    a.sv = sv;
    a.pitch = pitch_bytes / sizeof(double);
    a.n = n;
    a.stim_cell = stim ? dstim : nullptr;
    a.nx = 1;
    a.dx = 1.0;
    a.dt = dt;
    a.num_steps = num_steps;
    a.first_step = 1;    // V from row 0
    a.write_vm_row = 1;  // V* back to row 0
    MONO_CHECK(launch_ionic_step(model, a, nullptr));
    MONO_CHECK(hipDeviceSynchronize());
    return 0;
}

int mono_device_free(void *p) {
    if(p) MONO_CHECK(hipFree(p));
    return 0;
}

int mono_ionic_run_adaptive(int model, int precision, double *states, int64_t n, const double *stim, double current_t,
                            double *ode_dt, double *ode_time_new, double *ode_previous_dt, double min_dt, double max_dt,
                            double abs_tol, double rel_tol) {
    if(model != kCRN1998) {
        mono_set_error("mono_ionic_run_adaptive: the adaptive solver runs CRN 1998 (model %d given)", model);
        return -1;
    }
    if(!states || !ode_dt || !ode_time_new || !ode_previous_dt || n < 0 || !(min_dt > 0.0) || !(max_dt >= min_dt) ||
       !(abs_tol > 0.0) || !(rel_tol >= 0.0)) {
        mono_set_error("mono_ionic_run_adaptive: bad arguments");
        return -1;
    }
    if(precision != MONO_FP64 && precision != MONO_FP32_REFINE) {
        mono_set_error("mono_ionic_run_adaptive: unknown precision %d", precision);
        return -1;
    }
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    if(n == 0) return 0;
    const int neq = model_neq(model);
    const size_t pitch = (size_t)((n + 63) / 64 * 64);
    const size_t w = precision == MONO_FP64 ? 8 : 4;
    std::vector<char> h(w * pitch * neq, 0);
    for(int s = 0; s < neq; s++)
        for(int64_t i = 0; i < n; i++) {
            const double val = states[i * neq + s];
            if(w == 8) ((double *)h.data())[s * pitch + i] = val;
            else ((float *)h.data())[s * pitch + i] = (float)val;
        }
    std::vector<double> haux(3 * pitch, 0.0), hv(n);
    for(int64_t i = 0; i < n; i++) {
        haux[i] = ode_dt[i];
        haux[pitch + i] = ode_time_new[i];
        haux[2 * pitch + i] = ode_previous_dt[i];
        hv[i] = states[i * neq];
    }
    void *dsv = nullptr;
    double *dv = nullptr, *dstim = nullptr, *daux = nullptr;
    auto cleanup = [&]() {
        void *bufs[] = {dsv, dv, dstim, daux};
        for(void *p : bufs)
            if(p) (void)hipFree(p);
    };
    hipError_t e = hipMalloc(&dsv, h.size());
    if(e == hipSuccess) e = hipMalloc(&dv, sizeof(double) * n);
    if(e == hipSuccess) e = hipMalloc(&daux, sizeof(double) * 3 * pitch);
    if(e == hipSuccess && stim) e = hipMalloc(&dstim, sizeof(double) * n);
    if(e == hipSuccess) e = hipMemcpy(dsv, h.data(), h.size(), hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(dv, hv.data(), sizeof(double) * n, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(daux, haux.data(), sizeof(double) * 3 * pitch, hipMemcpyHostToDevice);
    if(e == hipSuccess && stim) e = hipMemcpy(dstim, stim, sizeof(double) * n, hipMemcpyHostToDevice);
    if(e == hipSuccess) {
        // final time of the call: current_t + min_dt (default_solvers.c:9-46)
        const double final_time = current_t + min_dt;
        if(precision == MONO_FP64) {
            IonicLaunch<double> a{};
            a.sv = (double *)dsv;
            a.pitch = pitch;
            a.v = dv;
            a.b = dv;
            a.alpha = 1.0;
            a.n = n;
            a.stim_cell = dstim;
            a.nx = 1;
            a.dx = 1.0;
            a.write_vm_row = 1;
            a.aux = daux;
            a.final_time = final_time;
            a.min_dt = min_dt;
            a.max_dt = max_dt;
            a.abs_tol = abs_tol;
            a.rel_tol = rel_tol;
            e = launch_ionic_adaptive(model, a, nullptr);
        } else {
            IonicLaunch<float> a{};
            a.sv = (float *)dsv;
            a.pitch = pitch;
            a.v = dv;
            a.b = dv;
            a.alpha = 1.0;
            a.n = n;
            a.stim_cell = dstim;
            a.nx = 1;
            a.dx = 1.0;
            a.write_vm_row = 1;
            a.aux = daux;
            a.final_time = final_time;
            a.min_dt = min_dt;
            a.max_dt = max_dt;
            a.abs_tol = abs_tol;
            a.rel_tol = rel_tol;
            e = launch_ionic_adaptive(model, a, nullptr);
        }
    }
    if(e == hipSuccess) e = hipDeviceSynchronize();
    if(e == hipSuccess) e = hipMemcpy(h.data(), dsv, h.size(), hipMemcpyDeviceToHost);
    if(e == hipSuccess) e = hipMemcpy(hv.data(), dv, sizeof(double) * n, hipMemcpyDeviceToHost);
    if(e == hipSuccess) e = hipMemcpy(haux.data(), daux, sizeof(double) * 3 * pitch, hipMemcpyDeviceToHost);
    cleanup();
    if(e != hipSuccess) {
        mono_set_error("mono_ionic_run_adaptive: %s", hipGetErrorString(e));
        return -1;
    }
    for(int s = 1; s < neq; s++)
        for(int64_t i = 0; i < n; i++)
            states[i * neq + s] = w == 8 ? ((double *)h.data())[s * pitch + i] : (double)((float *)h.data())[s * pitch + i];
    for(int64_t i = 0; i < n; i++) {
        states[i * neq] = hv[i];
        ode_dt[i] = haux[i];
        ode_time_new[i] = haux[pitch + i];
        ode_previous_dt[i] = haux[2 * pitch + i];
    }
    return 0;
}

int mono_assemble(const mono_tissue_desc *desc, double *coef27, double *alpha_out) {
    if(validate(desc)) return -1;
    if(!coef27) {
        mono_set_error("mono_assemble: null output");
        return -1;
    }
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    const int64_t n = (int64_t)desc->nx * desc->ny * desc->nz;
    double *tensor = nullptr, *g = nullptr, *dcoef = nullptr;
    uint8_t *act = nullptr;
    AssemblyParams ap{};
    if(assemble_into(*desc, n, &tensor, &g, &act, &dcoef, &ap, nullptr)) return -1;
    MONO_CHECK(hipDeviceSynchronize());
    MONO_CHECK(hipMemcpy(coef27, dcoef, sizeof(double) * 27 * n, hipMemcpyDeviceToHost));
    if(alpha_out) *alpha_out = ap.alpha;
    void *bufs[] = {tensor, g, act, dcoef};
    for(void *p : bufs)
        if(p) MONO_CHECK(hipFree(p));
    return 0;
}

int mono_cg_stencil(int nx, int ny, int nz, const double *coef27, const uint8_t *active, const double *b, double *x, double tol,
                    double rel_tol, int max_its, int jacobi, int precision, int *iterations, double *error) {
    if(nx <= 0 || ny <= 0 || nz <= 0 || !coef27 || !b || !x) {
        mono_set_error("mono_cg_stencil: bad arguments");
        return -1;
    }
    if((int64_t)nx * ny * nz > (int64_t)1 << 31) {
        mono_set_error("mono_cg_stencil: too many cells: %lld", (long long)nx * ny * nz);
        return -1;
    }
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    const int64_t n = (int64_t)nx * ny * nz;
    double *dcoef = nullptr, *db = nullptr, *dx = nullptr;
    uint8_t *dact = nullptr;
    MONO_CHECK(hipMalloc(&dcoef, sizeof(double) * 27 * n));
    MONO_CHECK(hipMalloc(&db, sizeof(double) * n));
    MONO_CHECK(hipMalloc(&dx, sizeof(double) * n));
    MONO_CHECK(hipMemcpy(dcoef, coef27, sizeof(double) * 27 * n, hipMemcpyHostToDevice));
    MONO_CHECK(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
    MONO_CHECK(hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice));
    if(active) {
        MONO_CHECK(hipMalloc(&dact, n));
        MONO_CHECK(hipMemcpy(dact, active, n, hipMemcpyHostToDevice));
    }
    AssemblyParams ap{};
    ap.nx = nx;
    ap.ny = ny;
    ap.nz = nz;
    ap.active = dact;
    Operator op;
    if(build_operator(ap, dcoef, &op, nullptr)) return -1;
    CgWorkspace ws;
    if(ws.init(n, n)) return -1;
    CgConfig cfg;
    cfg.tol = tol;
    cfg.rel_tol = rel_tol;
    cfg.max_its = max_its;
    cfg.jacobi = jacobi;
    CgStats st{};
    int rc = precision == MONO_FP64 ? ws.solve_fp64(op.view64(), db, dx, cfg, &st, nullptr)
                                    : ws.solve_mixed(op.view64(), op.view32(), db, dx, cfg, &st, nullptr);
    if(rc) return -1;
    MONO_CHECK(hipDeviceSynchronize());
    MONO_CHECK(hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost));
    if(iterations) *iterations = st.iterations;
    if(error) *error = st.error;
    if(ws.release() || op.release()) return -1;
    void *bufs[] = {dcoef, db, dx, dact};
    for(void *p : bufs)
        if(p) MONO_CHECK(hipFree(p));
    return 0;
}

}  // extern "C"
