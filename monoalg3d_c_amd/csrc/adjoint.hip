// Sensor loss and its gradient with respect to per-cell conductivity factors
// and per-cell Mitchell-Schaeffer tau_close factors, by the discrete adjoint of
// the operator-split step (mono_sensor_gradient in monoalg_hip.h).  No
// counterpart in the reference.
//
// Forward step n -> n+1 of the tissue loop (monodomain_solver.c:616-925 with
// the Mitchell-Schaeffer model, mitchell_shaeffer_2003.c:56-93, one ODE step
// per PDE step):
//   V*      = V_n + dt (-V_n/tau_out + h_n V_n^2 (1 - V_n)/tau_in + stim)
//   h_{n+1} = h_n + dt ((1 - h_n)/tau_open)            if V_n <  V_gate
//             h_n + dt (-h_n/(tau_close s))             otherwise
//   A(g) V_{n+1} = alpha V*
// Loss: L = sum over records m (every R steps, from step 0) and sensor cells c
// of (V_{mR}[c] - target[m][c])^2.
//
// Backward, with lambda = dL/dV_{n+1} and mu = dL/dh_{n+1} (A is symmetric):
//   w = A^-1 lambda                                   (adjoint solve, FP64 CG)
//   dL/dg_k -= sum over faces f of k: dc_f/dg_k (w_k - w_j)(V_k - V_j)   at V_{n+1}
//   lambda_n = alpha w dV*/dV_n
//   mu_n     = alpha w dV*/dh_n + mu dh_{n+1}/dh_n
//   dL/ds_k += mu dh_{n+1}/ds_k
//   lambda_n += 2 (V_n - target) on the sensors when n is a record step.
// c_f = 2 sa sb / (sa + sb) * area/distance with sa = sigma g_a (the isotropic
// face conductivity, assembly_common.c:911-1048).  The MS gate switch has zero
// derivative almost everywhere and is taken as such.
//
// Checkpoints: the forward run keeps (v, states) every K steps; the backward
// sweep recomputes each K-step segment from its checkpoint, keeping the
// segment's states, then walks it backwards.  With K <= 0 the forward keeps
// every step's V and h instead (16 bytes per cell and step) and nothing is
// recomputed.  The forward may run in FP32 + refinement; the adjoint always
// runs in FP64 on the stored states.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "../../include/monoalg_hip.h"
#include "cg.h"
#include "mono_common.h"
#include "tissue_impl.h"

namespace mono {

__global__ __launch_bounds__(256) void k_gather(const double *v, const int64_t *idx, int m, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < m) out[i] = v[idx[i]];
}

// lambda[idx[i]] += val[i] (sensor cells are distinct)
__global__ __launch_bounds__(256) void k_scatter_add(double *lam, const int64_t *idx, int m, const double *val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < m) lam[idx[i]] += val[i];
}

// V and h (state row 1, FP64) of the tissue in one launch
template <typename T>
__global__ __launch_bounds__(256) void k_copy_vh(const double *v, const T *h, double *vdst, double *hdst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) {
        vdst[i] = v[i];
        hdst[i] = (double)h[i];
    }
}

// One backward step per cell, given w = A^-1 lambda_{n+1} (and the previous
// solution wprev, for the next solve's extrapolated start wnext):
//   dL/dg_c -= sum over the six faces of c: dc/dg_c (w_c - w_j)(V_{n+1,c} - V_{n+1,j})
//   ionic adjoint of the forward-Euler Mitchell-Schaeffer step at (V_n, h_n):
//   lambda_n = alpha w dV*/dV, mu_n = alpha w dV*/dh + mu dh'/dh, dL/ds += mu dh'/ds
// two device-to-device copies in one launch (8-byte words): checkpoint save
// and restore of v and the state rows.  A kernel rather than hipMemcpyAsync:
// at 65k cells the runtime's device copies cost far more than the bytes moved.
__global__ __launch_bounds__(256) void k_copy2(const uint64_t *s0, uint64_t *d0, int64_t n0, const uint64_t *s1, uint64_t *d1,
                                               int64_t n1) {
    for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n0 + n1; i += (int64_t)gridDim.x * 256) {
        if(i < n0) d0[i] = s0[i];
        else d1[i - n0] = s1[i - n0];
    }
}

int copy2(const void *s0, void *d0, size_t b0, const void *s1, void *d1, size_t b1, hipStream_t s) {
    const int64_t n0 = (int64_t)(b0 / 8), n1 = (int64_t)(b1 / 8);
    int64_t g = (n0 + n1 + 255) / 256;
    if(g > 4096) g = 4096;
    if(g < 1) g = 1;
    k_copy2<<<(unsigned)g, 256, 0, s>>>((const uint64_t *)s0, (uint64_t *)d0, n0, (const uint64_t *)s1, (uint64_t *)d1, n1);
    MONO_CHECK(hipGetLastError());
    return 0;
}

__global__ __launch_bounds__(256) void k_adjoint_step(int nx, int ny, int nz, const uint8_t *active, const double *g, double sx,
                                                      double sy, double sz, double mx, double my, double mz, const double *w,
                                                      const double *wprev, double *wnext, const double *V1, const double *V,
                                                      const double *h, const float *scale,
                                                      double alpha, double dt, double *lam, double *mu, double *grad_g,
                                                      double *grad_s) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    if(active && !active[c]) {
        lam[c] = 0.0;
        mu[c] = 0.0;
        wnext[c] = 0.0;
        return;
    }
    const double wk = w[c];
    wnext[c] = 2.0 * wk - wprev[c];  // warm start of the next adjoint solve
    if(grad_g) {
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        const double gk = g ? g[c] : 1.0, vk = V1[c];
        double acc = 0.0;
        const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
        for(int f = 0; f < 6; f++) {
            const int xx = x + off[f][0], yy = y + off[f][1], zz = z + off[f][2];
            if(xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
            const int64_t j = xx + (int64_t)nx * (yy + (int64_t)ny * zz);
            if(active && !active[j]) continue;
            const double sig = f < 2 ? sx : (f < 4 ? sy : sz);
            const double mult = f < 2 ? mx : (f < 4 ? my : mz);
            const double sa = sig * gk, sb = sig * (g ? g[j] : 1.0);
            if(sa == 0.0 || sb == 0.0) continue;
            const double dcdg = mult * sig * 2.0 * sb * sb / ((sa + sb) * (sa + sb));
            acc += dcdg * (wk - w[j]) * (vk - V1[j]);
        }
        grad_g[c] -= acc;
    }
    const double tau_in = 0.3, tau_out = 6.0, V_gate = 0.13, tau_open = 120.0;
    const double s = scale ? (double)scale[4 * c] : 1.0;
    const double tau_close = 150.0 * s;
    const double v = V[c], hh = h[c];
    const double a = alpha * wk;
    const double dVdV = 1.0 + dt * (-1.0 / tau_out + hh * (2.0 * v - 3.0 * v * v) / tau_in);
    const double dVdh = dt * (v * v * (1.0 - v)) / tau_in;
    double dhdh, dhds;
    if(v < V_gate) {
        dhdh = 1.0 - dt / tau_open;
        dhds = 0.0;
    } else {
        dhdh = 1.0 - dt / tau_close;
        dhds = dt * hh * 150.0 / (tau_close * tau_close);
    }
    const double m1 = mu[c];
    if(grad_s) grad_s[c] += m1 * dhds;
    lam[c] = a * dVdV;
    mu[c] = a * dVdh + m1 * dhdh;
}

}  // namespace mono

using namespace mono;

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if(p) (void)hipFree(p);
    }
    template <typename T> T *as() const { return (T *)p; }
};

int check_spec(const mono_tissue_desc *d, int nsteps, const int64_t *sensors, int nsens, int record_every, int checkpoint_every) {
    if(!d || !sensors || nsens <= 0 || nsteps < 0 || record_every < 1) {
        mono_set_error("mono_sensor_gradient: bad arguments");
        return -1;
    }
    if(d->model != MONO_MODEL_MS2003 || d->aniso || d->tensor_source != MONO_TENSOR_UNIFORM || d->ode_adaptive) {
        mono_set_error("mono_sensor_gradient: the adjoint covers Mitchell-Schaeffer on an isotropic uniform tensor "
                       "(with per-cell factors g) and fixed-step ODEs");
        return -1;
    }
    if(d->dt_ode > 0 && d->dt_ode != d->dt_pde) {
        mono_set_error("mono_sensor_gradient: one ODE step per PDE step (dt_ode = dt_pde)");
        return -1;
    }
    const int64_t n = (int64_t)d->nx * d->ny * d->nz;
    for(int i = 0; i < nsens; i++) {
        if(sensors[i] < 0 || sensors[i] >= n || (d->active && !d->active[sensors[i]])) {
            mono_set_error("mono_sensor_gradient: sensor %d (cell %lld) outside the active tissue", i, (long long)sensors[i]);
            return -1;
        }
        for(int j = 0; j < i; j++)
            if(sensors[j] == sensors[i]) {
                mono_set_error("mono_sensor_gradient: sensor cell %lld listed twice", (long long)sensors[i]);
                return -1;
            }
    }
    return 0;
}

// v and the state rows at a checkpoint: views into one allocation for all
// checkpoints of a run (one hipMalloc instead of two per checkpoint; each
// allocation costs about a millisecond, as much as several steps at 65k cells)
struct Checkpoint {
    double *v = nullptr;
    void *sv = nullptr;
    double cur_time = 0.0, stim_start = 0.0;
    int64_t steps = 0;
    // the mixed solve's sweep schedule and polling sizes: a recomputed segment
    // then repeats the forward run's solves exactly
    std::vector<int> sched;
    int last_iters = 0, last_inner = 0;
};

int save(mono_tissue *t, Checkpoint &c, size_t sv_bytes) {
    if(copy2(t->v, c.v, sizeof(double) * t->n, t->sv, c.sv, sv_bytes, t->stream)) return -1;
    c.cur_time = t->cur_time;
    c.stim_start = t->stim_start;
    c.steps = t->st.steps;
    c.sched = t->cg.last_sched;
    c.last_iters = t->cg.last_iters;
    c.last_inner = t->cg.last_inner;
    return 0;
}

int restore(mono_tissue *t, const Checkpoint &c, size_t sv_bytes) {
    if(copy2(c.v, t->v, sizeof(double) * t->n, c.sv, t->sv, sv_bytes, t->stream)) return -1;
    t->cur_time = c.cur_time;
    t->st.time = c.cur_time;
    t->stim_start = c.stim_start;
    t->st.steps = c.steps;
    t->cg.last_sched = c.sched;
    t->cg.last_iters = c.last_iters;
    t->cg.last_inner = c.last_inner;
    return 0;
}

// V and h of the tissue into vdst, hdst (FP64)
int copy_vh(mono_tissue *t, double *vdst, double *hdst) {
    const unsigned grid = (unsigned)((t->n + 255) / 256);
    if(t->d.precision == MONO_FP64)
        k_copy_vh<double><<<grid, 256, 0, t->stream>>>(t->v, (const double *)t->sv + t->pitch, vdst, hdst, t->n);
    else
        k_copy_vh<float><<<grid, 256, 0, t->stream>>>(t->v, (const float *)t->sv + t->pitch, vdst, hdst, t->n);
    MONO_CHECK(hipGetLastError());
    return 0;
}

int record(mono_tissue *t, const DevBuf &idx, int nsens, DevBuf &tmp, double *host_row) {
    k_gather<<<(nsens + 255) / 256, 256, 0, t->stream>>>(t->v, idx.as<int64_t>(), nsens, tmp.as<double>());
    MONO_CHECK(hipGetLastError());
    MONO_CHECK(hipMemcpyAsync(host_row, tmp.p, sizeof(double) * nsens, hipMemcpyDeviceToHost, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

}  // namespace

extern "C" int mono_sensor_gradient(const mono_tissue_desc *desc, int nsteps, const int64_t *sensors, int nsens, int record_every,
                                    const double *target, int checkpoint_every, double *traces, double *loss, double *grad_g,
                                    double *grad_scale) {
    if(check_spec(desc, nsteps, sensors, nsens, record_every, checkpoint_every)) return -1;
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    mono_tissue *t = nullptr;
    // one ODE step per PDE step: a dt_ode left at 0 means dt_pde here
    mono_tissue_desc d1 = *desc;
    if(!(d1.dt_ode > 0)) d1.dt_ode = d1.dt_pde;
    if(int rc = mono_tissue_create(&d1, &t)) return rc;
    struct Guard {
        mono_tissue *t;
        ~Guard() { mono_tissue_destroy(t); }
    } guard{t};
    const int64_t n = t->n;
    const int nrec = nsteps / record_every + 1;
    const size_t sv_bytes = (size_t)t->pitch * t->neq * (desc->precision == MONO_FP64 ? 8 : 4);
    std::vector<double> tr((size_t)nrec * nsens);
    DevBuf idx, tmp;
    MONO_CHECK(hipMalloc(&idx.p, sizeof(int64_t) * nsens));
    MONO_CHECK(hipMalloc(&tmp.p, sizeof(double) * nsens));
    MONO_CHECK(hipMemcpy(idx.p, sensors, sizeof(int64_t) * nsens, hipMemcpyHostToDevice));

    // forward run with checkpoints every K steps, or (checkpoint_every <= 0)
    // keeping every step's states so that nothing is recomputed
    const bool store_all = checkpoint_every <= 0;
    const bool want_grad = target && (grad_g || grad_scale);
    const int K = store_all ? (nsteps > 0 ? nsteps : 1) : checkpoint_every;
    const int nseg = nsteps > 0 ? (nsteps + K - 1) / K : 0;
    std::vector<Checkpoint> ck(store_all ? 0 : nseg);
    const size_t nb = sizeof(double) * n;
    DevBuf ckpool;
    if(!ck.empty()) {
        const size_t vb = (nb + 255) / 256 * 256, per = vb + (sv_bytes + 255) / 256 * 256;
        if(hipMalloc(&ckpool.p, per * ck.size()) != hipSuccess) {
            mono_set_error("mono_sensor_gradient: no device memory for %zu checkpoints (%.2f GB)", ck.size(), per * ck.size() / 1e9);
            return -1;
        }
        for(size_t i = 0; i < ck.size(); i++) {
            ck[i].v = (double *)(ckpool.as<char>() + i * per);
            ck[i].sv = ckpool.as<char>() + i * per + vb;
        }
    }
    DevBuf Vseg, Hseg;  // states of one segment (or of the whole run)
    if(want_grad) {
        if(hipMalloc(&Vseg.p, nb * (K + 1)) != hipSuccess || hipMalloc(&Hseg.p, nb * (K + 1)) != hipSuccess) {
            mono_set_error("mono_sensor_gradient: no device memory for %d steps of states (%.2f GB); use a checkpoint interval",
                           K + 1, 2.0 * nb * (K + 1) / 1e9);
            return -1;
        }
    }
    if(want_grad && store_all && copy_vh(t, Vseg.as<double>(), Hseg.as<double>())) return -1;
    const bool trace = getenv("MONO_TRACE_ADJOINT") != nullptr;
    auto now = [&]() {
        if(trace) (void)hipStreamSynchronize(t->stream);
        return std::chrono::steady_clock::now();
    };
    const auto f0 = now();
    for(int k = 0; k < nsteps; k++) {
        if(!store_all && k % K == 0 && save(t, ck[k / K], sv_bytes)) return -1;
        if(k % record_every == 0 && record(t, idx, nsens, tmp, &tr[(size_t)(k / record_every) * nsens])) return -1;
        const int rc = mono_tissue_step(t, 1);
        if(rc < 0) return -1;
        if(rc == 1) {
            mono_set_error("mono_sensor_gradient: the forward run stopped on a NaN at step %d", k);
            return -1;
        }
        if(want_grad && store_all && copy_vh(t, Vseg.as<double>() + (size_t)(k + 1) * n, Hseg.as<double>() + (size_t)(k + 1) * n))
            return -1;
    }
    if(nsteps % record_every == 0 && record(t, idx, nsens, tmp, &tr[(size_t)(nsteps / record_every) * nsens])) return -1;
    const double t_fw = std::chrono::duration<double>(now() - f0).count();
    if(traces) memcpy(traces, tr.data(), sizeof(double) * tr.size());
    if(!target) return 0;

    // loss and its direct gradient on the sensors, 2 (V - target)
    std::vector<double> res(tr.size());
    double L = 0.0;
    for(size_t i = 0; i < tr.size(); i++) {
        const double e = tr[i] - target[i];
        L += e * e;
        res[i] = 2.0 * e;
    }
    if(loss) *loss = L;
    if(!grad_g && !grad_scale) return 0;

    // backward sweep
    DevBuf lam, mu, gg, gs, dres, wb[3];
    MONO_CHECK(hipMalloc(&lam.p, nb));
    MONO_CHECK(hipMalloc(&mu.p, nb));
    for(DevBuf &b : wb) MONO_CHECK(hipMalloc(&b.p, nb));
    MONO_CHECK(hipMalloc(&gg.p, nb));
    MONO_CHECK(hipMalloc(&gs.p, nb));
    MONO_CHECK(hipMalloc(&dres.p, sizeof(double) * res.size()));
    for(DevBuf *b : {&lam, &mu, &gg, &gs, &wb[0], &wb[1], &wb[2]}) MONO_CHECK(hipMemsetAsync(b->p, 0, nb, t->stream));
    // adjoint solutions rotate through three buffers: the solve of step k
    // starts from the linear extrapolation 2 w_{k+1} - w_{k+2} of the two
    // previous solutions, written by the previous step's k_adjoint_step
    int wi = 0;
    MONO_CHECK(hipMemcpyAsync(dres.p, res.data(), sizeof(double) * res.size(), hipMemcpyHostToDevice, t->stream));

    const mono_tissue_desc &d = t->d;
    const double sx = desc->tensor[0], sy = desc->tensor[1], sz = desc->tensor[2];  // uniform tensor (host)
    const double mx = (d.dy * d.dz) / d.dx, my = (d.dx * d.dz) / d.dy, mz = (d.dx * d.dy) / d.dz;
    CgConfig acfg = t->cgcfg;
    acfg.mixed = false;
    acfg.jacobi = 0;
    acfg.tol = 0.0;
    acfg.tol_dev = nullptr;
    // adjoint solves to ||r|| <= 1e-8 ||lambda|| (or the forward's relative
    // tolerance when that is looser): the gradient needs far less than the
    // forward's accuracy, and each digit costs about two CG iterations
    acfg.rel_tol = d.cg_rel_tol > 1e-8 ? d.cg_rel_tol : 1e-8;
    const unsigned grid = (unsigned)((n + 255) / 256);
    const OpView<double> op = t->op.view64();
    bool live = false;  // lambda, mu non-zero yet
    // MONO_TRACE_ADJOINT=1: phase times and adjoint CG iterations on stderr
    double t_re = 0.0, t_adj = 0.0;
    long adj_its = 0, adj_solves = 0;
    for(int seg = nseg - 1; seg >= 0; seg--) {
        const int k0 = seg * K, k1 = (k0 + K < nsteps) ? k0 + K : nsteps;
        auto c0 = now();
        double *VS = Vseg.as<double>(), *HS = Hseg.as<double>();
        if(!store_all) {
            // recompute the segment from its checkpoint, keeping its states
            if(restore(t, ck[seg], sv_bytes)) return -1;
            if(copy_vh(t, VS, HS)) return -1;
            for(int k = k0; k < k1; k++) {
                if(mono_tissue_step(t, 1) < 0) return -1;
                if(copy_vh(t, VS + (size_t)(k - k0 + 1) * n, HS + (size_t)(k - k0 + 1) * n)) return -1;
            }
        }
        auto c1 = now();
        t_re += std::chrono::duration<double>(c1 - c0).count();
        for(int k = k1 - 1; k >= k0; k--) {
            // lambda += dL/dV_{k+1} on the sensors
            if((k + 1) % record_every == 0) {
                k_scatter_add<<<(nsens + 255) / 256, 256, 0, t->stream>>>(lam.as<double>(), idx.as<int64_t>(), nsens,
                                                                         dres.as<double>() + (size_t)((k + 1) / record_every) * nsens);
                MONO_CHECK(hipGetLastError());
                live = true;
            }
            if(!live) continue;
            const double *Vn = VS + (size_t)(k - k0) * n, *Vn1 = VS + (size_t)(k - k0 + 1) * n;
            const double *Hn = HS + (size_t)(k - k0) * n;
            // w = A^-1 lambda, warm-started from the previous w
            CgStats cs{};
            double *w = wb[wi].as<double>(), *wprev = wb[(wi + 2) % 3].as<double>(), *wnext = wb[(wi + 1) % 3].as<double>();
            if(t->cg.solve_fp64(op, lam.as<double>(), w, acfg, &cs, t->stream)) return -1;
            adj_its += cs.iterations;
            adj_solves++;
            k_adjoint_step<<<grid, 256, 0, t->stream>>>(d.nx, d.ny, d.nz, t->active, t->g, sx, sy, sz, mx, my, mz, w, wprev, wnext,
                                                        Vn1, Vn, Hn, t->ionic_scale, t->alpha, d.dt_pde, lam.as<double>(),
                                                        mu.as<double>(), grad_g ? gg.as<double>() : nullptr,
                                                        grad_scale ? gs.as<double>() : nullptr);
            MONO_CHECK(hipGetLastError());
            wi = (wi + 1) % 3;
        }
        t_adj += std::chrono::duration<double>(now() - c1).count();
    }
    MONO_CHECK(hipStreamSynchronize(t->stream));
    if(trace)
        fprintf(stderr, "[mono] adjoint: forward %.3f s, recompute %.3f s, adjoint sweep %.3f s, %ld solves, %.2f CG iterations per solve\n",
                t_fw, t_re, t_adj, adj_solves, adj_solves ? (double)adj_its / adj_solves : 0.0);
    if(grad_g) MONO_CHECK(hipMemcpy(grad_g, gg.p, nb, hipMemcpyDeviceToHost));
    if(grad_scale) MONO_CHECK(hipMemcpy(grad_scale, gs.p, nb, hipMemcpyDeviceToHost));
    return 0;
}
