// Checkpoint and restart of a tissue (mono_tissue_save_state /
// mono_tissue_restore_state, monoalg_hip.h).
//
// Three files carry the reference's checkpoint layout, so that a checkpoint
// written by the reference can be restarted here (save_state.c:16-155,
// restore_state.c:45-256; every field written one by one, no padding, little
// endian, real = double as the reference builds its models):
//
//   grid_checkpoint.dat               mesh header {cube side xyz, mesh side xyz
//                                     (double), number_of_cells, num_active_cells
//                                     (uint32)} then per cell {centre xyz, v,
//                                     six face fluxes, b (double), can_change,
//                                     active (bool), extra info size (uint64),
//                                     extra info bytes}
//   monodomain_solver_checkpoint.dat  struct monodomain_solver (88 bytes,
//                                     monodomain_solver.h:16-36) then struct
//                                     time_info {current_t, final_t, dt (double),
//                                     iteration (int), 4 bytes padding}
//   ode_solver_checkpoint.dat         adaptive (bool), max_dt, min_dt, rel_tol,
//                                     abs_tol (real), num_cells_to_solve (uint64),
//                                     cells-to-solve flag (bool), gpu (bool),
//                                     gpu_id (int), pitch (uint64),
//                                     original_num_cells (uint32), then the states
//                                     of the active cells: gpu = 1 state-major
//                                     [neq (+3 adaptive rows dt, time_new,
//                                     previous_dt)][N], gpu = 0 cell-major [N][neq]
//
// The k-th active cell record of the grid file owns column k of the states
// (sv_position = grid_position, monodomain_solver.c:1116, order_grid_cells,
// grid.c:245-263); cells are matched to this tissue's cells by their centres,
// so the reference's space-filling cell order restores correctly.  Files
// written here list the cells in lexicographic order with gpu = 1; the
// reference restores their potentials (matched by centre) but would read
// their states in its own cell order, so the way back is for this library.
//
// A fourth file, hip_solver_checkpoint.dat, holds what the reference's files
// cannot: the exact stimulus start after periodic shifts, the mixed-precision
// solver's sweep schedule and polling sizes, the run statistics and the activation-time observer.
// With it a restart reproduces the uninterrupted run bit for bit in both
// precisions.  Without it (a reference checkpoint) the run continues from the
// configuration's stimulus start as the reference's restart does (the shifted
// start lives only in its in-memory configuration, ode_solver.c:224-230).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/monoalg_hip.h"
#include "mono_common.h"
#include "tissue_impl.h"

using namespace mono;

namespace {

constexpr char kMagic[8] = {'M', 'O', 'N', 'O', 'H', 'I', 'P', 'C'};
constexpr uint32_t kVersion = 2;  // 2: mono_stats gained ionic_block
constexpr size_t kSolverBytes = 88, kTimeInfoBytes = 32;  // struct monodomain_solver, struct time_info

struct File {
    FILE *f = nullptr;
    std::string path;
    ~File() {
        if(f) fclose(f);
    }
    bool open(const char *dir, const char *name, const char *mode) {
        path = std::string(dir) + "/" + name;
        f = fopen(path.c_str(), mode);
        return f != nullptr;
    }
    bool put(const void *p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes; }
    bool get(void *p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }
    template <class T> bool put(T x) { return put(&x, sizeof(T)); }
    template <class T> bool get(T &x) { return get(&x, sizeof(T)); }
    long size() {
        const long at = ftell(f);
        fseek(f, 0, SEEK_END);
        const long s = ftell(f);
        fseek(f, at, SEEK_SET);
        return s;
    }
    bool close() {
        const bool ok = fclose(f) == 0;
        f = nullptr;
        return ok;
    }
};

size_t state_bytes(const mono_tissue *t) {
    return (t->d.precision == MONO_FP64 ? sizeof(double) : sizeof(float)) * t->pitch * t->neq;
}

// states (neq rows of `pitch`) as FP64, whatever the tissue's precision
int download_states(mono_tissue *t, std::vector<double> &sv) {
    sv.assign((size_t)t->pitch * t->neq, 0.0);
    if(t->d.precision == MONO_FP64) {
        MONO_CHECK(hipMemcpyAsync(sv.data(), t->sv, state_bytes(t), hipMemcpyDeviceToHost, t->stream));
        MONO_CHECK(hipStreamSynchronize(t->stream));
    } else {
        std::vector<float> f((size_t)t->pitch * t->neq);
        MONO_CHECK(hipMemcpyAsync(f.data(), t->sv, state_bytes(t), hipMemcpyDeviceToHost, t->stream));
        MONO_CHECK(hipStreamSynchronize(t->stream));
        for(size_t i = 0; i < f.size(); i++) sv[i] = f[i];
    }
    return 0;
}

int upload_states(mono_tissue *t, const std::vector<double> &sv) {
    if(t->d.precision == MONO_FP64) {
        MONO_CHECK(hipMemcpyAsync(t->sv, sv.data(), state_bytes(t), hipMemcpyHostToDevice, t->stream));
        MONO_CHECK(hipStreamSynchronize(t->stream));
    } else {
        std::vector<float> f(sv.size());
        for(size_t i = 0; i < f.size(); i++) f[i] = (float)sv[i];  // exact for values saved from FP32
        MONO_CHECK(hipMemcpyAsync(t->sv, f.data(), state_bytes(t), hipMemcpyHostToDevice, t->stream));
        MONO_CHECK(hipStreamSynchronize(t->stream));
    }
    return 0;
}

template <class T> int download(mono_tissue *t, const T *dev, size_t count, std::vector<T> &out) {
    out.resize(count);
    if(count == 0) return 0;
    MONO_CHECK(hipMemcpyAsync(out.data(), dev, sizeof(T) * count, hipMemcpyDeviceToHost, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

template <class T> int upload(mono_tissue *t, T *dev, const std::vector<T> &in) {
    if(in.empty()) return 0;
    MONO_CHECK(hipMemcpyAsync(dev, in.data(), sizeof(T) * in.size(), hipMemcpyHostToDevice, t->stream));
    MONO_CHECK(hipStreamSynchronize(t->stream));
    return 0;
}

// the octree cube the reference's structured domains are cut from: the mesh
// extent rounded up to a power-of-two number of cells (a cube only for equal
// spacings; otherwise the mesh extent)
void cube_side(const mono_tissue_desc &d, double out[3]) {
    if(d.dx == d.dy && d.dy == d.dz) {
        int64_t m = d.nx > d.ny ? d.nx : d.ny;
        m = m > d.nz ? m : d.nz;
        int64_t p = 1;
        while(p < m) p <<= 1;
        out[0] = out[1] = out[2] = (double)p * d.dx;
    } else {
        out[0] = d.nx * d.dx;
        out[1] = d.ny * d.dy;
        out[2] = d.nz * d.dz;
    }
}

int write_fail(const File &f) {
    mono_set_error("mono_tissue_save_state: cannot write %s", f.path.c_str());
    return -1;
}

int read_fail(const File &f, const char *what) {
    mono_set_error("mono_tissue_restore_state: %s: %s", f.path.c_str(), what);
    return -1;
}

}  // namespace

extern "C" {

int mono_tissue_save_state(mono_tissue *t, const char *dir) {
    if(!t || !dir) {
        mono_set_error("mono_tissue_save_state: null argument");
        return -1;
    }
    const mono_tissue_desc &d = t->d;
    const int64_t n = t->n;
    std::vector<double> v, b, sv, aux;
    std::vector<uint8_t> act;
    if(download(t, t->v, n, v) || download(t, t->b, n, b) || download_states(t, sv)) return -1;
    if(t->aux && download(t, t->aux, 3 * t->pitch, aux)) return -1;
    if(t->active && download(t, t->active, n, act)) return -1;
    auto is_active = [&](int64_t i) { return act.empty() || act[i] != 0; };
    std::vector<int64_t> cells;  // active cells, lexicographic
    for(int64_t i = 0; i < n; i++)
        if(is_active(i)) cells.push_back(i);
    const uint32_t na = (uint32_t)cells.size();

    {  // grid
        File f;
        if(!f.open(dir, "grid_checkpoint.dat", "wb")) return write_fail(f);
        double cube[3], mesh[3] = {d.nx * d.dx, d.ny * d.dy, d.nz * d.dz};
        cube_side(d, cube);
        bool ok = f.put(cube, sizeof(cube)) && f.put(mesh, sizeof(mesh)) && f.put((uint32_t)n) && f.put(na);
        const double zero6[6] = {0, 0, 0, 0, 0, 0};
        for(int64_t i = 0; ok && i < n; i++) {
            const int64_t x = i % d.nx, y = (i / d.nx) % d.ny, z = i / ((int64_t)d.nx * d.ny);
            const double c[3] = {(x + 0.5) * d.dx, (y + 0.5) * d.dy, (z + 0.5) * d.dz};
            ok = f.put(c, sizeof(c)) && f.put(v[i]) && f.put(zero6, sizeof(zero6)) && f.put(b[i]) && f.put((uint8_t)1) &&
                 f.put((uint8_t)is_active(i)) && f.put((uint64_t)0);
        }
        if(!ok || !f.close()) return write_fail(f);
    }
    {  // monodomain solver + time info
        File f;
        if(!f.open(dir, "monodomain_solver_checkpoint.dat", "wb")) return write_fail(f);
        const double zero = 0.0;
        // num_threads, final_time (the configuration's, not known here), beta,
        // cm, refine_each, derefine_each, refinement/derefinement bounds,
        // start_adapting_at, abort_on_no_activity, only_abort_after_dt, dt
        bool ok = f.put((int32_t)1) && f.put((int32_t)0) && f.put(zero) && f.put(d.beta) && f.put(d.cm) && f.put((int32_t)0) &&
                  f.put((int32_t)0) && f.put(zero) && f.put(zero) && f.put(zero) && f.put((uint8_t)0) &&
                  f.put("\0\0\0\0\0\0\0", 7) && f.put(zero) && f.put(d.dt_pde);
        ok = ok && f.put(t->cur_time) && f.put(zero) && f.put(d.dt_pde) && f.put((int32_t)t->st.steps) && f.put((int32_t)0);
        if(!ok || !f.close()) return write_fail(f);
    }
    {  // ODE solver states
        File f;
        if(!f.open(dir, "ode_solver_checkpoint.dat", "wb")) return write_fail(f);
        const uint8_t adaptive = d.ode_adaptive ? 1 : 0;
        const double max_dt = adaptive ? d.dt_pde : 0.0;
        const double rel = adaptive ? d.ode_rel_tol : 0.0, abs_ = adaptive ? d.ode_abs_tol : 0.0;
        int dev = 0;
        MONO_CHECK(hipGetDevice(&dev));
        bool ok = f.put(adaptive) && f.put(max_dt) && f.put(t->dt_ode) && f.put(rel) && f.put(abs_) && f.put((uint64_t)na) &&
                  f.put((uint8_t)0) && f.put((uint8_t)1) && f.put((int32_t)dev) && f.put((uint64_t)na * sizeof(double)) && f.put(na);
        std::vector<double> row(na);
        const int rows = t->neq + (adaptive ? 3 : 0);
        for(int s = 0; ok && s < rows; s++) {
            const double *src = s < t->neq ? sv.data() + (size_t)s * t->pitch : aux.data() + (size_t)(s - t->neq) * t->pitch;
            for(uint32_t k = 0; k < na; k++) row[k] = src[cells[k]];
            ok = f.put(row.data(), sizeof(double) * na);
        }
        if(!ok || !f.close()) return write_fail(f);
    }
    {  // this library's own state
        File f;
        if(!f.open(dir, "hip_solver_checkpoint.dat", "wb")) return write_fail(f);
        const std::vector<int> &sched = t->cg.last_sched;
        bool ok = f.put(kMagic, 8) && f.put(kVersion) && f.put((int32_t)d.precision) && f.put((int32_t)d.model) && f.put(n) &&
                  f.put(t->cur_time) && f.put(t->stim_start) && f.put(t->st) && f.put((int32_t)t->cg.last_iters) && f.put((int32_t)t->cg.last_inner) &&
                  f.put((int32_t)sched.size()) && (sched.empty() || f.put(sched.data(), sizeof(int) * sched.size()));
        const Observer &o = t->obs;
        const int32_t has_obs = o.n_act != nullptr;
        ok = ok && f.put(has_obs);
        if(ok && has_obs) {
            const size_t kk = (size_t)(o.max_events > 0 ? o.max_events : 1);
            std::vector<float> last_v, was, ac, ap;
            std::vector<double> last_vd;
            std::vector<int32_t> nact, napd;
            if(download(t, o.last_v, n, last_v) || download(t, o.was_active, n, was) || download(t, o.n_act, n, nact) ||
               download(t, o.n_apd, n, napd) || download(t, o.act, kk * n, ac) || download(t, o.apd, kk * n, ap))
                return -1;
            if(o.vm_double && download(t, o.last_vd, n, last_vd)) return -1;
            ok = f.put(o.act_thr) && f.put(o.apd_thr) && f.put(o.time_thr) && f.put((int32_t)o.max_events) &&
                 f.put((int32_t)o.vm_double) && f.put(last_v.data(), 4 * n) && f.put(was.data(), 4 * n) &&
                 f.put(nact.data(), 4 * n) && f.put(napd.data(), 4 * n) && f.put(ac.data(), 4 * kk * n) &&
                 f.put(ap.data(), 4 * kk * n) && (!o.vm_double || f.put(last_vd.data(), 8 * n));
        }
        if(!ok || !f.close()) return write_fail(f);
    }
    return 0;
}

int mono_tissue_restore_state(mono_tissue *t, const char *dir) {
    if(!t || !dir) {
        mono_set_error("mono_tissue_restore_state: null argument");
        return -1;
    }
    const mono_tissue_desc &d = t->d;
    const int64_t n = t->n;
    std::vector<uint8_t> act;
    if(t->active && download(t, t->active, n, act)) return -1;
    std::vector<double> v;
    if(download(t, t->v, n, v)) return -1;

    double cur_time = 0.0;
    int32_t iteration = 0;
    {  // monodomain solver: only the time info is used
        File f;
        if(!f.open(dir, "monodomain_solver_checkpoint.dat", "rb")) return read_fail(f, "cannot open");
        if(f.size() != (long)(kSolverBytes + kTimeInfoBytes)) return read_fail(f, "not a monodomain solver checkpoint (size)");
        fseek(f.f, kSolverBytes, SEEK_SET);
        double final_t, dt;
        if(!f.get(cur_time) || !f.get(final_t) || !f.get(dt) || !f.get(iteration)) return read_fail(f, "short read");
        if(iteration < 0) return read_fail(f, "negative iteration count");
        if(dt != d.dt_pde)
            return read_fail(f, "saved with another dt_pde");
    }

    std::vector<int64_t> cells;  // column k of the states -> cell of this tissue
    {  // grid
        File f;
        if(!f.open(dir, "grid_checkpoint.dat", "rb")) return read_fail(f, "cannot open");
        double hdr[6];
        uint32_t nc, na;
        if(!f.get(hdr, sizeof(hdr)) || !f.get(nc) || !f.get(na)) return read_fail(f, "short header");
        std::vector<uint8_t> seen(n, 0);
        for(uint32_t r = 0; r < nc; r++) {
            double c[3], cv, flux[6], cb;
            uint8_t can_change, active;
            uint64_t extra;
            if(!f.get(c, sizeof(c)) || !f.get(cv) || !f.get(flux, sizeof(flux)) || !f.get(cb) || !f.get(can_change) ||
               !f.get(active) || !f.get(extra))
                return read_fail(f, "short cell record");
            if(extra) fseek(f.f, (long)extra, SEEK_CUR);
            if(!active) continue;
            const double fx = c[0] / d.dx - 0.5, fy = c[1] / d.dy - 0.5, fz = c[2] / d.dz - 0.5;
            const int64_t x = llround(fx), y = llround(fy), z = llround(fz);
            if(fabs(fx - x) > 1e-6 || fabs(fy - y) > 1e-6 || fabs(fz - z) > 1e-6 || x < 0 || y < 0 || z < 0 || x >= d.nx ||
               y >= d.ny || z >= d.nz) {
                mono_set_error("mono_tissue_restore_state: active cell centred at (%g, %g, %g) is not a cell of this tissue",
                               c[0], c[1], c[2]);
                return -1;
            }
            const int64_t i = x + (int64_t)d.nx * (y + (int64_t)d.ny * z);
            if(!act.empty() && !act[i]) {
                mono_set_error("mono_tissue_restore_state: cell (%lld, %lld, %lld) is active in the checkpoint, inactive here",
                               (long long)x, (long long)y, (long long)z);
                return -1;
            }
            if(seen[i]++) return read_fail(f, "a cell appears twice");
            v[i] = cv;
            cells.push_back(i);
        }
        if(cells.size() != na) return read_fail(f, "active cell count disagrees with its header");
        for(int64_t i = 0; i < n; i++)
            if((act.empty() || act[i]) && !seen[i]) {
                mono_set_error("mono_tissue_restore_state: %lld active cells of this tissue are not in the checkpoint",
                               (long long)(n - (int64_t)cells.size()));
                return -1;
            }
    }

    std::vector<double> sv, aux;
    if(download_states(t, sv)) return -1;
    if(t->aux && download(t, t->aux, 3 * t->pitch, aux)) return -1;
    bool aux_restored = false;
    {  // ODE solver states
        File f;
        if(!f.open(dir, "ode_solver_checkpoint.dat", "rb")) return read_fail(f, "cannot open");
        uint8_t adaptive, flag, gpu;
        double max_dt, min_dt, rel, abs_;
        uint64_t ncts, pitch;
        int32_t gpu_id;
        uint32_t N;
        if(!f.get(adaptive) || !f.get(max_dt) || !f.get(min_dt) || !f.get(rel) || !f.get(abs_) || !f.get(ncts) || !f.get(flag))
            return read_fail(f, "short header");
        if(flag) return read_fail(f, "holds a cells-to-solve list (adaptive meshes are not restored)");
        if(!f.get(gpu) || !f.get(gpu_id) || !f.get(pitch) || !f.get(N)) return read_fail(f, "short header");
        if(N != cells.size()) return read_fail(f, "state count disagrees with the grid checkpoint's active cells");
        if((adaptive != 0) != (d.ode_adaptive != 0))
            return read_fail(f, adaptive ? "saved by the adaptive ODE solver, this tissue uses a fixed step"
                                         : "saved by the fixed-step ODE solver, this tissue is adaptive");
        const int rows = t->neq + (adaptive && gpu ? 3 : 0);
        const long body = f.size() - ftell(f.f);
        size_t w;
        if(N > 0 && body == (long)((size_t)rows * N * sizeof(double))) w = sizeof(double);
        else if(N > 0 && body == (long)((size_t)rows * N * sizeof(float))) w = sizeof(float);
        else {
            mono_set_error("mono_tissue_restore_state: %s: %ld bytes of states for %u cells do not match %d states per cell",
                           f.path.c_str(), body, N, rows);
            return -1;
        }
        std::vector<char> raw(body);
        if(!f.get(raw.data(), raw.size())) return read_fail(f, "short states");
        auto val = [&](size_t k) {
            if(w == sizeof(double)) {
                double x;
                memcpy(&x, raw.data() + 8 * k, 8);
                return x;
            }
            float x;
            memcpy(&x, raw.data() + 4 * k, 4);
            return (double)x;
        };
        for(int s = 0; s < rows; s++)
            for(uint32_t k = 0; k < N; k++) {
                const double x = gpu ? val((size_t)s * N + k) : val((size_t)k * t->neq + s);
                if(s < t->neq) sv[(size_t)s * t->pitch + cells[k]] = x;
                else aux[(size_t)(s - t->neq) * t->pitch + cells[k]] = x;
            }
        aux_restored = adaptive && gpu;
    }
    if(t->aux && !aux_restored) {
        // a checkpoint of the reference's CPU adaptive solver keeps no per-cell
        // step sizes: start every cell from this time with the minimum step
        for(int64_t i = 0; i < n; i++) {
            aux[i] = t->dt_ode;
            aux[t->pitch + i] = cur_time;
            aux[2 * t->pitch + i] = 0.0;
        }
    }

    // this library's own state, when present
    double stim_start = d.stim_start;
    mono_stats st = t->st;
    st.steps = iteration;
    st.time = cur_time;
    st.cg_iterations_total = st.inner_iterations_total = st.outer_sweeps_total = 0;
    st.nan_detected = 0;
    int32_t last_iters = 0, last_inner = 0;
    std::vector<int> sched;
    struct ObsData {
        float a = 0, p = 0, tt = 0;
        int32_t kmax = 0, vmd = 0;
        std::vector<float> last_v, was, ac, ap;
        std::vector<double> last_vd;
        std::vector<int32_t> nact, napd;
    } od;
    int32_t has_obs = 0;
    {
        File f;
        if(f.open(dir, "hip_solver_checkpoint.dat", "rb")) {
            char magic[8];
            uint32_t version;
            int32_t prec, model, nsched;
            int64_t fn;
            double ft;
            mono_stats fst;
            if(!f.get(magic, 8) || memcmp(magic, kMagic, 8) != 0) return read_fail(f, "not a checkpoint of this library");
            if(!f.get(version) || version != kVersion) return read_fail(f, "unknown version");
            if(!f.get(prec) || !f.get(model) || !f.get(fn) || !f.get(ft) || !f.get(stim_start) || !f.get(fst) ||
               !f.get(last_iters) || !f.get(last_inner) || !f.get(nsched))
                return read_fail(f, "short read");
            if(model != d.model || fn != n) return read_fail(f, "saved from a tissue of another model or size");
            if(ft != cur_time || fst.steps != iteration) return read_fail(f, "time disagrees with the solver checkpoint");
            if(nsched < 0 || nsched > 1 << 16) return read_fail(f, "bad sweep schedule length");
            sched.resize(nsched);
            if(nsched && !f.get(sched.data(), sizeof(int) * nsched)) return read_fail(f, "short read");
            // the schedule and statistics belong to the precision that wrote them
            if(prec != d.precision) {
                sched.clear();
                last_iters = last_inner = 0;
            } else {
                fst.class_operator = st.class_operator;
                fst.stencil_points = st.stencil_points;
                st = fst;
            }
            if(!f.get(has_obs)) return read_fail(f, "short read");
            if(has_obs) {
                if(!f.get(od.a) || !f.get(od.p) || !f.get(od.tt) || !f.get(od.kmax) || !f.get(od.vmd) || od.kmax < 0 ||
                   od.kmax > 1 << 16)
                    return read_fail(f, "bad observer header");
                const size_t kk = (size_t)(od.kmax > 0 ? od.kmax : 1);
                od.last_v.resize(n);
                od.was.resize(n);
                od.nact.resize(n);
                od.napd.resize(n);
                od.ac.resize(kk * n);
                od.ap.resize(kk * n);
                bool ok = f.get(od.last_v.data(), 4 * n) && f.get(od.was.data(), 4 * n) && f.get(od.nact.data(), 4 * n) &&
                          f.get(od.napd.data(), 4 * n) && f.get(od.ac.data(), 4 * kk * n) && f.get(od.ap.data(), 4 * kk * n);
                if(ok && od.vmd) {
                    od.last_vd.resize(n);
                    ok = f.get(od.last_vd.data(), 8 * n);
                }
                if(!ok) return read_fail(f, "short observer arrays");
            }
        }
    }

    // everything read and checked: install it
    if(upload(t, t->v, v) || upload_states(t, sv)) return -1;
    if(t->aux && upload(t, t->aux, aux)) return -1;
    if(has_obs) {
        Observer &o = t->obs;
        if(o.init(n, od.a, od.p, od.tt, od.kmax, od.vmd, t->stream)) return -1;
        if(upload(t, o.last_v, od.last_v) || upload(t, o.was_active, od.was) || upload(t, o.n_act, od.nact) ||
           upload(t, o.n_apd, od.napd) || upload(t, o.act, od.ac) || upload(t, o.apd, od.ap))
            return -1;
        if(od.vmd && upload(t, o.last_vd, od.last_vd)) return -1;
    }
    t->cur_time = cur_time;
    t->stim_start = stim_start;
    t->st = st;
    t->cg.last_sched = sched;
    t->cg.last_iters = last_iters;
    t->cg.last_inner = last_inner;
    return 0;
}

}  // extern "C"
