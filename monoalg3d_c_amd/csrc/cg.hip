// Device-resident conjugate gradient for the implicit diffusion solve.
//
// Control flow and stopping rule follow cpu_conjugate_gradient
// (linear_system_solver.c:47-183): r = b - A x0, p = r (or z = D^-1 r), then
// q = A p, alpha = r.z / p.q, x += alpha p, r -= alpha q, beta = r1.z1 / r.z,
// iteration count starting at 1, stop when r.r <= tol or the count reaches
// max_its.  Each reduction is fused into the kernel that produces its
// operands; partial sums are combined by the last block to finish (fixed order,
// so results are run-to-run deterministic), which also computes the scalars for
// the next kernel.  Every kernel reads a device `done` flag first, so the host
// can enqueue several iterations ahead without synchronising.
//
// Mixed precision (FP32 CG with FP64 defect correction): the outer residual
// r = b - A x is formed in FP64 with the FP64 operator; an FP32 CG solves
// A d = r to a relative tolerance with FP64 dot accumulation; x += d in FP64.
#include <type_traits>

#include "cg.h"

namespace mono {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;
constexpr int kMaxRed = 3;

__host__ __device__ __forceinline__ int pos_cls(int i, int len) { return i == 0 ? 0 : (i == len - 1 ? 2 : 1); }

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous eighth of the chunks so stencil neighbours share its L2.
__device__ __forceinline__ int xcd_chunk(int b, int nblocks) {
    if(nblocks % 8 != 0) return b;
    return (b % 8) * (nblocks / 8) + b / 8;
}

__device__ __forceinline__ void chunk_range(int64_t n, int64_t &beg, int64_t &end) {
    const int nb = gridDim.x;
    const int L = xcd_chunk(blockIdx.x, nb);
    int64_t per = (n + nb - 1) / nb;
    per = (per + kBlock - 1) / kBlock * kBlock;
    beg = (int64_t)L * per;
    end = beg + per < n ? beg + per : n;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

// Block-wide sums of up to kMaxRed values; result valid in thread 0.
template <int NR> __device__ __forceinline__ void block_sum(double (&v)[NR]) {
    __shared__ double lds[NR][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for(int k = 0; k < NR; k++) {
        const double s = wave_sum(v[k]);
        if(lane == 0) lds[k][w] = s;
    }
    __syncthreads();
    if(threadIdx.x == 0) {
#pragma unroll
        for(int k = 0; k < NR; k++) {
            double s = 0.0;
            for(int i = 0; i < kBlock / 64; i++) s += lds[k][i];
            v[k] = s;
        }
    }
}

// Publish this block's partial sums; returns true (in every thread) in the
// last block to finish, whose thread 0 then holds the grid totals in v.
template <int NR>
__device__ __forceinline__ bool grid_sum(double (&v)[NR], double *partials, unsigned *counter) {
    block_sum<NR>(v);
    __shared__ int is_last;
    if(threadIdx.x == 0) {
#pragma unroll
        for(int k = 0; k < NR; k++)
            __hip_atomic_store(&partials[k * kMaxBlocks + blockIdx.x], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if(!is_last) return false;
    // fixed-order combination of the partials
    double acc[NR];
#pragma unroll
    for(int k = 0; k < NR; k++) acc[k] = 0.0;
    for(int i = threadIdx.x; i < (int)gridDim.x; i += kBlock) {
#pragma unroll
        for(int k = 0; k < NR; k++)
            acc[k] += __hip_atomic_load(&partials[k * kMaxBlocks + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    block_sum<NR>(acc);
    if(threadIdx.x == 0) {
#pragma unroll
        for(int k = 0; k < NR; k++) v[k] = acc[k];
        *counter = 0u;
    }
    return true;
}

// Stencil walk.  Blocks are dispatched round-robin over the 8 XCDs (block b runs
// on XCD b % 8).  XCD k owns the y-strip [y0, y1) of every z plane and its
// blocks walk that strip plane by plane in 256-cell chunks, so the z- and
// y-neighbours a row product reads were fetched into the same XCD's L2 shortly
// before.  Cell coordinates advance incrementally (no divisions per cell).
struct Walk {
    int x, y, z, y0, y1;
    int64_t l, end;
};

__device__ __forceinline__ bool walk_begin(int nx, int ny, int nz, Walk &w) {
    const int nxcd = (gridDim.x % 8 == 0 && ny >= 8) ? 8 : 1;
    const int k = blockIdx.x % nxcd, j = blockIdx.x / nxcd, nbk = gridDim.x / nxcd;
    w.y0 = (int)((int64_t)ny * k / nxcd);
    w.y1 = (int)((int64_t)ny * (k + 1) / nxcd);
    const int64_t spp = (int64_t)(w.y1 - w.y0) * nx;
    const int64_t total = spp * nz;
    int64_t per = (total + nbk - 1) / nbk;
    per = (per + kBlock - 1) / kBlock * kBlock;
    w.l = (int64_t)j * per + threadIdx.x;
    w.end = (int64_t)(j + 1) * per < total ? (int64_t)(j + 1) * per : total;
    if(w.l >= w.end) return false;
    w.z = (int)(w.l / spp);
    const int64_t off = w.l - (int64_t)w.z * spp;
    const int yy = (int)(off / nx);
    w.x = (int)(off - (int64_t)yy * nx);
    w.y = w.y0 + yy;
    return true;
}

__device__ __forceinline__ bool walk_next(int nx, Walk &w) {
    w.l += kBlock;
    if(w.l >= w.end) return false;
    w.x += kBlock;
    while(w.x >= nx) {
        w.x -= nx;
        if(++w.y >= w.y1) {
            w.y = w.y0;
            ++w.z;
        }
    }
    return true;
}

// Axis fast range for the interior row path: 1..L-2, or the only cell when L == 1.
__device__ __forceinline__ bool in_fast(int i, int L) { return L == 1 ? true : (i > 0 && i < L - 1); }

// y_c = (A x)_c over the stored slots; diagonal first, as the reference forms
// its row products (linear_system_solver.c:78-85).  Slot loops have a
// compile-time length (S or NI) and no data-dependent branches, so all the
// neighbour loads of a row can be in flight together.  When every lane of the
// wave holds a cell of one z layer inside the fast ranges (the common case),
// the coefficient row is wave-uniform; otherwise each lane reads its class row
// and out-of-box slots use the cell itself with a zero coefficient.
template <typename TV, typename TC, int S, int NI>
__device__ __forceinline__ TV stencil_row(const OpView<TC> &op, const TV *__restrict__ x, int64_t c, int xi, int yi, int zi) {
    const bool fast = in_fast(xi, op.nx) && in_fast(yi, op.ny) && in_fast(zi, op.nz);
    TV acc = TV(0);
    if(op.table) {
        const int z0 = __builtin_amdgcn_readfirstlane(zi);
        if(__all(fast && zi == z0)) {
            const TC *row = op.table_int + z0 * NI;
#pragma unroll
            for(int k = 0; k < NI; k++) acc += (TV)row[k] * x[c + op.lin_int[k]];
            return acc;
        }
        const TC *row = op.table + ((zi * 3 + pos_cls(yi, op.ny)) * 3 + pos_cls(xi, op.nx)) * S;
#pragma unroll
        for(int k = 0; k < S; k++) {
            const int xx = xi + op.off[k][0], yy = yi + op.off[k][1], zz = zi + op.off[k][2];
            const bool in = xx >= 0 && yy >= 0 && zz >= 0 && xx < op.nx && yy < op.ny && zz < op.nz;
            const TC a = in ? row[k] : TC(0);
            acc += (TV)a * x[in ? c + op.lin[k] : c];
        }
        return acc;
    }
    const TC *col = op.percell + c;
    const int64_t n = op.n;
    if(__all(fast)) {
#pragma unroll
        for(int k = 0; k < S; k++) acc += (TV)col[k * n] * x[c + op.lin[k]];
        return acc;
    }
#pragma unroll
    for(int k = 0; k < S; k++) {
        const int xx = xi + op.off[k][0], yy = yi + op.off[k][1], zz = zi + op.off[k][2];
        const bool in = xx >= 0 && yy >= 0 && zz >= 0 && xx < op.nx && yy < op.ny && zz < op.nz;
        const TC a = in ? col[k * n] : TC(0);
        acc += (TV)a * x[in ? c + op.lin[k] : c];
    }
    return acc;
}

template <typename TC> __device__ __forceinline__ double diag_of(const OpView<TC> &op, int64_t c) {
    if(op.table) {
        const int xi = (int)(c % op.nx);
        const int64_t t = c / op.nx;
        const int yi = (int)(t % op.ny), zi = (int)(t / op.ny);
        return (double)op.table[((zi * 3 + pos_cls(yi, op.ny)) * 3 + pos_cls(xi, op.nx)) * op.nslots];
    }
    return (double)op.percell[c];
}

__device__ __forceinline__ bool is_active(const uint8_t *a, int64_t c) { return !a || a[c]; }

// ---------------------------------------------------------------------------
// FP64 CG (plain path): x is the FP64 potential.
// ---------------------------------------------------------------------------
template <int S, int NI>
__global__ __launch_bounds__(kBlock) void k_res64(OpView<double> op, const double *__restrict__ x, const double *__restrict__ b,
                                                  double *__restrict__ r, double *__restrict__ p, double *__restrict__ z,
                                                  int jacobi, double tol, double rel_tol, int max_its, CgScalars *sc,
                                                  double *partials, unsigned *counter) {
    double v[3] = {0.0, 0.0, 0.0};  // r.r, r.z, b.b
    Walk w;
    if(walk_begin(op.nx, op.ny, op.nz, w)) do {
        const int64_t c = w.x + (int64_t)op.nx * (w.y + (int64_t)op.ny * w.z);
        const double ax = stencil_row<double, double, S, NI>(op, x, c, w.x, w.y, w.z);
        if(!is_active(op.active, c)) continue;
        const double bc = b[c];
        const double rc = bc - ax;
        r[c] = rc;
        if(jacobi) {
            double d = diag_of(op, c);
            if(d == 0.0) d = 1.0;
            const double zc = (1.0 / d) * rc;
            z[c] = zc;
            p[c] = zc;
            v[1] += rc * zc;
        } else {
            p[c] = rc;
        }
        v[0] += rc * rc;
        v[2] += bc * bc;
    } while(walk_next(op.nx, w));
    if(grid_sum<3>(v, partials, counter) && threadIdx.x == 0) {
        const double tol_abs = rel_tol > 0.0 ? rel_tol * rel_tol * v[2] : tol;
        sc->rr_true = v[0];
        sc->rr = jacobi ? v[1] : v[0];
        sc->rr0 = v[0];
        sc->tol_abs = tol_abs;
        sc->iters = 1;
        sc->max_its = max_its;
        sc->done = !(v[0] >= tol_abs) || max_its <= 1;
    }
}

template <typename TV, typename TC, int S, int NI>
__global__ __launch_bounds__(kBlock) void k_spmv_dot(OpView<TC> op, const TV *__restrict__ p, TV *__restrict__ q, CgScalars *sc,
                                                     double *partials, unsigned *counter) {
    if(sc->done) return;
    double v[1] = {0.0};
    Walk w;
    if(walk_begin(op.nx, op.ny, op.nz, w)) do {
        const int64_t c = w.x + (int64_t)op.nx * (w.y + (int64_t)op.ny * w.z);
        const TV qc = stencil_row<TV, TC, S, NI>(op, p, c, w.x, w.y, w.z);
        if(!is_active(op.active, c)) continue;
        q[c] = qc;
        v[0] += (double)p[c] * (double)qc;
    } while(walk_next(op.nx, w));
    if(grid_sum<1>(v, partials, counter) && threadIdx.x == 0) {
        sc->pAp = v[0];
        sc->alpha = sc->rr / v[0];
    }
}

// Elementwise kernels walk contiguous chunks with four cells in flight per
// thread (loads of the four are issued before any use).
constexpr int kIlp = 4;

template <typename TV, typename TX, typename TC>
__global__ __launch_bounds__(kBlock) void k_update(OpView<TC> op, TX *__restrict__ x, const TV *__restrict__ p, const TV *__restrict__ q,
                                                   TV *__restrict__ r, TV *__restrict__ z, int jacobi, CgScalars *sc,
                                                   double *partials, unsigned *counter) {
    if(sc->done) return;
    const double alpha = sc->alpha;
    const TV a = (TV)alpha;
    int64_t beg, end;
    chunk_range(op.n, beg, end);
    double v[2] = {0.0, 0.0};  // r.r, r.z
    for(int64_t c0 = beg + threadIdx.x; c0 < end; c0 += kIlp * kBlock) {
        int64_t cc[kIlp];
        TX xv[kIlp];
        TV pv[kIlp], qv[kIlp], rv[kIlp];
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            cc[u] = c < end ? c : end - 1;
            xv[u] = x[cc[u]];
            pv[u] = p[cc[u]];
            qv[u] = q[cc[u]];
            rv[u] = r[cc[u]];
        }
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            if(c >= end || !is_active(op.active, c)) continue;
            x[c] = xv[u] + (TX)(a * pv[u]);
            const TV rc = rv[u] - a * qv[u];
            r[c] = rc;
            if(jacobi) {
                double d = diag_of(op, c);
                if(d == 0.0) d = 1.0;
                const TV zc = (TV)(1.0 / d) * rc;
                z[c] = zc;
                v[1] += (double)zc * (double)rc;
            }
            v[0] += (double)rc * (double)rc;
        }
    }
    if(grid_sum<2>(v, partials, counter) && threadIdx.x == 0) {
        const double rz_new = jacobi ? v[1] : v[0];
        sc->beta = rz_new / sc->rr;
        sc->rr = rz_new;
        sc->rr_true = v[0];
        sc->iters = sc->iters + 1;
        sc->done = (v[0] <= sc->tol_abs) || (sc->iters >= sc->max_its) || (v[0] != v[0]);
    }
}

template <typename TV, typename TC>
__global__ __launch_bounds__(kBlock) void k_dir(OpView<TC> op, TV *__restrict__ p, const TV *__restrict__ rz, CgScalars *sc) {
    if(sc->done) return;
    const TV beta = (TV)sc->beta;
    int64_t beg, end;
    chunk_range(op.n, beg, end);
    for(int64_t c0 = beg + threadIdx.x; c0 < end; c0 += kIlp * kBlock) {
        int64_t cc[kIlp];
        TV pv[kIlp], rv[kIlp];
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            cc[u] = c < end ? c : end - 1;
            pv[u] = p[cc[u]];
            rv[u] = rz[cc[u]];
        }
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            if(c >= end || !is_active(op.active, c)) continue;
            p[c] = rv[u] + beta * pv[u];
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed precision: outer residual in FP64, FP32 inner CG, FP64 correction.
// ---------------------------------------------------------------------------
template <int S, int NI>
__global__ __launch_bounds__(kBlock) void k_res_outer(OpView<double> op, const double *__restrict__ x, const double *__restrict__ b,
                                                      float *__restrict__ r32, float *__restrict__ p32, float *__restrict__ d32,
                                                      int first, double tol, double rel_tol, int max_outer, double inner_rel,
                                                      int max_inner, CgScalars *so, CgScalars *si, double *partials,
                                                      unsigned *counter) {
    if(!first && so->done) return;
    double v[2] = {0.0, 0.0};  // r.r, b.b
    Walk w;
    if(walk_begin(op.nx, op.ny, op.nz, w)) do {
        const int64_t c = w.x + (int64_t)op.nx * (w.y + (int64_t)op.ny * w.z);
        const double ax = stencil_row<double, double, S, NI>(op, x, c, w.x, w.y, w.z);
        if(!is_active(op.active, c)) continue;
        const double bc = b[c];
        const double rc = bc - ax;
        const float rf = (float)rc;
        r32[c] = rf;
        p32[c] = rf;
        d32[c] = 0.0f;
        v[0] += rc * rc;
        v[1] += bc * bc;
    } while(walk_next(op.nx, w));
    if(grid_sum<2>(v, partials, counter) && threadIdx.x == 0) {
        if(first) {
            so->tol_abs = rel_tol > 0.0 ? rel_tol * rel_tol * v[1] : tol;
            so->iters = 0;
            so->max_its = max_outer;
            so->rr0 = v[0];
        }
        so->rr_true = v[0];
        so->rr = v[0];
        so->done = !(v[0] > so->tol_abs) || so->iters >= so->max_its;
        so->iters = so->iters + 1;
        si->rr = v[0];
        si->rr_true = v[0];
        si->rr0 = v[0];
        // inner target: relative reduction, but no tighter than half the outer target
        const double t_rel = inner_rel * inner_rel * v[0];
        si->tol_abs = t_rel > 0.5 * so->tol_abs ? t_rel : 0.5 * so->tol_abs;
        si->iters = 0;
        si->max_its = max_inner;
        si->done = so->done;
    }
}

__global__ __launch_bounds__(kBlock) void k_correct(int64_t n, const uint8_t *active, double *__restrict__ x,
                                                    const float *__restrict__ d32, const CgScalars *so) {
    if(so->done) return;
    int64_t beg, end;
    chunk_range(n, beg, end);
    for(int64_t c0 = beg + threadIdx.x; c0 < end; c0 += kIlp * kBlock) {
        int64_t cc[kIlp];
        double xv[kIlp];
        float dv[kIlp];
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            cc[u] = c < end ? c : end - 1;
            xv[u] = x[cc[u]];
            dv[u] = d32[cc[u]];
        }
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            if(c >= end || !is_active(active, c)) continue;
            x[c] = xv[u] + (double)dv[u];
        }
    }
}

// ---------------------------------------------------------------------------
// CSR (non-uniform meshes), FP64: one wave per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_csr_spmv(int n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                     const double *__restrict__ val, const double *__restrict__ x,
                                                     double *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if(row >= n) return;
    double acc = 0.0;
    for(int64_t k = rp[row] + lane; k < rp[row + 1]; k += 64) acc += val[k] * x[col[k]];
    acc = wave_sum(acc);
    if(lane == 0) y[row] = acc;
}

// dot products and vector updates of the CSR solve reuse the stencil kernels
// through an operator view with no off-diagonal slots (only the active mask
// and the diagonal for Jacobi are consulted).

__global__ __launch_bounds__(kBlock) void k_csr_res(int n, const double *__restrict__ Ax, const double *__restrict__ b,
                                                    const double *__restrict__ diag, double *__restrict__ r, double *__restrict__ p,
                                                    double *__restrict__ z, int jacobi, double tol, double rel_tol, int max_its,
                                                    CgScalars *sc, double *partials, unsigned *counter) {
    int64_t beg, end;
    chunk_range(n, beg, end);
    double v[3] = {0.0, 0.0, 0.0};
    for(int64_t c = beg + threadIdx.x; c < end; c += kBlock) {
        const double rc = b[c] - Ax[c];
        r[c] = rc;
        if(jacobi) {
            double d = diag[c];
            if(d == 0.0) d = 1.0;
            const double zc = (1.0 / d) * rc;
            z[c] = zc;
            p[c] = zc;
            v[1] += rc * zc;
        } else {
            p[c] = rc;
        }
        v[0] += rc * rc;
        v[2] += b[c] * b[c];
    }
    if(grid_sum<3>(v, partials, counter) && threadIdx.x == 0) {
        const double tol_abs = rel_tol > 0.0 ? rel_tol * rel_tol * v[2] : tol;
        sc->rr_true = v[0];
        sc->rr = jacobi ? v[1] : v[0];
        sc->rr0 = v[0];
        sc->tol_abs = tol_abs;
        sc->iters = 1;
        sc->max_its = max_its;
        sc->done = !(v[0] >= tol_abs) || max_its <= 1;
    }
}

__global__ __launch_bounds__(kBlock) void k_dot_pq(int n, const double *__restrict__ p, const double *__restrict__ q, CgScalars *sc,
                                                   double *partials, unsigned *counter) {
    if(sc->done) return;
    int64_t beg, end;
    chunk_range(n, beg, end);
    double v[1] = {0.0};
    for(int64_t c = beg + threadIdx.x; c < end; c += kBlock) v[0] += p[c] * q[c];
    if(grid_sum<1>(v, partials, counter) && threadIdx.x == 0) {
        sc->pAp = v[0];
        sc->alpha = sc->rr / v[0];
    }
}

__global__ __launch_bounds__(kBlock) void k_csr_update(int n, double *__restrict__ x, const double *__restrict__ p,
                                                       const double *__restrict__ q, double *__restrict__ r, double *__restrict__ z,
                                                       const double *__restrict__ diag, int jacobi, CgScalars *sc,
                                                       double *partials, unsigned *counter) {
    if(sc->done) return;
    const double a = sc->alpha;
    int64_t beg, end;
    chunk_range(n, beg, end);
    double v[2] = {0.0, 0.0};
    for(int64_t c = beg + threadIdx.x; c < end; c += kBlock) {
        x[c] += a * p[c];
        const double rc = r[c] - a * q[c];
        r[c] = rc;
        if(jacobi) {
            double d = diag[c];
            if(d == 0.0) d = 1.0;
            const double zc = (1.0 / d) * rc;
            z[c] = zc;
            v[1] += zc * rc;
        }
        v[0] += rc * rc;
    }
    if(grid_sum<2>(v, partials, counter) && threadIdx.x == 0) {
        const double rz_new = jacobi ? v[1] : v[0];
        sc->beta = rz_new / sc->rr;
        sc->rr = rz_new;
        sc->rr_true = v[0];
        sc->iters = sc->iters + 1;
        sc->done = (v[0] <= sc->tol_abs) || (sc->iters >= sc->max_its) || (v[0] != v[0]);
    }
}

__global__ __launch_bounds__(kBlock) void k_csr_dir(int n, double *__restrict__ p, const double *__restrict__ rz, CgScalars *sc) {
    if(sc->done) return;
    const double beta = sc->beta;
    int64_t beg, end;
    chunk_range(n, beg, end);
    for(int64_t c = beg + threadIdx.x; c < end; c += kBlock) p[c] = rz[c] + beta * p[c];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <int A> using IC = std::integral_constant<int, A>;

// Call f(IC<S>, IC<NI>) for the operator's padded slot counts.
template <class F> static void with_slots(int S, int NI, F &&f) {
    if(S == 7) f(IC<7>{}, IC<7>{});
    else if(S == 19) {
        if(NI == 19) f(IC<19>{}, IC<19>{});
        else f(IC<19>{}, IC<7>{});
    } else {
        if(NI == 27) f(IC<27>{}, IC<27>{});
        else if(NI == 19) f(IC<27>{}, IC<19>{});
        else f(IC<27>{}, IC<7>{});
    }
}
static int grid_for(int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if(g > kMaxBlocks) g = kMaxBlocks;
    if(g >= 64) g = g / 8 * 8;  // multiple of 8 for the XCD remap
    return (int)(g < 1 ? 1 : g);
}

int CgWorkspace::init(int64_t n_, int64_t) {
    n = n_;
    const size_t b64 = sizeof(double) * (size_t)(n > 0 ? n : 1), b32 = sizeof(float) * (size_t)(n > 0 ? n : 1);
    MONO_CHECK(hipMalloc(&r, b64));
    MONO_CHECK(hipMalloc(&p, b64));
    MONO_CHECK(hipMalloc(&q, b64));
    MONO_CHECK(hipMalloc(&z, b64));
    MONO_CHECK(hipMalloc(&r32, b32));
    MONO_CHECK(hipMalloc(&p32, b32));
    MONO_CHECK(hipMalloc(&q32, b32));
    MONO_CHECK(hipMalloc(&d32, b32));
    // inactive cells are never written: keep them at zero
    MONO_CHECK(hipMemset(r, 0, b64));
    MONO_CHECK(hipMemset(p, 0, b64));
    MONO_CHECK(hipMemset(q, 0, b64));
    MONO_CHECK(hipMemset(z, 0, b64));
    MONO_CHECK(hipMemset(r32, 0, b32));
    MONO_CHECK(hipMemset(p32, 0, b32));
    MONO_CHECK(hipMemset(q32, 0, b32));
    MONO_CHECK(hipMemset(d32, 0, b32));
    MONO_CHECK(hipMalloc(&partials, sizeof(double) * kMaxBlocks * kMaxRed));
    MONO_CHECK(hipMalloc(&counter, sizeof(unsigned)));
    MONO_CHECK(hipMemset(counter, 0, sizeof(unsigned)));
    MONO_CHECK(hipMalloc(&sc, sizeof(CgScalars)));
    MONO_CHECK(hipMalloc(&sc_in, sizeof(CgScalars)));
    MONO_CHECK(hipMemset(sc, 0, sizeof(CgScalars)));
    MONO_CHECK(hipMemset(sc_in, 0, sizeof(CgScalars)));
    MONO_CHECK(hipHostMalloc(&host_sc, 2 * sizeof(CgScalars), hipHostMallocDefault));
    return 0;
}

int CgWorkspace::release() {
    double *d64[] = {r, p, q, z, partials};
    for(double *ptr : d64)
        if(ptr) MONO_CHECK(hipFree(ptr));
    float *d32s[] = {r32, p32, q32, d32};
    for(float *ptr : d32s)
        if(ptr) MONO_CHECK(hipFree(ptr));
    if(counter) MONO_CHECK(hipFree(counter));
    if(sc) MONO_CHECK(hipFree(sc));
    if(sc_in) MONO_CHECK(hipFree(sc_in));
    if(host_sc) MONO_CHECK(hipHostFree(host_sc));
    r = p = q = z = partials = nullptr;
    r32 = p32 = q32 = d32 = nullptr;
    counter = nullptr;
    sc = sc_in = host_sc = nullptr;
    return 0;
}

static int poll(CgScalars *dev, CgScalars *host, hipStream_t s) {
    MONO_CHECK(hipMemcpyAsync(host, dev, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    MONO_CHECK(hipStreamSynchronize(s));
    return 0;
}

int CgWorkspace::solve_fp64(const OpView<double> &op, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s) {
    const int grid = grid_for(op.n);
    with_slots(op.nslots, op.nint, [&](auto S_, auto NI_) {
        k_res64<decltype(S_)::value, decltype(NI_)::value><<<grid, kBlock, 0, s>>>(op, x, b, r, p, z, cfg.jacobi, cfg.tol, cfg.rel_tol,
                                                                                   cfg.max_its, sc, partials, counter);
    });
    MONO_CHECK(hipGetLastError());
    const double *rz = cfg.jacobi ? z : r;
    int enq = 0;
    int chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : (last_iters > 2 ? last_iters - 1 : 2);
    for(;;) {
        for(int k = 0; k < chunk; k++) {
            with_slots(op.nslots, op.nint, [&](auto S_, auto NI_) {
                k_spmv_dot<double, double, decltype(S_)::value, decltype(NI_)::value>
                    <<<grid, kBlock, 0, s>>>(op, p, q, sc, partials, counter);
            });
            k_update<double, double, double><<<grid, kBlock, 0, s>>>(op, x, p, q, r, z, cfg.jacobi, sc, partials, counter);
            k_dir<double, double><<<grid, kBlock, 0, s>>>(op, p, rz, sc);
        }
        MONO_CHECK(hipGetLastError());
        enq += chunk;
        if(poll(sc, host_sc, s)) return -1;
        if(host_sc->done || enq > cfg.max_its + 2) break;
        chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : 2;
    }
    last_iters = host_sc->iters;
    if(st) {
        st->iterations = host_sc->iters;
        st->error = host_sc->rr_true;
        st->outer = 1;
        st->inner_total = host_sc->iters - 1;
    }
    return 0;
}

int CgWorkspace::solve_mixed(const OpView<double> &op64, const OpView<float> &op32, const double *b, double *x, const CgConfig &cfg,
                             CgStats *st, hipStream_t s) {
    const int grid = grid_for(op64.n);
    int inner_total = 0;
    int chunk0 = last_inner > 2 ? last_inner - 1 : 4;
    CgScalars *h_out = host_sc, *h_in = host_sc + 1;
    for(int sweep = 0;; sweep++) {
        with_slots(op64.nslots, op64.nint, [&](auto S_, auto NI_) {
            k_res_outer<decltype(S_)::value, decltype(NI_)::value><<<grid, kBlock, 0, s>>>(
                op64, x, b, r32, p32, d32, sweep == 0, cfg.tol, cfg.rel_tol, cfg.max_outer, cfg.inner_rel, cfg.max_its, sc, sc_in,
                partials, counter);
        });
        MONO_CHECK(hipGetLastError());
        int chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : chunk0;
        int enq = 0;
        bool outer_done = false;
        for(;;) {
            for(int k = 0; k < chunk; k++) {
                with_slots(op32.nslots, op32.nint, [&](auto S_, auto NI_) {
                    k_spmv_dot<float, float, decltype(S_)::value, decltype(NI_)::value>
                        <<<grid, kBlock, 0, s>>>(op32, p32, q32, sc_in, partials, counter);
                });
                k_update<float, float, float><<<grid, kBlock, 0, s>>>(op32, d32, p32, q32, r32, nullptr, 0, sc_in, partials, counter);
                k_dir<float, float><<<grid, kBlock, 0, s>>>(op32, p32, r32, sc_in);
            }
            MONO_CHECK(hipGetLastError());
            enq += chunk;
            MONO_CHECK(hipMemcpyAsync(h_out, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
            MONO_CHECK(hipMemcpyAsync(h_in, sc_in, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
            MONO_CHECK(hipStreamSynchronize(s));
            if(h_out->done) {
                outer_done = true;
                break;
            }
            if(h_in->done || enq > cfg.max_its + 2) break;
            chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : 2;
        }
        if(outer_done) break;
        inner_total += h_in->iters;
        if(sweep == 0) last_inner = h_in->iters;
        k_correct<<<grid, kBlock, 0, s>>>(op64.n, op64.active, x, d32, sc);
        MONO_CHECK(hipGetLastError());
        chunk0 = 2;
    }
    if(st) {
        st->iterations = inner_total + 1;
        st->error = h_out->rr_true;
        st->outer = h_out->iters;
        st->inner_total = inner_total;
    }
    return 0;
}

int CgWorkspace::solve_csr(const CsrMatrix &A, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s) {
    const int grid = grid_for(A.n);
    const unsigned sgrid = (unsigned)(((int64_t)A.n * 64 + kBlock - 1) / kBlock);
    k_csr_spmv<<<sgrid, kBlock, 0, s>>>(A.n, A.row_ptr, A.col, A.val, x, q);
    k_csr_res<<<grid, kBlock, 0, s>>>(A.n, q, b, A.diag, r, p, z, cfg.jacobi, cfg.tol, cfg.rel_tol, cfg.max_its, sc, partials, counter);
    MONO_CHECK(hipGetLastError());
    const double *rz = cfg.jacobi ? z : r;
    int enq = 0;
    int chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : 4;
    for(;;) {
        for(int k = 0; k < chunk; k++) {
            k_csr_spmv<<<sgrid, kBlock, 0, s>>>(A.n, A.row_ptr, A.col, A.val, p, q);
            k_dot_pq<<<grid, kBlock, 0, s>>>(A.n, p, q, sc, partials, counter);
            k_csr_update<<<grid, kBlock, 0, s>>>(A.n, x, p, q, r, z, A.diag, cfg.jacobi, sc, partials, counter);
            k_csr_dir<<<grid, kBlock, 0, s>>>(A.n, p, rz, sc);
        }
        MONO_CHECK(hipGetLastError());
        enq += chunk;
        if(poll(sc, host_sc, s)) return -1;
        if(host_sc->done || enq > cfg.max_its + 2) break;
    }
    if(st) {
        st->iterations = host_sc->iters;
        st->error = host_sc->rr_true;
        st->outer = 1;
        st->inner_total = host_sc->iters - 1;
    }
    return 0;
}

}  // namespace mono
