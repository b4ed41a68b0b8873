// Device-resident conjugate gradient for the implicit diffusion solve.
//
// Control flow and stopping rule follow cpu_conjugate_gradient
// (linear_system_solver.c:47-183): r = b - A x0, p = r (or z = D^-1 r), then
// q = A p, alpha = r.z / p.q, x += alpha p, r -= alpha q, beta = r1.z1 / r.z,
// iteration count starting at 1, stop when r.r <= tol or the count reaches
// max_its.
//
// Reductions use no atomics: every block of a kernel writes its partial sums,
// and the next kernel, which needs the scalar anyway, adds the partials in a
// fixed order in each of its blocks (results are run-to-run deterministic).
// Block 0 of that kernel records the scalars and the convergence decision in a
// device CgScalars record; every later kernel of the solve reads its `done`
// flag first, so the host enqueues iterations ahead and only polls now and
// then.  (Device-scope atomics on one address from thousands of blocks cost
// hundreds of microseconds on this part; a second kernel launch would cost a
// few more.)
//
// Mixed precision (FP32 CG with FP64 defect correction): the outer residual
// r = b - A x is formed in FP64 with the FP64 operator; an FP32 CG solves
// A d = r to a relative tolerance with FP64 dot accumulation; x += d in FP64.
#include <stdlib.h>
#include <string.h>

#include <type_traits>
#include <vector>

#include "aniso_row.h"
#include "cg.h"

namespace mono {

constexpr int kBlock = 256;
#ifndef MONO_SLAB_RUNS
#define MONO_SLAB_RUNS 1
#endif
constexpr int kElemBlocks = 1024;  // grid of the elementwise kernels

__host__ __device__ __forceinline__ int pos_cls(int i, int len) { return i == 0 ? 0 : (i == len - 1 ? 2 : 1); }

// Elementwise kernels: contiguous chunks per block, chunks of one XCD adjacent.
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

// Block-wide sums; the totals are returned to every thread.
template <int NR> __device__ __forceinline__ void block_sum_all(double (&v)[NR]) {
    __shared__ double lds[NR][kBlock / 64];
    __shared__ double tot[NR];
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
            tot[k] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < NR; k++) v[k] = tot[k];
}

// This block's contribution to the kernel's reductions.
template <int NR> __device__ __forceinline__ void write_partials(double (&v)[NR], double *P) {
    block_sum_all<NR>(v);
    if(threadIdx.x == 0) {
#pragma unroll
        for(int k = 0; k < NR; k++) P[k * kMaxBlocks + blockIdx.x] = v[k];
    }
}

// The same into slot `slot` (a run of the plane-loop kernels).
template <int NR> __device__ __forceinline__ void write_partials_at(double (&v)[NR], double *P, int slot) {
    block_sum_all<NR>(v);
    if(threadIdx.x == 0) {
#pragma unroll
        for(int k = 0; k < NR; k++) P[k * kMaxBlocks + slot] = v[k];
    }
}

// Totals of the previous kernel's partials (fixed order; every thread gets them).
template <int NR> __device__ __forceinline__ void sum_partials(const double *__restrict__ P, int nprev, double (&tot)[NR]) {
#pragma unroll
    for(int k = 0; k < NR; k++) tot[k] = 0.0;
    for(int i = threadIdx.x; i < nprev; i += kBlock) {
#pragma unroll
        for(int k = 0; k < NR; k++) tot[k] += P[k * kMaxBlocks + i];
    }
    block_sum_all<NR>(tot);
}

// Parameters of one solve passed to its kernels.
struct CgSetup {
    double tol, rel_tol, inner_rel;
    int max_its, jacobi, max_outer, sweep;
    const double *tol_dev;  // if set: the r.r threshold, read on the device (CgConfig::tol_dev)
};

// Start of an iteration kernel: settle the reduction of the previous kernel.
//   mode 0: after the FP64 residual (P0 = r.r, r.z, b.b): tolerance, first test
//   mode 1: after the update of iteration k-1 (P0 = r.r, r.z): beta, stopping test
//   mode 2: after the mixed-precision outer residual (P0 = r.r, b.b): outer test
//           and the inner solve's target
// Returns false (in every thread) when the solve is finished; `now` is then
// true when this kernel made the decision (false: an earlier kernel did).
__device__ __forceinline__ bool iter_prologue(int mode, int k, const double *__restrict__ P0, int nprev, CgScalars *sc, CgScalars *so,
                                              const CgSetup &cs, double &beta, bool &now) {
    now = false;
    if(mode == 1) {
        // Block 0 of this very kernel may set `done` while other blocks start:
        // only a decision of an earlier kernel (done_iter < k) is taken from
        // the record, so every block of the deciding kernel decides itself
        // (the same way, from the same partials).  One read per block keeps
        // the outcome uniform across its waves.
        __shared__ int s_done;
        if(threadIdx.x == 0) s_done = sc->done && sc->done_iter < k;
        __syncthreads();
        if(s_done) return false;
    }
    double t[3];
    if(mode == 0) {
        sum_partials<3>(P0, nprev, t);
    } else {
        double t2[2];
        sum_partials<2>(P0, nprev, t2);
        t[0] = t2[0];
        t[1] = t2[1];
        t[2] = 0.0;
    }
    bool done;
    if(mode == 0) {
        const double rr = t[0], rz = cs.jacobi ? t[1] : t[0], bb = t[2];
        const double tol_abs = cs.tol_dev ? *cs.tol_dev : (cs.rel_tol > 0.0 ? cs.rel_tol * cs.rel_tol * bb : cs.tol);
        done = !(rr >= tol_abs) || cs.max_its <= 1;
        beta = 0.0;
        if(blockIdx.x == 0 && threadIdx.x == 0) {
            sc->tol_abs = tol_abs;
            sc->rr0 = rr;
            sc->rr_true = rr;
            sc->rrs[0] = rz;
            sc->iters = 1;
            sc->max_its = cs.max_its;
            sc->beta = 0.0;
            sc->done_iter = done ? k : 0x7fffffff;
            __threadfence();
            sc->done = done;
        }
    } else if(mode == 1) {
        const double rr = t[0], rz = cs.jacobi ? t[1] : t[0];
        const int iters = 1 + k;
        done = (rr <= sc->tol_abs) || (iters >= sc->max_its) || (rr != rr);
        beta = rz / sc->rrs[(k - 1) & 1];
        if(blockIdx.x == 0 && threadIdx.x == 0) {
            sc->rrs[k & 1] = rz;
            sc->rr_true = rr;
            sc->iters = iters;
            sc->beta = beta;
            // done_iter is written only with a decision, and before `done`:
            // a block that sees done = 1 of this kernel never pairs it with a
            // stale done_iter < k
            if(done) {
                sc->done_iter = k;
                __threadfence();
            }
            sc->done = done;
        }
    } else {
        const double rr = t[0], bb = t[1];
        const double tol_o = cs.sweep == 0 ? (cs.tol_dev ? *cs.tol_dev : (cs.rel_tol > 0.0 ? cs.rel_tol * cs.rel_tol * bb : cs.tol))
                                           : so->tol_abs;
        done = !(rr > tol_o) || cs.sweep >= cs.max_outer || (rr != rr);
        const double t_rel = cs.inner_rel * cs.inner_rel * rr;
        beta = 0.0;
        if(blockIdx.x == 0 && threadIdx.x == 0) {
            if(cs.sweep == 0) {
                so->tol_abs = tol_o;
                so->rr0 = rr;
            }
            so->rr_true = rr;
            so->iters = cs.sweep + 1;
            so->done = done;
            sc->tol_abs = t_rel > 0.5 * tol_o ? t_rel : 0.5 * tol_o;
            sc->rr0 = rr;
            sc->rr_true = rr;
            sc->rrs[0] = rr;
            sc->iters = 1;
            sc->max_its = cs.max_its;
            sc->beta = 0.0;
            sc->done_iter = done ? k : 0x7fffffff;
            __threadfence();
            sc->done = done;
        }
    }
    now = done;
    return !done;
}

template <typename TC> __device__ __forceinline__ double diag_of(const OpView<TC> &op, int64_t c) {
    if(op.table) {
        const int xi = (int)(c % op.nx);
        const int64_t t = c / op.nx;
        const int yi = (int)(t % op.ny), zi = (int)(t / op.ny);
        return (double)op.table[((zi * 3 + pos_cls(yi, op.ny)) * 3 + pos_cls(xi, op.nx)) * 27 + 13];
    }
    if(op.codes) return (double)op.ctab[(int)op.codes[c] * op.pat];
    return (double)op.percell[c];
}


// ---------------------------------------------------------------------------
// Tiled stencil sweep.  A block owns a 64 x 8 tile of an xy plane and marches
// through z planes, keeping planes z-1, z, z+1 of the tile (plus a one-cell
// halo, zero outside the box) in LDS; the loads of planes z+2 and z+3 are in
// flight while plane z is computed.  Rows follow compile-time patterns, so
// every LDS read uses an immediate offset.  A wave computes one 64-cell row
// segment of the tile.
//   per-cell mode: every cell reads its own row (pattern PAT);
//   class mode: interior cells use their layer's interior row (pattern PAT,
//     wave-uniform, in scalar registers); rows on a y face and the two
//     boundary planes use the 27-slot row of their class (also wave-uniform);
//     the cells on an x face (at most 16 per tile plane) are left out of the
//     wave-wide pass and evaluated by 16 lanes of wave 0 with per-lane rows.
// Work items are (xy tile, z plane) pairs in tile-major order; each block
// takes one contiguous run of `per` items (a few tiles' z ranges at most), and
// the grid is one wave of resident blocks, so all blocks finish together.
// With FUSE the value staged for the stencil is the new search direction
// p = rz + beta * p_in (written for the block's own cells to p_out, a buffer
// distinct from p_in, so halo reads never see updated values).
// ---------------------------------------------------------------------------
#ifndef MONO_TILE_X
#define MONO_TILE_X 64
#endif
#ifndef MONO_TILE_Y
#define MONO_TILE_Y 8
#endif
constexpr int kTX = MONO_TILE_X, kTY = MONO_TILE_Y, kW = kTX + 2, kH = kTY + 2, kPL = kW * kH;
static_assert(kBlock % kTX == 0 && (kTX * kTY) % kBlock == 0 && 2 * kTY <= kWave, "tile shape");
constexpr int kNLD = (kPL + kBlock - 1) / kBlock;
constexpr int kPLS = kNLD * kBlock;  // LDS plane stride: every lane of every load set has a slot
constexpr int kRPT = kTX * kTY / kBlock;  // rows per thread
#ifndef MONO_PERCELL_PREFETCH
#define MONO_PERCELL_PREFETCH 1
#endif

// Work split of the plane-loop kernels.  The items are (xy tile, z plane)
// pairs in tile-major order (a tile's planes are consecutive); run r is the
// contiguous item range [r * per, (r + 1) * per), so a run walks down one
// tile's planes and possibly continues on the next tile.  There are as many
// runs as the kernel has resident blocks (at least kMinRun items each), one
// per block, so all blocks finish together.  The resident count comes from
// the occupancy query for the kernel on this device, which is the same in
// every process: the runs, and so the order in which the next kernel adds
// their partial sums, are a fixed function of the box, the kernel and the
// device model, and a restarted run repeats the reductions bit for bit.
// Blocks of one XCD take adjacent runs, so the tiles in flight on an XCD are
// neighbours and their shared halo rows come from its L2.
struct TileGeom {
    int ntx, nty, nblocks;
    int ntiles, nruns, per;
    int L;                 // slab mode: planes per slab (0: runs are item ranges)
    int lo[3], hi[3];      // interior range of each axis (extent 1: its cell; extent 2: empty)
};

template <typename TC> using scalar_ptr = const __attribute__((address_space(4))) TC *;

// Wave-uniform rows that no kernel writes are read through the constant
// address space, so they land in scalar registers.
template <typename TC> __device__ __forceinline__ scalar_ptr<TC> as_scalar(const TC *p) { return (scalar_ptr<TC>)p; }

template <int P, typename TC, typename TV>
__device__ __forceinline__ TV row_lds(const TC *row, const TV *const (&P3)[3], int base) {
    TV acc = TV(0);
#pragma unroll
    for(int k = 0; k < P; k++)
        acc += (TV)row[k] * P3[RowPattern<P>::o[k][2] + 1][base + RowPattern<P>::o[k][1] * kW + RowPattern<P>::o[k][0]];
    return acc;
}

// 27-slot class row (slot = (oz+1)*9 + (oy+1)*3 + (ox+1)), staged in LDS,
// against the LDS planes: one oz plane (9 slots) at a time, so these rows do
// not raise the kernel's register count.
template <typename TC, typename TV>
__device__ __forceinline__ TV row27_lds(const TC *row, const TV *const (&P3)[3], int base) {
    TV acc = TV(0);
#pragma unroll 1
    for(int oz = 0; oz < 3; oz++) {
        const TV *P = (oz == 0 ? P3[0] : (oz == 1 ? P3[1] : P3[2])) + base;
        const TC *r = row + 9 * oz;
#pragma unroll
        for(int j = 0; j < 9; j++) acc += (TV)r[j] * P[(j / 3 - 1) * kW + (j % 3 - 1)];
    }
    return acc;
}

// Planes z0..z1-1 of xy tile t.
template <typename TV, typename TC, int CM, int PAT, bool FUSE, class Epi>
__device__ __forceinline__ void tile_segment(const OpView<TC> &op, const TileGeom &g, int t, int z0, int z1, const TV *__restrict__ x,
                                             const TV *__restrict__ rz, TV beta, TV *__restrict__ p_out, const Epi &epi,
                                             double (&v)[Epi::NR]) {
    static_assert(9 * 27 <= kBlock, "one class-row coefficient per thread");
    // CM: 0 per-cell rows, 1 class table, 2 code table (7-point rows), 3 rows
    // formed from neighbourhood masks (27-point)
    constexpr bool CLS = CM == 1;
    static_assert(CM != 2 || PAT == 7, "code tables hold 7-point rows");
    static_assert(CM != 3 || PAT == 27, "mask rows are 27-point rows");
    __shared__ TV pl[3 * kPLS];
    __shared__ TC crows[CLS ? kBlock : 1];  // 9 x 27 class rows, padded to one slot per thread
    __shared__ TC ctab[CM == 2 ? kCodes * PAT : 1];
    if constexpr(CM == 2) {
        // the code table (identical in every segment: rewriting it is benign);
        // the first barrier of the plane loop orders it before any use
        for(int i = threadIdx.x; i < kCodes * PAT; i += kBlock) ctab[i] = op.ctab[i];
    }
    const int x0 = (t % g.ntx) * kTX, y0 = (t / g.ntx) * kTY;
    const int tid = threadIdx.x;
    const int64_t sz = (int64_t)op.nx * op.ny;
    int goff[kNLD];
    bool gok[kNLD];
#pragma unroll
    for(int j = 0; j < kNLD; j++) {
        const int idx = tid + j * kBlock;
        const int lx = idx % kW, ly = idx / kW;
        const int gx = x0 + lx - 1, gy = y0 + ly - 1;
        gok[j] = idx < kPL && gx >= 0 && gx < op.nx && gy >= 0 && gy < op.ny;
        goff[j] = gok[j] ? gx + op.nx * gy : 0;
    }
    // two register sets keep the loads of planes z+2 and z+3 in flight while
    // plane z is computed.  They hold the raw loaded words: the direction
    // update p = r + beta p and the zero padding are applied when the plane is
    // put into LDS, so no arithmetic on a loaded value (which would make the
    // wave wait for the load) sits between a fetch and its put.
    struct Fetched {
        TV a[kNLD], b[FUSE ? kNLD : 1];
        float c[Epi::kCorr ? kNLD : 1];
        bool zin;
    };
    // correcting residual: x + d is staged (and written to p_out) unless the
    // outer solve had already converged, when d is stale and x is copied
    bool apply = false;
    if constexpr(Epi::kCorr) apply = !epi.so->done;
    Fetched ra, rb;
    // Every load of the plane loop is issued unconditionally (a plane that is
    // not needed reads element 0 of the tile's first row), so the number of
    // loads in flight is the same on every path and the compiler's wait counts
    // can leave the next planes' loads outstanding.
    // (the scheduling barriers keep each plane's loads together and in
    // program order, which the wait counts rely on)
    auto fetch = [&](Fetched &f, int z, bool need) {
        f.zin = need && z >= 0 && z < op.nz;
        const int64_t zo = sz * (f.zin ? z : 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for(int j = 0; j < kNLD; j++) {
            const int64_t c = need ? zo + goff[j] : 0;  // always a valid address
            f.a[j] = x[c];
            if constexpr(FUSE) f.b[j] = rz[c];
            if constexpr(Epi::kCorr) f.c[j] = epi.d[c];
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto put = [&](const Fetched &f, int slot) {
#pragma unroll
        for(int j = 0; j < kNLD; j++) {
            TV val;
            if constexpr(FUSE) val = f.b[j] + beta * f.a[j];
            else if constexpr(Epi::kCorr) val = f.a[j] + (apply ? (double)f.c[j] : 0.0);
            else val = f.a[j];
            pl[slot * kPLS + tid + j * kBlock] = (f.zin && gok[j]) ? val : TV(0);  // lanes past kPL fill padding
        }
    };
    // class mode: the 9 class rows of a plane (27 slots each) are staged in LDS,
    // one plane ahead, for tiles on an x or y face and on the boundary planes
    const bool face_xy = CLS && (x0 < g.lo[0] || x0 + kTX - 1 > g.hi[0] || y0 < g.lo[1] || y0 + kTY - 1 > g.hi[1]);
    auto need_rows = [&](int z) { return CLS && (face_xy || z < g.lo[2] || z > g.hi[2]); };
    TC rrow = TC(0);
    auto fetch_rows = [&](int z, bool need) {
        if constexpr(CLS) rrow = op.table[(need && tid < 9 * 27 && need_rows(z)) ? (int64_t)z * 9 * 27 + tid : 0];
    };
    const int tx = tid % kTX, ty = tid / kTX;
    const int gx = x0 + tx;
    const bool gx_in = gx < op.nx;
    // class mode: lanes on an x face are evaluated by the face pass below
    const bool xf = !CLS || (gx >= g.lo[0] && gx <= g.hi[0]);
    // face pass: lanes 0..2 kTY-1 of wave 0 take the tile's x-face cells (row
    // l % kTY, left face for l < kTY when the tile holds x = 0, right face otherwise)
    const int fx = tid < kTY ? 0 : op.nx - 1;
    const bool face_lane =
        CLS && tid < 2 * kTY && fx >= x0 && fx < x0 + kTX && (fx < g.lo[0] || fx > g.hi[0]) && y0 + (tid % kTY) < op.ny;
    auto ycls = [&](int gy) { return gy == 0 ? 0 : (gy == op.ny - 1 ? 2 : 1); };
    // epilogue inputs of plane z (e.g. b), loaded a plane ahead
    using Pre = typename Epi::Pre;
    Pre pa[kRPT], pb[kRPT];
    auto cell_of = [&](int q, int z) -> int64_t {
        const int gy = y0 + ty + q * (kBlock / kTX);
        return (gx_in && gy < op.ny) ? gx + (int64_t)op.nx * gy + sz * z : 0;
    };
    auto preload = [&](Pre(&pre)[kRPT], int z, bool need) {
#pragma unroll
        for(int q = 0; q < kRPT; q++) pre[q] = epi.load(need ? cell_of(q, z) : 0);
    };
    auto face_pre = [&](int z, bool need) -> Pre {
        return epi.load((need && face_lane) ? fx + (int64_t)op.nx * (y0 + (tid % kTY)) + sz * z : 0);
    };
    // per-cell rows: the rows of plane z are loaded at the end of the previous
    // plane's compute, so their loads are in flight across the LDS staging and
    // both barriers instead of being waited for row by row inside compute
    // (read once per sweep: non-temporal loads).
    // Slot arrays start at uniform addresses; a 32-bit cell offset lets every
    // load use a scalar base (callers keep n <= 2^31).
    constexpr bool PF = CM == 0 && MONO_PERCELL_PREFETCH;
    TC crg[PF ? kRPT : 1][PF ? PAT : 1];
    // mask rows: the masks of plane z, loaded likewise a plane ahead
    constexpr bool MM = CM == 3;
    uint32_t mk[MM ? kRPT : 1];
    auto fetch_coef = [&](int z, bool need) {
        if constexpr(MM) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for(int q = 0; q < kRPT; q++) mk[q] = __builtin_nontemporal_load(op.nbm + (uint32_t)(need ? cell_of(q, z) : 0));
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr(PF) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for(int q = 0; q < kRPT; q++) {
                const uint32_t cu = (uint32_t)(need ? cell_of(q, z) : 0);
#pragma unroll
                for(int k = 0; k < PAT; k++) crg[q][k] = __builtin_nontemporal_load(op.percell + (int64_t)k * op.n + cu);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // plane z from LDS (planes z-1, z, z+1 and, when needed, its class rows staged)
    auto compute = [&](int z, const Pre(&pre)[kRPT], Pre fpre) {
        const TV *const P3[3] = {pl + ((z + 2) % 3) * kPLS, pl + (z % 3) * kPLS, pl + ((z + 1) % 3) * kPLS};
        TC crow[PAT];
        // mask rows: the tensors of layers z-1, z, z+1 (clamped at the box
        // faces, where no neighbour uses them), wave-uniform
        TC tz[MM ? 3 : 1][6];
        if constexpr(MM) {
            const scalar_ptr<TC> lt = as_scalar(op.lt);
#pragma unroll
            for(int l = 0; l < 3; l++) {
                const int zl = z - 1 + l < 0 ? 0 : (z - 1 + l >= op.nz ? op.nz - 1 : z - 1 + l);
#pragma unroll
                for(int k = 0; k < 6; k++) tz[l][k] = lt[zl * 6 + k];
            }
        }
        bool zf = true;
        if constexpr(CLS) {
            zf = z >= g.lo[2] && z <= g.hi[2];
            if(zf) {
                const scalar_ptr<TC> row = as_scalar(op.table_int + z * PAT);
#pragma unroll
                for(int k = 0; k < PAT; k++) crow[k] = row[k];
            }
        }
        // per-cell rows without prefetch: one row at a time (its PAT
        // coefficient loads are in flight)
        constexpr int kUq = CLS || PF || MM ? kRPT : 1;
#pragma unroll kUq
        for(int q = 0; q < kRPT; q++) {
            const int rr = ty + q * (kBlock / kTX);
            const int gy = y0 + rr;
            if(!gx_in || gy >= op.ny) continue;
            const int64_t c = gx + (int64_t)op.nx * gy + sz * z;
            const int base = (rr + 1) * kW + tx + 1;
            TV acc;
            double dg = 0.0;  // the row's diagonal (Jacobi epilogues only)
            if constexpr(MM) {
                TC row[27];
                mask_aniso_row<TC, true>(mk[q], tz, op.geo[0], op.geo[1], op.geo[2], op.geo[3], row);
                acc = TV(0);
#pragma unroll
                for(int k = 0; k < PAT; k++)
                    acc += (TV)row[slot_of(RowPattern<PAT>::o[k][0], RowPattern<PAT>::o[k][1], RowPattern<PAT>::o[k][2])] *
                           P3[RowPattern<PAT>::o[k][2] + 1][base + RowPattern<PAT>::o[k][1] * kW + RowPattern<PAT>::o[k][0]];
                if constexpr(Epi::kDiag) dg = (double)row[13];
            } else if constexpr(PF) {
                acc = TV(0);
#pragma unroll
                for(int k = 0; k < PAT; k++)
                    acc += (TV)crg[q][k] *
                           P3[RowPattern<PAT>::o[k][2] + 1][base + RowPattern<PAT>::o[k][1] * kW + RowPattern<PAT>::o[k][0]];
                if constexpr(Epi::kDiag) dg = (double)crg[q][0];
            } else if constexpr(CM == 0) {
                const uint32_t cu = (uint32_t)c;
                acc = TV(0);
#pragma unroll
                for(int k = 0; k < PAT; k++)
                    acc += (TV)(op.percell + (int64_t)k * op.n)[cu] *
                           P3[RowPattern<PAT>::o[k][2] + 1][base + RowPattern<PAT>::o[k][1] * kW + RowPattern<PAT>::o[k][0]];
                if constexpr(Epi::kDiag) dg = (double)op.percell[cu];
            } else if constexpr(CM == 2) {
                const TC *row = ctab + (int)op.codes[c] * PAT;
                acc = TV(0);
#pragma unroll
                for(int k = 0; k < PAT; k++)
                    acc += (TV)row[k] * P3[RowPattern<PAT>::o[k][2] + 1][base + RowPattern<PAT>::o[k][1] * kW + RowPattern<PAT>::o[k][0]];
                if constexpr(Epi::kDiag) dg = (double)row[0];
            } else {
                if(!xf) continue;
                if(zf && gy >= g.lo[1] && gy <= g.hi[1]) {
                    acc = row_lds<PAT>(crow, P3, base);
                    if constexpr(Epi::kDiag) dg = (double)crow[0];
                } else {
                    // y face or boundary plane: the row of class (y class, interior x),
                    // the same for the whole wave
                    const TC *row = crows + (ycls(gy) * 3 + (op.nx == 1 ? 0 : 1)) * 27;
                    acc = row27_lds(row, P3, base);
                    if constexpr(Epi::kDiag) dg = (double)row[13];
                }
            }
            const TV xc = P3[1][base];
            if constexpr(FUSE || Epi::kCorr) p_out[c] = xc;
            epi(c, acc, xc, dg, pre[q], v);
        }
        if constexpr(CLS) {
            if(face_lane) {
                const int rr = tid % kTY, gy = y0 + rr;
                const int64_t c = fx + (int64_t)op.nx * gy + sz * z;
                const int base = (rr + 1) * kW + (fx - x0) + 1;
                const TC *row = crows + (ycls(gy) * 3 + (fx == 0 ? 0 : 2)) * 27;
                const TV xc = P3[1][base];
                if constexpr(FUSE || Epi::kCorr) p_out[c] = xc;
                epi(c, row27_lds(row, P3, base), xc, (double)row[13], fpre, v);
            }
        }
    };
    fetch(ra, z0 - 1, true);
    put(ra, (z0 + 2) % 3);
    fetch(ra, z0, true);
    put(ra, z0 % 3);
    fetch_rows(z0, true);
    fetch(ra, z0 + 1, true);
    fetch(rb, z0 + 2, z0 + 2 <= z1);
    preload(pa, z0, true);
    Pre fa = face_pre(z0, true);
    preload(pb, z0 + 1, z0 + 1 < z1);
    Pre fb = face_pre(z0 + 1, z0 + 1 < z1);
    fetch_coef(z0, true);
    // Each half: the plane and class rows fetched earlier go into LDS (the
    // class rows unconditionally, so that their load is waited for on every
    // path), then the class rows and the plane three ahead are fetched, in
    // that order.
    // The loop body holds both halves with no exit between them (an odd last
    // plane is handled after the loop), so every path into the loop head has
    // issued the same loads in the same order and the compiler's wait counts
    // there leave the other register set's loads in flight.
    auto half_a = [&](int z) {
        put(ra, (z + 1) % 3);  // plane z+1
        if constexpr(CLS) crows[tid] = rrow;
        fetch_rows(z + 1, z + 1 < z1);
        fetch(ra, z + 3, z + 3 <= z1);
        __syncthreads();
        compute(z, pa, fa);
        preload(pa, z + 2, z + 2 < z1);
        fa = face_pre(z + 2, z + 2 < z1);
        fetch_coef(z + 1, z + 1 < z1);
        __syncthreads();
    };
    int z = z0;
    for(; z + 1 < z1; z += 2) {
        half_a(z);
        put(rb, (z + 2) % 3);  // plane z+2
        if constexpr(CLS) crows[tid] = rrow;
        fetch_rows(z + 2, z + 2 < z1);
        fetch(rb, z + 4, z + 4 <= z1);
        __syncthreads();
        compute(z + 1, pb, fb);
        preload(pb, z + 3, z + 3 < z1);
        fb = face_pre(z + 3, z + 3 < z1);
        fetch_coef(z + 2, z + 2 < z1);
        __syncthreads();
    }
    if(z < z1) half_a(z);
}

// Call f(tile, z0, z1) for every tile and plane range of run r.
template <class F> __device__ __forceinline__ void for_run(const TileGeom &g, int nz, int r, F &&f) {
    if(g.L > 0) {  // slab mode: run r is tile r % ntiles over slab r / ntiles
        const int slab = r / g.ntiles, t = r % g.ntiles;
        const int z0 = slab * g.L, z1 = z0 + g.L < nz ? z0 + g.L : nz;
        f(t, z0, z1);
        return;
    }
    const int64_t nitems = (int64_t)g.ntiles * nz;
    int64_t w = (int64_t)r * g.per;
    const int64_t end = w + g.per < nitems ? w + g.per : nitems;
    while(w < end) {
        const int t = (int)(w / nz), z0 = (int)(w % nz);
        const int z1 = (int64_t)nz - z0 < end - w ? nz : z0 + (int)(end - w);
        f(t, z0, z1);
        w += z1 - z0;
    }
}

// This block's run: its partial sums go to the run's own slot of Pout.
template <typename TV, typename TC, int CM, int PAT, bool FUSE, class Epi>
__device__ __forceinline__ void tile_sweep(const OpView<TC> &op, const TileGeom &g, const TV *__restrict__ x, const TV *__restrict__ rz,
                                           TV beta, TV *__restrict__ p_out, const Epi &epi, double *Pout) {
    const int r = xcd_chunk(blockIdx.x, gridDim.x);
    if(r < g.nruns) {
        double v[Epi::NR];
#pragma unroll
        for(int j = 0; j < Epi::NR; j++) v[j] = 0.0;
        for_run(g, op.nz, r, [&](int t, int z0, int z1) {
            tile_segment<TV, TC, CM, PAT, FUSE, Epi>(op, g, t, z0, z1, x, rz, beta, p_out, epi, v);
        });
        write_partials_at<Epi::NR>(v, Pout, r);
    }
}

// x += alpha p over the cells of this block's (tile, plane) items: the
// deferred correction when a solve ends before the sweep that would apply it.
template <typename TV, typename TC>
__device__ __forceinline__ void defer_apply(const OpView<TC> &op, const TileGeom &g, TV *__restrict__ x, const TV *__restrict__ p,
                                            TV alpha) {
    const int r = xcd_chunk(blockIdx.x, gridDim.x);
    const int64_t sz = (int64_t)op.nx * op.ny;
    const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
    if(r < g.nruns)
        for_run(g, op.nz, r, [&](int t, int z0, int z1) {
            const int gx = (t % g.ntx) * kTX + tx, y0 = (t / g.ntx) * kTY;
            if(gx >= op.nx) return;
            for(int z = z0; z < z1; z++) {
#pragma unroll
                for(int q = 0; q < kRPT; q++) {
                    const int gy = y0 + ty + q * (kBlock / kTX);
                    if(gy >= op.ny) continue;
                    const int64_t c = gx + (int64_t)op.nx * gy + sz * z;
                    x[c] = x[c] + alpha * p[c];
                }
            }
        });
}

// One-plane domains (nz == 1) with 7-point rows (5 in the plane): one wave per
// (64-column strip, run of kRows2d rows).  A lane owns one column; all loads of
// the run (its rows and the one above and below, the strip's two edge columns,
// the row codes and the epilogue inputs) are issued first, in straight-line
// code, then the rows are evaluated in order, x neighbours by lane shifts.  No
// LDS staging and no barriers.  The in-plane terms are summed in the order of
// RowPattern<7> as the plane kernel does (its z terms are products with the
// zero padding), so both kernels give the same rows bit for bit.
#ifndef MONO_ROWS_2D
#define MONO_ROWS_2D 8
#endif
constexpr int kRows2d = MONO_ROWS_2D;

template <typename TV> __device__ __forceinline__ TV lane_shift(TV v, int src) { return __shfl(v, src); }

template <typename TV, typename TC, int CM, bool FUSE, class Epi>
__device__ __forceinline__ void strip_run(const OpView<TC> &op, int strip, int chunk, const TV *__restrict__ x,
                                          const TV *__restrict__ rz, TV beta, TV *__restrict__ p_out, const Epi &epi,
                                          const TC *ctab, bool apply, double (&v)[Epi::NR]) {
    constexpr int R = kRows2d, W = R + 2;
    const int lane = threadIdx.x & 63;
    const int x0 = strip * 64, gx = x0 + lane;
    const bool xin = gx < op.nx;
    const int ex = lane < 32 ? x0 - 1 : x0 + 64;  // the edge column this lane loads
    const bool exin = ex >= 0 && ex < op.nx;
    const int y0 = chunk * R;
    // raw loads of rows y0-1 .. y0+R (window slot w holds row y0-1+w)
    TV ca[W], ea[W], cb[FUSE ? W : 1], eb[FUSE ? W : 1];
    float cd[Epi::kCorr ? W : 1], ed[Epi::kCorr ? W : 1];
    bool yin[W];
#pragma unroll
    for(int w = 0; w < W; w++) {
        const int y = y0 - 1 + w;
        yin[w] = y >= 0 && y < op.ny;
        const int64_t row = (int64_t)(yin[w] ? y : 0) * op.nx;
        const int64_t c = (yin[w] && xin) ? row + gx : 0, ce = (yin[w] && exin) ? row + ex : 0;
        ca[w] = x[c];
        ea[w] = x[ce];
        if constexpr(FUSE) {
            cb[w] = rz[c];
            eb[w] = rz[ce];
        }
        if constexpr(Epi::kCorr) {
            cd[w] = epi.d[c];
            ed[w] = epi.d[ce];
        }
    }
    using Pre = typename Epi::Pre;
    Pre pre[R];
    uint8_t code[CM == 2 ? R : 1];
#pragma unroll
    for(int r = 0; r < R; r++) {
        const int y = y0 + r;
        const int64_t c = (xin && y < op.ny) ? gx + (int64_t)op.nx * y : 0;
        pre[r] = epi.load(c);
        if constexpr(CM == 2) code[r] = op.codes[c];
    }
    auto val = [&](TV a, TV b, float d, bool ok) -> TV {
        TV t;
        if constexpr(FUSE) t = b + beta * a;
        else if constexpr(Epi::kCorr) t = a + (apply ? (double)d : 0.0);
        else t = a;
        return ok ? t : TV(0);
    };
    // combined values of the window (centre column and edge column)
    TV cv[W], ev[W];
#pragma unroll
    for(int w = 0; w < W; w++) {
        cv[w] = val(ca[w], FUSE ? cb[FUSE ? w : 0] : TV(0), Epi::kCorr ? cd[Epi::kCorr ? w : 0] : 0.0f, yin[w] && xin);
        ev[w] = val(ea[w], FUSE ? eb[FUSE ? w : 0] : TV(0), Epi::kCorr ? ed[Epi::kCorr ? w : 0] : 0.0f, yin[w] && exin);
    }
#pragma unroll
    for(int r = 0; r < R; r++) {
        const int y = y0 + r, w = r + 1;
        const TV left = lane_shift(cv[w], lane - 1), right = lane_shift(cv[w], lane + 1);
        // lane 0's left neighbour and lane 63's right neighbour are the edge
        // columns, loaded by lanes 0..31 and 32..63 respectively
        const TV xl = lane == 0 ? ev[w] : left;
        const TV xr = lane == 63 ? ev[w] : right;
        const TV nb[5] = {cv[w], xr, xl, cv[w + 1], cv[w - 1]};  // RowPattern<7> slots 0..4
        if(!xin || y >= op.ny) continue;
        const int64_t c = gx + (int64_t)op.nx * y;
        TV acc = TV(0);
        double dg = 0.0;
        if constexpr(CM == 0) {
            const uint32_t cu = (uint32_t)c;
#pragma unroll
            for(int k = 0; k < 5; k++) acc += (TV)(op.percell + (int64_t)k * op.n)[cu] * nb[k];
            if constexpr(Epi::kDiag) dg = (double)op.percell[cu];
        } else {
            const TC *rowc = ctab + (int)code[r] * 7;
#pragma unroll
            for(int k = 0; k < 5; k++) acc += (TV)rowc[k] * nb[k];
            if constexpr(Epi::kDiag) dg = (double)rowc[0];
        }
        if constexpr(FUSE || Epi::kCorr) p_out[c] = cv[w];
        epi(c, acc, cv[w], dg, pre[r], v);
    }
}

template <typename TV, typename TC, int CM, bool FUSE, class Epi>
__global__ __launch_bounds__(kBlock) void k_stencil2d(OpView<TC> op, int nitems, const TV *__restrict__ x, const TV *__restrict__ rz,
                                                      TV *__restrict__ p_out, Epi epi, CgSetup cs, int mode, int k,
                                                      const double *Pin, int nprev, double *Pout, CgScalars *sc, CgScalars *so) {
    double betad = 0.0;
    bool now;
    if(mode >= 0 && !iter_prologue(mode, k, Pin, nprev, sc, so, cs, betad, now)) return;
    const TV beta = (TV)betad;
    __shared__ TC ctab[CM == 2 ? kCodes * 7 : 1];
    if constexpr(CM == 2) {
        for(int i = threadIdx.x; i < kCodes * 7; i += kBlock) ctab[i] = op.ctab[i];
        __syncthreads();
    }
    bool apply = false;
    if constexpr(Epi::kCorr) apply = !epi.so->done;
    double v[Epi::NR];
#pragma unroll
    for(int j = 0; j < Epi::NR; j++) v[j] = 0.0;
    const int nstrips = (op.nx + 63) / 64;
    // items (chunk, strip), strip fastest; each XCD takes a contiguous range
    const int waves = gridDim.x * (kBlock / 64);
    for(int item = xcd_chunk(blockIdx.x, gridDim.x) * (kBlock / 64) + (int)(threadIdx.x >> 6); item < nitems; item += waves)
        strip_run<TV, TC, CM, FUSE, Epi>(op, item % nstrips, item / nstrips, x, rz, beta, p_out, epi, ctab, apply, v);
    write_partials<Epi::NR>(v, Pout);
}

// Residual kernels: no prologue.  Iteration kernel (q = A p): prologue first.
// Waves per SIMD asked of the class-mode FP32 sweeps: at 6 the compiler keeps
// them at 73 VGPRs without spills (unbounded they take 81, 5 waves), and the
// occupancy query gives the work split 6 blocks per CU.  Measured on the 16M
// box: FP32+refine step 2.36 ms at 6, 2.40 at 7, 2.42 unbounded; the FP64
// class sweep bounded to 4 waves (124 VGPRs) was slower.  Other storage
// modes spill under such bounds and keep the compiler's choice (1: none).
#ifndef MONO_SW32
#define MONO_SW32 6
#endif
template <typename TV, int CM> struct SweepWaves {
    static constexpr int value = (CM == 1 && sizeof(TV) == 4) ? MONO_SW32 : 1;
};
template <typename TV, typename TC, int CM, int PAT, bool FUSE, class Epi>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SweepWaves<TV, CM>::value))) void k_stencil(OpView<TC> op, TileGeom g, const TV *__restrict__ x, const TV *__restrict__ rz,
                                                    TV *__restrict__ p_out, Epi epi, CgSetup cs, int mode, int k, const double *Pin,
                                                    int nprev, double *Pout, CgScalars *sc, CgScalars *so) {
    double betad = 0.0;
    bool now = false;
    Epi e = epi;
    if constexpr(Epi::kDefer) {
        // x += alpha_{k-1} p_{k-1} of the previous iteration (p_{k-1} is this
        // kernel's input direction), folded into this sweep
        e.on = mode == 1;
        e.alpha = mode == 1 ? (TV)sc->alpha : TV(0);
    }
    if(mode >= 0 && !iter_prologue(mode, k, Pin, nprev, sc, so, cs, betad, now)) {
        // the kernel that ends the solve still owes x its last correction
        if constexpr(Epi::kDefer)
            if(now && e.on) defer_apply<TV, TC>(op, g, e.x, x, e.alpha);
        return;
    }
    const TV beta = (TV)betad;
    tile_sweep<TV, TC, CM, PAT, FUSE, Epi>(op, g, x, rz, beta, p_out, e, Pout);
}

// q = A p, partial p.q
// Epilogue interface: NR partial sums; kDiag: needs the row's diagonal;
// Pre load(c): the epilogue's inputs of cell c, loaded a plane ahead.
struct NoPre {};

template <typename TV> struct EpiSpmv {
    static constexpr int NR = 1;
    static constexpr bool kDiag = false, kCorr = false, kDefer = false;
    static constexpr const float *d = nullptr;
    static constexpr const CgScalars *so = nullptr;
    using Pre = NoPre;
    TV *q;
    __device__ Pre load(int64_t) const { return {}; }
    __device__ void operator()(int64_t c, TV ax, TV pc, double, Pre, double (&v)[NR]) const {
        q[c] = ax;
        v[0] += (double)pc * (double)ax;
    }
};

// q = A p, partial p.q, and the previous iteration's x += alpha p_in (its
// inputs loaded a plane ahead), so the update kernel need not touch x or p.
// x is touched once per iteration: non-temporal, so it does not displace the
// vectors the next kernels reuse from the caches.
template <typename TV> struct EpiSpmvX {
    static constexpr int NR = 1;
    static constexpr bool kDiag = false, kCorr = false, kDefer = true;
    static constexpr const float *d = nullptr;
    static constexpr const CgScalars *so = nullptr;
    struct Pre {
        TV x, p;
    };
    TV *q;
    TV *x;
    const TV *p_in;
    bool on = false;   // set in the kernel: iterations k >= 1
    TV alpha = TV(0);  // set in the kernel from the previous update
    __device__ Pre load(int64_t c) const { return on ? Pre{__builtin_nontemporal_load(&x[c]), p_in[c]} : Pre{TV(0), TV(0)}; }
    __device__ void operator()(int64_t c, TV ax, TV pc, double, Pre pre, double (&v)[NR]) const {
        q[c] = ax;
        if(on) __builtin_nontemporal_store(pre.x + alpha * pre.p, &x[c]);
        v[0] += (double)pc * (double)ax;
    }
};

// r = b - A x, p = r (or z = D^-1 r, p = z); partial r.r, r.z, b.b
template <bool JACOBI> struct EpiRes64 {
    static constexpr int NR = 3;
    static constexpr bool kDiag = JACOBI, kCorr = false, kDefer = false;
    static constexpr const float *d = nullptr;
    static constexpr const CgScalars *so = nullptr;
    using Pre = double;
    const double *b;
    double *r, *p, *z;
    __device__ Pre load(int64_t c) const { return b[c]; }
    __device__ void operator()(int64_t c, double ax, double, double dg, Pre bc, double (&v)[NR]) const {
        const double rc = bc - ax;
        r[c] = rc;
        if constexpr(JACOBI) {
            const double d = dg == 0.0 ? 1.0 : dg;
            const double zc = (1.0 / d) * rc;
            z[c] = zc;
            p[c] = zc;
            v[1] += rc * zc;
        } else {
            p[c] = rc;
        }
        v[0] += rc * rc;
        v[2] += bc * bc;
    }
};

// mixed-precision outer residual r = b - A x in FP64, handed to the FP32 solve
// (whose first direction is r itself and whose correction starts from zero,
// so neither is written here).  CORR: x is first corrected by the previous
// sweep's FP32 correction d, x + d going to the kernel's p_out (a second FP64
// buffer: other blocks still read x as their halo).
template <bool CORR> struct EpiResOuterT {
    static constexpr int NR = 2;  // r.r, b.b
    static constexpr bool kDiag = false, kCorr = CORR, kDefer = false;
    using Pre = double;
    const double *b;
    float *r32;
    const float *d;        // CORR: the correction
    const CgScalars *so;   // CORR: outer solve state (done: d is stale)
    __device__ Pre load(int64_t c) const { return b[c]; }
    __device__ void operator()(int64_t c, double ax, double, double, Pre bc, double (&v)[NR]) const {
        const double rc = bc - ax;
        r32[c] = (float)rc;
        v[0] += rc * rc;
        v[1] += bc * bc;
    }
};
using EpiResOuter = EpiResOuterT<false>;
using EpiResOuterCorr = EpiResOuterT<true>;

// Elementwise kernels keep four cells in flight per thread (all loads issued
// before any use).
constexpr int kIlp = 4;

// x += alpha p, r -= alpha q (z = D^-1 r); partial r.r, r.z.  Prologue: p.q.
// x (the mixed solve's correction d) is read and written once per iteration
// with the non-temporal hint, leaving the caches to p, q and r (16M-cell
// FP32+refine step 2.53 -> 2.45 ms).
template <typename TV, typename TX, typename TC>
__global__ __launch_bounds__(kBlock) void k_update(OpView<TC> op, TX *__restrict__ x, const TV *__restrict__ p, const TV *__restrict__ q,
                                                   TV *__restrict__ r, TV *__restrict__ z, int jacobi, int k, const double *Pin,
                                                   int nprev, double *Pout, CgScalars *sc, int fresh, int defer) {
    if(sc->done) return;
    double pq[1];
    sum_partials<1>(Pin, nprev, pq);
    const double alpha = sc->rrs[k & 1] / pq[0];
    if(blockIdx.x == 0 && threadIdx.x == 0) {
        sc->pAp = pq[0];
        sc->alpha = alpha;
    }
    const TV a = (TV)alpha;
    // fresh: x starts from zero (the mixed solve's correction), so the first
    // iteration writes a p without reading x
    const bool x0 = fresh && k == 0;
    // defer: x += alpha p is left to the next iteration's stencil sweep
    if(defer) {
        int64_t beg, end;
        chunk_range(op.n, beg, end);
        double v[2] = {0.0, 0.0};
        for(int64_t c0 = beg + threadIdx.x; c0 < end; c0 += kIlp * kBlock) {
            int64_t cc[kIlp];
            TV qv[kIlp], rv[kIlp];
#pragma unroll
            for(int u = 0; u < kIlp; u++) {
                const int64_t c = c0 + u * kBlock;
                cc[u] = c < end ? c : end - 1;
                qv[u] = q[cc[u]];
                rv[u] = r[cc[u]];
            }
#pragma unroll
            for(int u = 0; u < kIlp; u++) {
                const int64_t c = c0 + u * kBlock;
                if(c >= end) continue;
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
        write_partials<2>(v, Pout);
        return;
    }
    int64_t beg, end;
    chunk_range(op.n, beg, end);
    double v[2] = {0.0, 0.0};
    for(int64_t c0 = beg + threadIdx.x; c0 < end; c0 += kIlp * kBlock) {
        int64_t cc[kIlp];
        TX xv[kIlp];
        TV pv[kIlp], qv[kIlp], rv[kIlp];
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            cc[u] = c < end ? c : end - 1;
            xv[u] = x0 ? TX(0) : __builtin_nontemporal_load(&x[cc[u]]);
            pv[u] = p[cc[u]];
            qv[u] = q[cc[u]];
            rv[u] = r[cc[u]];
        }
#pragma unroll
        for(int u = 0; u < kIlp; u++) {
            const int64_t c = c0 + u * kBlock;
            if(c >= end) continue;
            __builtin_nontemporal_store(xv[u] + (TX)(a * pv[u]), &x[c]);
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
    write_partials<2>(v, Pout);
}

// inner iterations run by the sweep that just ended, or -1 when the outer
// solve had already converged (its inner and correction kernels were empty)
__global__ void k_note_sweep(const CgScalars *si, const CgScalars *so, int *out) {
    if(threadIdx.x == 0) out[0] = so->done ? -1 : si->iters - 1;
}

// ---------------------------------------------------------------------------
// CSR (non-uniform meshes), FP64: one wave per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_csr_spmv(int n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                     const double *__restrict__ val, const double *__restrict__ x,
                                                     double *__restrict__ y, const CgScalars *sc) {
    if(sc && sc->done) return;
    const int lane = threadIdx.x & 63;
    const int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if(row >= n) return;
    double acc = 0.0;
    for(int64_t kk = rp[row] + lane; kk < rp[row + 1]; kk += 64) acc += val[kk] * x[col[kk]];
    acc = wave_sum(acc);
    if(lane == 0) y[row] = acc;
}

__global__ __launch_bounds__(kBlock) void k_csr_res(int n, const double *__restrict__ Ax, const double *__restrict__ b,
                                                    const double *__restrict__ diag, double *__restrict__ r, double *__restrict__ p,
                                                    double *__restrict__ z, int jacobi, double *Pout) {
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
    write_partials<3>(v, Pout);
}

// p = rz + beta p after the stopping test (prologue over the previous partials)
__global__ __launch_bounds__(kBlock) void k_csr_dir(int n, double *__restrict__ p, const double *__restrict__ rz, CgSetup cs, int mode,
                                                    int k, const double *Pin, int nprev, CgScalars *sc) {
    double beta = 0.0;
    bool now;
    if(!iter_prologue(mode, k, Pin, nprev, sc, nullptr, cs, beta, now)) return;
    int64_t beg, end;
    chunk_range(n, beg, end);
    for(int64_t c = beg + threadIdx.x; c < end; c += kBlock) p[c] = rz[c] + beta * p[c];
}

__global__ __launch_bounds__(kBlock) void k_csr_dot(int n, const double *__restrict__ p, const double *__restrict__ q, double *Pout,
                                                    const CgScalars *sc) {
    if(sc->done) return;
    int64_t beg, end;
    chunk_range(n, beg, end);
    double v[1] = {0.0};
    for(int64_t c = beg + threadIdx.x; c < end; c += kBlock) v[0] += p[c] * q[c];
    write_partials<1>(v, Pout);
}

__global__ __launch_bounds__(kBlock) void k_csr_update(int n, double *__restrict__ x, const double *__restrict__ p,
                                                       const double *__restrict__ q, double *__restrict__ r, double *__restrict__ z,
                                                       const double *__restrict__ diag, int jacobi, int k, const double *Pin, int nprev,
                                                       double *Pout, CgScalars *sc) {
    if(sc->done) return;
    double pq[1];
    sum_partials<1>(Pin, nprev, pq);
    const double a = sc->rrs[k & 1] / pq[0];
    if(blockIdx.x == 0 && threadIdx.x == 0) {
        sc->pAp = pq[0];
        sc->alpha = a;
    }
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
    write_partials<2>(v, Pout);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <int A> using IC = std::integral_constant<int, A>;

template <class F> static void with_pat(int P, F &&f) {
    if(P == 7) f(IC<7>{});
    else if(P == 19) f(IC<19>{});
    else f(IC<27>{});
}

// Call f(IC<S>, IC<PAT>) for the operator's class-row slot count and row pattern.
// Call f(IC<CM>, IC<PAT>) for the operator's storage mode (0 per-cell rows,
// 1 class table, 2 code table, 3 mask rows) and row pattern.
template <typename TC, class F> static void with_op(const OpView<TC> &op, F &&f) {
    with_pat(op.pat, [&](auto P_) {
        if(op.table) f(IC<1>{}, P_);
        else if constexpr(decltype(P_)::value == 7) {
            if(op.codes) f(IC<2>{}, P_);
            else f(IC<0>{}, P_);
        } else if constexpr(decltype(P_)::value == 27) {
            if(op.nbm) f(IC<3>{}, P_);
            else f(IC<0>{}, P_);
        } else f(IC<0>{}, P_);
    });
}

// Runs of the plane-loop kernels (see TileGeom): one per resident block (at
// most kMaxBlocks, the partial-sum slots; at least kMinRun items each), in
// whole rounds of the XCD dispatch order.
constexpr int kMinRun = 8;
constexpr int kFill = 2;  // planes of pipeline fill at the start of a segment
template <typename TC> static TileGeom make_geom(const OpView<TC> &op, int slots) {
    TileGeom g{};
    const int L[3] = {op.nx, op.ny, op.nz};
    for(int a = 0; a < 3; a++) {
        g.lo[a] = L[a] == 1 ? 0 : 1;
        g.hi[a] = L[a] == 1 ? 0 : L[a] - 2;
    }
    g.ntx = (op.nx + kTX - 1) / kTX;
    g.nty = (op.ny + kTY - 1) / kTY;
    g.ntiles = g.ntx * g.nty;
    // item ranges: one contiguous run of (tile, plane) items per block
    const int64_t nitems = (int64_t)g.ntiles * op.nz;
    int64_t nb = (nitems + kMinRun - 1) / kMinRun;
    if(nb > slots) nb = slots;
    if(nb > kMaxBlocks) nb = kMaxBlocks;
    if(nb >= 16) nb = nb / 8 * 8;
    if(nb < 1) nb = 1;
    const int per = (int)((nitems + nb - 1) / nb);
#if MONO_SLAB_RUNS
    // slab mode: every tile of the plane in each of nslab z slabs, one run per
    // block, all runs resident at once; the runs of a slab are consecutive,
    // so an XCD's blocks march neighbouring tiles through the same planes
    // together and their shared halo lines are fetched from HBM once.  Its
    // runs are whole slabs, so when the resident slots are not close to a
    // multiple of the tile count its longest run exceeds the item-range one:
    // the split with the shorter longest run wins, a run of item ranges
    // counted with two pipeline fills (it may span two tiles) and a quarter
    // more for the halo lines its unaligned neighbours fetch again (measured:
    // 16M-cell box, FP32 class sweep 74.5 us as 1050 slab runs of 34 planes
    // against 80.5 us as item runs of 28; FP64 class residual 163 us as 700
    // slab runs of 50 planes on 1024 slots against 130 us as item runs of 35).
    if(g.ntiles <= slots) {
        int64_t nslab = slots / g.ntiles;
        if(nslab > op.nz) nslab = op.nz;
        const int Ls = (int)((op.nz + nslab - 1) / nslab);
        nslab = (op.nz + Ls - 1) / Ls;
        const double cost_slab = Ls + kFill, cost_item = 1.25 * (per + 2 * kFill);
        const char *force = getenv("MONO_SPLIT");  // "slab" / "item": measurement override
        const bool slab = force && !strcmp(force, "slab") ? true
                          : force && !strcmp(force, "item") ? false
                                                             : cost_slab <= cost_item;
        if(slab && g.ntiles * nslab <= kMaxBlocks) {
            g.L = Ls;
            g.nruns = (int)(g.ntiles * nslab);
            g.per = 0;
            int64_t nbs = g.nruns;
            if(nbs >= 16) nbs = (nbs + 7) / 8 * 8;
            g.nblocks = (int)nbs;
            return g;
        }
    }
#endif
    g.L = 0;
    g.per = per;
    g.nruns = (int)((nitems + g.per - 1) / g.per);
    g.nblocks = (int)nb;
    return g;
}

static int check_pattern(int pat) {
    if(pat != 7 && pat != 19 && pat != 27) {
        mono_set_error("unsupported row pattern %d", pat);
        return -1;
    }
    return 0;
}

static int grid_for(int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if(g > kElemBlocks) g = kElemBlocks;
    if(g >= 64) g = g / 8 * 8;  // multiple of 8 for the XCD chunk order
    return (int)(g < 1 ? 1 : g);
}


int CgWorkspace::init(int64_t n_, int64_t, bool fp32) {
    n = n_;
    const size_t b64 = sizeof(double) * (size_t)(n > 0 ? n : 1), b32 = sizeof(float) * (size_t)(n > 0 ? n : 1);
    double **v64[] = {&r, &p, &p2, &q, &z};
    for(double **ptr : v64) {
        MONO_CHECK(hipMalloc(ptr, b64));
        MONO_CHECK(hipMemset(*ptr, 0, b64));  // inactive cells are never written
    }
    float **v32[] = {&r32, &p32, &p32b, &q32, &d32};
    for(float **ptr : v32) {
        if(!fp32) break;
        MONO_CHECK(hipMalloc(ptr, b32));
        MONO_CHECK(hipMemset(*ptr, 0, b32));
    }
    MONO_CHECK(hipMalloc(&P0, sizeof(double) * kMaxBlocks * kMaxRed));
    MONO_CHECK(hipMalloc(&P1, sizeof(double) * kMaxBlocks * kMaxRed));
    MONO_CHECK(hipMalloc(&sc, sizeof(CgScalars)));
    MONO_CHECK(hipMalloc(&sc_in, sizeof(CgScalars)));
    MONO_CHECK(hipMemset(sc, 0, sizeof(CgScalars)));
    MONO_CHECK(hipMemset(sc_in, 0, sizeof(CgScalars)));
    MONO_CHECK(hipHostMalloc(&host_sc, 2 * sizeof(CgScalars), hipHostMallocDefault));
    MONO_CHECK(hipMalloc(&sweep_its, sizeof(int) * kMaxBlindSweeps));
    MONO_CHECK(hipHostMalloc(&host_its, sizeof(int) * kMaxBlindSweeps, hipHostMallocDefault));
    const char *ug = getenv("MONO_GRAPH");
    use_graphs = ug && ug[0] == '1';
    return 0;
}

int CgWorkspace::release() {
    double *d64[] = {r, p, p2, q, z, P0, P1};
    for(double *ptr : d64)
        if(ptr) MONO_CHECK(hipFree(ptr));
    float *d32s[] = {r32, p32, p32b, q32, d32};
    for(float *ptr : d32s)
        if(ptr) MONO_CHECK(hipFree(ptr));
    if(sc) MONO_CHECK(hipFree(sc));
    if(sc_in) MONO_CHECK(hipFree(sc_in));
    if(host_sc) MONO_CHECK(hipHostFree(host_sc));
    if(sweep_its) MONO_CHECK(hipFree(sweep_its));
    if(host_its) MONO_CHECK(hipHostFree(host_its));
    for(PrefixGraph &g : graphs)
        if(g.exec) MONO_CHECK(hipGraphExecDestroy(g.exec));
    graphs.clear();
    sweep_its = nullptr;
    host_its = nullptr;
    last_sched.clear();
    r = p = p2 = q = z = P0 = P1 = nullptr;
    r32 = p32 = p32b = q32 = d32 = nullptr;
    sc = sc_in = host_sc = nullptr;
    return 0;
}

template <typename T> static bool same_view(const OpView<T> &a, const OpView<T> &b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.n == b.n && a.pat == b.pat && a.table == b.table &&
           a.table_int == b.table_int && a.percell == b.percell && a.codes == b.codes && a.ctab == b.ctab && a.active == b.active &&
           a.nbm == b.nbm && a.lt == b.lt && a.geo[0] == b.geo[0] && a.geo[1] == b.geo[1] && a.geo[2] == b.geo[2] && a.geo[3] == b.geo[3];
}

static bool same_key(const CgWorkspace::PrefixKey &a, const CgWorkspace::PrefixKey &b) {
    const CgConfig &u = a.cfg, &v = b.cfg;
    return a.kind == b.kind && a.sched == b.sched && same_view(a.op64, b.op64) && same_view(a.op32, b.op32) && a.b == b.b &&
           a.x == b.x && u.tol == v.tol && u.rel_tol == v.rel_tol && u.max_its == v.max_its && u.jacobi == v.jacobi &&
           u.mixed == v.mixed && u.inner_rel == v.inner_rel && u.max_outer == v.max_outer && u.poll_chunk == v.poll_chunk &&
           u.tol_dev == v.tol_dev;
}

// First sight of a key: run eagerly and remember it.  Second sight: record the
// launches into a graph (falling back to eager launches if the recording
// fails) and launch it.  Later: launch the graph.  `state` is host state the
// enqueue leaves behind (the mixed solve's buffer parity); a replay restores
// the value the recording left.
template <class F> int CgWorkspace::run_prefix(const PrefixKey &key, hipStream_t s, int &state, F &&enqueue) {
    if(!use_graphs || s == nullptr) return enqueue();
    PrefixGraph *hit = nullptr;
    for(PrefixGraph &g : graphs)
        if(same_key(g.key, key)) hit = &g;
    if(!hit) {
        if(graphs.size() >= (size_t)kMaxPrefixGraphs) {
            size_t old = 0;
            for(size_t i = 1; i < graphs.size(); i++)
                if(graphs[i].last_use < graphs[old].last_use) old = i;
            if(graphs[old].exec) MONO_CHECK(hipGraphExecDestroy(graphs[old].exec));
            graphs.erase(graphs.begin() + old);
        }
        PrefixGraph g;
        g.key = key;
        g.last_use = ++graph_clock;
        graphs.push_back(g);
        return enqueue();
    }
    hit->last_use = ++graph_clock;
    if(hit->failed) return enqueue();
    if(!hit->exec) {
        const int state0 = state;
        hipGraph_t graph = nullptr;
        if(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            hit->failed = true;
            return enqueue();
        }
        const int rc = enqueue();
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        hipError_t ei = hipErrorUnknown;
        if(rc == 0 && ec == hipSuccess && graph) ei = hipGraphInstantiate(&hit->exec, graph, nullptr, nullptr, 0);
        if(graph) MONO_CHECK(hipGraphDestroy(graph));
        if(ei != hipSuccess) {
            (void)hipGetLastError();
            hit->exec = nullptr;
            hit->failed = true;
            state = state0;
            return enqueue();
        }
        hit->post = state;
        graph_captures++;
    }
    MONO_CHECK(hipGraphLaunch(hit->exec, s));
    state = hit->post;
    graph_launches++;
    return 0;
}

static int poll(CgScalars *dev, CgScalars *host, hipStream_t s) {
    MONO_CHECK(hipMemcpyAsync(host, dev, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    MONO_CHECK(hipStreamSynchronize(s));
    return 0;
}

// Launch one stencil kernel instance over its geometry; returns the number of
// partial sums it leaves (its runs; the strip kernel: its blocks), or -1 on a
// HIP error.
template <typename TV, typename TC, int C, int P, bool FUSE, class Epi>
static int launch_stencil(const OpView<TC> &op, const TV *x, const TV *rz, TV *p_out, const Epi &epi, const CgSetup &cs, int mode,
                          int k, const double *Pin, int nprev, double *Pout, CgScalars *sc, CgScalars *so, hipStream_t s) {
    if constexpr(P == 7 && (C == 0 || C == 2)) {
        if(op.nz == 1) {  // one plane: the strip kernel
            const int nitems = ((op.nx + 63) / 64) * ((op.ny + kRows2d - 1) / kRows2d);
            int nb = (nitems + kBlock / 64 - 1) / (kBlock / 64);
            if(nb > kMaxBlocks) nb = kMaxBlocks;  // the waves then loop over the items
            if(nb >= 16) nb = nb / 8 * 8;
            k_stencil2d<TV, TC, C, FUSE, Epi><<<nb, kBlock, 0, s>>>(op, nitems, x, rz, p_out, epi, cs, mode, k, Pin, nprev, Pout, sc, so);
            return nb;
        }
    }
    constexpr auto K = &k_stencil<TV, TC, C, P, FUSE, Epi>;
    static int slots[64] = {};
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if(!slots[dev]) {
        int cus = 0, per_cu = 0;
        MONO_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        MONO_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(K), kBlock, 0));
        slots[dev] = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
        if(getenv("MONO_TRACE_SLOTS"))
            fprintf(stderr, "[mono] stencil<%s,%d,%d,%d,NR=%d>: %d resident blocks (%d x %d CUs)\n", sizeof(TV) == 8 ? "f64" : "f32",
                    C, P, (int)FUSE, Epi::NR, slots[dev], per_cu, cus);
    }
    const TileGeom g = make_geom(op, slots[dev]);
    if(getenv("MONO_TRACE_SLOTS")) {
        static bool said[64] = {};
        if(!said[dev])
            fprintf(stderr, "[mono]   %dx%dx%d: %s, %d runs of %d %s\n", op.nx, op.ny, op.nz, g.L ? "slab runs" : "item runs", g.nruns,
                    g.L ? g.L : g.per, g.L ? "planes" : "items");
        said[dev] = true;
    }
    K<<<g.nblocks, kBlock, 0, s>>>(op, g, x, rz, p_out, epi, cs, mode, k, Pin, nprev, Pout, sc, so);
    return g.nruns;
}

// The launches of an FP64 solve: the residual (k0 == 0; its partial count goes
// to nres) and iterations k0 .. k0 + count - 1.
int CgWorkspace::fp64_launches(const OpView<double> &op, const double *b, double *x, const CgConfig &cfg, int k0, int count,
                               int &nres, hipStream_t s) {
    const int grid = grid_for(op.n);
    const CgSetup cs{cfg.tol, cfg.rel_tol, cfg.inner_rel, cfg.max_its, cfg.jacobi, cfg.max_outer, 0, cfg.tol_dev};
    double *pb[2] = {p, p2};
    const double *rz = cfg.jacobi ? z : r;
    // boxes of more than one plane (the tiled stencil kernel): each
    // iteration's x += alpha p is applied by the next iteration's stencil
    // sweep, which reads p anyway (72 instead of 80 bytes per cell and
    // iteration); the sweep that ends the solve applies the last one
    const bool defer = op.nz > 1;
    if(k0 == 0) {
        // residual: r = b - A x, p = r (or z), partials r.r, r.z, b.b -> P0
        with_op(op, [&](auto C_, auto P_) {
            constexpr int C = decltype(C_)::value, P = decltype(P_)::value;
            if(cfg.jacobi)
                nres = launch_stencil<double, double, C, P, false>(op, x, nullptr, nullptr, EpiRes64<true>{b, r, pb[0], z}, cs, -1, 0,
                                                                   nullptr, 0, P0, sc, nullptr, s);
            else
                nres = launch_stencil<double, double, C, P, false>(op, x, nullptr, nullptr, EpiRes64<false>{b, r, pb[0], z}, cs, -1,
                                                                   0, nullptr, 0, P0, sc, nullptr, s);
        });
        if(nres < 0) return -1;
        MONO_CHECK(hipGetLastError());
    }
    for(int k = k0; k < k0 + count; k++) {
        double *pin = pb[k & 1], *pout = pb[(k + 1) & 1];
        // iteration k: stopping test and beta from P0; p = rz + beta p; q = A p; p.q -> P1
        int nst = 0;
        with_op(op, [&](auto C_, auto P_) {
            constexpr int C = decltype(C_)::value, P = decltype(P_)::value;
            if(defer)
                nst = launch_stencil<double, double, C, P, true>(op, pin, rz, pout, EpiSpmvX<double>{q, x, pin}, cs, k == 0 ? 0 : 1, k,
                                                                 P0, k == 0 ? nres : grid, P1, sc, nullptr, s);
            else
                nst = launch_stencil<double, double, C, P, true>(op, pin, rz, pout, EpiSpmv<double>{q}, cs, k == 0 ? 0 : 1, k, P0,
                                                                 k == 0 ? nres : grid, P1, sc, nullptr, s);
        });
        if(nst < 0) return -1;
        // alpha from P1; x += alpha p (unless deferred); r -= alpha q; r.r, r.z -> P0
        k_update<double, double, double><<<grid, kBlock, 0, s>>>(op, x, pout, q, r, z, cfg.jacobi, k, P1, nst, P0, sc, 0,
                                                                 defer ? 1 : 0);
    }
    MONO_CHECK(hipGetLastError());
    return 0;
}

int CgWorkspace::solve_fp64(const OpView<double> &op, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s) {
    if(check_pattern(op.pat)) return -1;
    int nres = 0;
    auto iterations = [&](int k0, int count) -> int { return fp64_launches(op, b, x, cfg, k0, count, nres, s); };
    // the last solve's iteration count includes the kernel that detected
    // convergence: enqueue that many, so one poll usually sees the solve done
    int chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : (last_iters > 2 ? last_iters : 2);
    PrefixKey key;
    key.kind = 0;
    key.sched.assign(1, chunk);
    key.op64 = op;
    key.b = b;
    key.x = x;
    key.cfg = cfg;
    int no_state = 0;
    if(run_prefix(key, s, no_state, [&]() -> int { return iterations(0, chunk); })) return -1;
    int enq = chunk;
    for(;;) {
        if(poll(sc, host_sc, s)) return -1;
        if(host_sc->done || enq > cfg.max_its + 2) break;
        chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : 2;
        if(iterations(enq, chunk)) return -1;
        enq += chunk;
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
    if(check_pattern(op64.pat)) return -1;
    float *pb[2] = {p32, p32b};
    int inner_total = 0, corrections = 0;
    int chunk0 = last_inner > 2 ? last_inner + 1 : 4;
    CgScalars *h_out = host_sc, *h_in = host_sc + 1;
    // The correction x += d of a sweep is applied by the next sweep's
    // residual, which reads x from one FP64 buffer and writes x + d to the
    // other (x and the workspace's FP64 vector r, unused by this solve).
    double *xbuf[2] = {x, r};
    int cur = 0;
    auto residual = [&](int sweep, const CgSetup &cs) -> int {
        // outer residual in FP64 -> r32; r.r, b.b -> P0
        int nres = 0;
        with_op(op64, [&](auto C_, auto P_) {
            constexpr int C = decltype(C_)::value, P = decltype(P_)::value;
            if(sweep == 0)
                nres = launch_stencil<double, double, C, P, false>(op64, xbuf[cur], nullptr, nullptr, EpiResOuter{b, r32, nullptr, nullptr},
                                                                   cs, -1, 0, nullptr, 0, P0, sc, nullptr, s);
            else
                nres = launch_stencil<double, double, C, P, false>(op64, xbuf[cur], nullptr, xbuf[cur ^ 1],
                                                                   EpiResOuterCorr{b, r32, d32, sc}, cs, -1, 0, nullptr, 0, P0, sc,
                                                                   nullptr, s);
        });
        if(sweep > 0) cur ^= 1;
        return nres;
    };
    auto inner = [&](const CgSetup &cs, int k, int nres) -> int {
        // the first direction is the residual itself
        float *pin = k == 0 ? r32 : pb[k & 1], *pout = pb[(k + 1) & 1];
        int nst = 0;
        with_op(op32, [&](auto C_, auto P_) {
            constexpr int C = decltype(C_)::value, P = decltype(P_)::value;
            nst = launch_stencil<float, float, C, P, true>(op32, pin, r32, pout, EpiSpmv<float>{q32}, cs, k == 0 ? 2 : 1, k, P0,
                                                           k == 0 ? nres : grid, P1, sc_in, sc, s);
        });
        if(nst < 0) return -1;
        k_update<float, float, float><<<grid, kBlock, 0, s>>>(op32, d32, pout, q32, r32, nullptr, 0, k, P1, nst, P0, sc_in, 1, 0);
        return 0;
    };
    // Sweeps enqueued without polling, as many as the previous solve needed,
    // each with the inner iterations it took plus two.  Kernels of a finished
    // (inner or outer) solve exit on the device flags, so a schedule that is
    // too long costs empty launches, one that is too short a partial
    // correction; either way the polled sweeps below finish the solve to the
    // same outer test.  k_note_sweep records what each blind sweep did.
    const int blind = cfg.poll_chunk > 0 ? 0 : (int)last_sched.size();
    auto blind_sweeps = [&]() -> int {
        cur = 0;
        for(int sweep = 0; sweep < blind; sweep++) {
            const CgSetup cs{cfg.tol, cfg.rel_tol, cfg.inner_rel, cfg.max_its, 0, cfg.max_outer, sweep, cfg.tol_dev};
            const int nres = residual(sweep, cs);
            if(nres < 0) return -1;
            for(int k = 0; k < last_sched[sweep] + 2; k++)
                if(inner(cs, k, nres)) return -1;
            k_note_sweep<<<1, 64, 0, s>>>(sc_in, sc, sweep_its + sweep);
            MONO_CHECK(hipGetLastError());
        }
        return 0;
    };
    if(blind > 0) {
        PrefixKey key;
        key.kind = 1;
        key.sched = last_sched;
        key.op64 = op64;
        key.op32 = op32;
        key.b = b;
        key.x = x;
        key.cfg = cfg;
        if(run_prefix(key, s, cur, blind_sweeps)) return -1;
    }
    std::vector<int> sched;
    bool blind_read = blind == 0;
    if(blind > 0) chunk0 = 1;  // the next sweep mostly confirms convergence
    for(int sweep = blind;; sweep++) {
        const CgSetup cs{cfg.tol, cfg.rel_tol, cfg.inner_rel, cfg.max_its, 0, cfg.max_outer, sweep, cfg.tol_dev};
        const int nres = residual(sweep, cs);
        if(nres < 0) return -1;
        MONO_CHECK(hipGetLastError());
        int chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : chunk0;
        int enq = 0;
        for(;;) {
            for(int j = 0; j < chunk; j++)
                if(inner(cs, enq + j, nres)) return -1;
            MONO_CHECK(hipGetLastError());
            enq += chunk;
            MONO_CHECK(hipMemcpyAsync(h_out, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
            MONO_CHECK(hipMemcpyAsync(h_in, sc_in, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
            if(!blind_read) MONO_CHECK(hipMemcpyAsync(host_its, sweep_its, sizeof(int) * blind, hipMemcpyDeviceToHost, s));
            MONO_CHECK(hipStreamSynchronize(s));
            if(!blind_read) {
                // inner iterations of the blind sweeps; -1 marks a sweep after convergence
                for(int k = 0; k < blind && host_its[k] >= 0; k++) {
                    sched.push_back(host_its[k]);
                    inner_total += host_its[k];
                    corrections++;
                }
                blind_read = true;
            }
            if(h_in->done || enq > cfg.max_its + 2) break;
            chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : 2;
        }
        if(h_out->done) break;
        inner_total += h_in->iters - 1;
        sched.push_back(h_in->iters - 1);
        corrections++;
        if(sweep == 0) last_inner = h_in->iters - 1;
        chunk0 = 2;
    }
    // the solution lives in the buffer the last residual wrote
    if(cur != 0) MONO_CHECK(hipMemcpyAsync(x, xbuf[cur], sizeof(double) * op64.n, hipMemcpyDeviceToDevice, s));
    if(sched.size() > (size_t)kMaxBlindSweeps) sched.resize(kMaxBlindSweeps);
    last_sched = sched;
    if(st) {
        st->iterations = inner_total + 1;
        st->error = h_out->rr_true;
        st->outer = corrections + 1;
        st->inner_total = inner_total;
    }
    return 0;
}

int CgWorkspace::solve_csr(const CsrMatrix &A, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s) {
    CsrBuffers w;
    w.r = r;
    w.p = p;
    w.q = q;
    w.z = z;
    w.P0 = P0;
    w.P1 = P1;
    w.sc = sc;
    w.host_sc = host_sc;
    return csr_solve(A, b, x, cfg, st, s, w);
}

int csr_solve(const CsrMatrix &A, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s, const CsrBuffers &w) {
    const int grid = grid_for(A.n);
    const unsigned sgrid = (unsigned)(((int64_t)A.n * 64 + kBlock - 1) / kBlock);
    const CgSetup cs{cfg.tol, cfg.rel_tol, cfg.inner_rel, cfg.max_its, cfg.jacobi, cfg.max_outer, 0, nullptr};
    double *r = w.r, *p = w.p, *q = w.q, *z = w.z, *P0 = w.P0, *P1 = w.P1;
    CgScalars *sc = w.sc, *host_sc = w.host_sc;
    k_csr_spmv<<<sgrid, kBlock, 0, s>>>(A.n, A.row_ptr, A.col, A.val, x, q, nullptr);
    k_csr_res<<<grid, kBlock, 0, s>>>(A.n, q, b, A.diag, r, p, z, cfg.jacobi, P0);
    MONO_CHECK(hipGetLastError());
    const double *rz = cfg.jacobi ? z : r;
    int enq = 0;
    int chunk = cfg.poll_chunk > 0 ? cfg.poll_chunk : 4;
    for(;;) {
        for(int j = 0; j < chunk; j++) {
            const int k = enq + j;
            k_csr_dir<<<grid, kBlock, 0, s>>>(A.n, p, rz, cs, k == 0 ? 0 : 1, k, P0, grid, sc);
            k_csr_spmv<<<sgrid, kBlock, 0, s>>>(A.n, A.row_ptr, A.col, A.val, p, q, sc);
            k_csr_dot<<<grid, kBlock, 0, s>>>(A.n, p, q, P1, sc);
            k_csr_update<<<grid, kBlock, 0, s>>>(A.n, x, p, q, r, z, A.diag, cfg.jacobi, k, P1, grid, P0, sc);
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

__global__ __launch_bounds__(kBlock) void k_csr_diag(int n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                     const double *__restrict__ val, double *__restrict__ diag) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if(i >= n) return;
    double d = 0.0;
    for(int64_t k = rp[i]; k < rp[i + 1]; k++)
        if(col[k] == (int32_t)i) d = val[k];
    diag[i] = d;
}

hipError_t launch_csr_diag(const CsrMatrix &A, double *diag, hipStream_t s) {
    const unsigned g = (unsigned)(((int64_t)A.n + kBlock - 1) / kBlock);
    k_csr_diag<<<g, kBlock, 0, s>>>(A.n, A.row_ptr, A.col, A.val, diag);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_rhs(int64_t n, const double *__restrict__ alpha, const double *__restrict__ vm,
                                                double *__restrict__ b) {
    int64_t beg, end;
    chunk_range(n, beg, end);
    for(int64_t c = beg + threadIdx.x; c < end; c += kBlock) b[c] = vm[c] * alpha[c];
}

hipError_t launch_rhs(int64_t n, const double *alpha, const double *vm, double *b, hipStream_t s) {
    if(n <= 0) return hipSuccess;
    k_rhs<<<grid_for(n), kBlock, 0, s>>>(n, alpha, vm, b);
    return hipGetLastError();
}

}  // namespace mono
