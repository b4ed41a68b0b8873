// Conjugate-gradient solver for the diffusion step (host interface of cg.hip).
//
// Replaces SOLVE_LINEAR_SYSTEM / INIT_LINEAR_SYSTEM (linear_system_solver.c:38-231,
// gpu_solvers_cublas_12.c:46-232).  All vectors, scalars and the convergence
// test stay on the device; the host only enqueues kernels and polls a small
// status record every few iterations.
#pragma once

#include <vector>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mono_common.h"
#include "operator.h"

namespace mono {

// Per-block partial sums: at most kMaxBlocks blocks (or runs) per reduction
// kernel, at most kMaxRed sums each, stored sum-major (sum k of block b at
// k * kMaxBlocks + b).
constexpr int kMaxBlocks = 8192;
constexpr int kMaxRed = 3;

struct CsrMatrix {
    int n = 0;
    int64_t nnz = 0;
    const int64_t *row_ptr = nullptr;  // device
    const int32_t *col = nullptr;      // device
    const double *val = nullptr;       // device
    const double *diag = nullptr;      // device, optional (Jacobi)
};

// Stopping rule: the reference's absolute test on r.r (linear_system_solver.c:95,159-163)
// or a relative one (r.r <= rel^2 * b.b).
struct CgConfig {
    double tol = 1e-16;      // absolute r.r threshold (reference semantics)
    double rel_tol = 0.0;    // if > 0: stop on ||r|| <= rel_tol * ||b|| instead
    int max_its = 200;
    int jacobi = 0;
    // mixed precision (FP32 inner CG, FP64 outer residual and correction)
    int mixed = 0;
    double inner_rel = 1e-4;  // inner stop: ||r_in|| <= inner_rel * ||r_in0||
    int max_outer = 20;
    int poll_chunk = 0;       // iterations enqueued between status polls (0 = adaptive)
    // if set: the outer r.r threshold, a device scalar written before the
    // solve (replaces tol and rel_tol; batches of instances set it to
    // rel^2 * min_i ||b_i||^2 each step)
    const double *tol_dev = nullptr;
};

// Device buffers of a CSR solve: r, p, q, z (n doubles each), two partial-sum
// buffers (kMaxBlocks * kMaxRed doubles each), the device scalar record and a
// host record the solve polls into (pinned or not).
struct CsrBuffers {
    double *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr;
    double *P0 = nullptr, *P1 = nullptr;
    CgScalars *sc = nullptr;
    CgScalars *host_sc = nullptr;
};

struct CgStats {
    int iterations = 0;       // as the reference counts them (1 + CG iterations)
    double error = 0.0;       // final r.r (FP64)
    int outer = 0;            // mixed: outer sweeps
    int inner_total = 0;      // mixed: total inner iterations
};

// CSR conjugate gradient (FP64) on caller-provided buffers.
int csr_solve(const CsrMatrix &A, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s,
              const CsrBuffers &w);
// diag[i] = A(i, i) (0 where the row has no diagonal entry; the last entry wins)
hipError_t launch_csr_diag(const CsrMatrix &A, double *diag, hipStream_t s);
// b[i] = alpha[i] * vm[i] (UPDATE_MONODOMAIN)
hipError_t launch_rhs(int64_t n, const double *alpha, const double *vm, double *b, hipStream_t s);
class CgWorkspace {
  public:
    // fp32 = false: only the FP64 vectors (CSR and FP64 solves)
    int init(int64_t n, int64_t pitch_hint, bool fp32 = true);
    int release();
    // Solve A x = b in FP64 (x is the warm start and the result).
    int solve_fp64(const OpView<double> &op, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s);
    // Mixed precision: FP64 outer defect correction around an FP32 CG.
    int solve_mixed(const OpView<double> &op64, const OpView<float> &op32, const double *b, double *x, const CgConfig &cfg,
                    CgStats *st, hipStream_t s);
    // CSR variant (non-uniform meshes), FP64.
    int solve_csr(const CsrMatrix &A, const double *b, double *x, const CgConfig &cfg, CgStats *st, hipStream_t s);

    int64_t n = 0;
    // FP64 vectors
    double *r = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr, *z = nullptr;
    // FP32 vectors (mixed inner solve)
    float *r32 = nullptr, *p32 = nullptr, *p32b = nullptr, *q32 = nullptr, *d32 = nullptr;
    double *P0 = nullptr, *P1 = nullptr;  // per-block partial sums (alternating between kernels)
    CgScalars *sc = nullptr;        // device scalars (outer / FP64 solve)
    CgScalars *sc_in = nullptr;     // device scalars (inner solve)
    CgScalars *host_sc = nullptr;   // pinned mirror for polling
    int last_iters = 0;
    int last_inner = 0;
    // mixed solve: inner iterations of each correction sweep of the last solve
    // (the next solve enqueues that many sweeps without polling)
    std::vector<int> last_sched;
    int *sweep_its = nullptr;  // device: per blind sweep, inner iterations or -1
    int *host_its = nullptr;   // pinned mirror

    // The launches a solve enqueues before its first poll (the mixed solve's
    // blind sweeps; the FP64 solve's residual and first chunk of iterations)
    // are a fixed sequence for a given schedule and operands.  On a stream of
    // its own (not the null stream) the workspace records that sequence as a
    // hipGraph the second time it meets the same key (by then every kernel
    // instance in it has run, so none needs its first-use calibration) and
    // replays the graph afterwards: one launch instead of ~40.  The sequence
    // holds no memset nodes.  Opt-in (MONO_GRAPH=1 in the environment at
    // init): the eager launches already keep the GPU busy back to back (idle
    // gaps are ~1% of a 16M-cell step and ~3.5% of a batch step), and the
    // replays measured 0.4% (16M cells) to 2% (batch) slower.
    struct PrefixKey {
        int kind = 0;  // 0 FP64, 1 mixed
        std::vector<int> sched;
        OpView<double> op64{};
        OpView<float> op32{};
        const double *b = nullptr;
        const double *x = nullptr;
        CgConfig cfg;
    };
    struct PrefixGraph {
        PrefixKey key;
        hipGraphExec_t exec = nullptr;
        bool failed = false;  // recording failed once: stay eager for this key
        int post = 0;         // host state the recorded enqueue left
        unsigned long last_use = 0;
    };
    int use_graphs = 0;
    unsigned long graph_clock = 0;
    std::vector<PrefixGraph> graphs;
    long graph_launches = 0, graph_captures = 0;
    // Enqueue `enqueue`'s launches on s, through a recorded graph when the key
    // has been seen before; returns 0 or -1 (HIP error).
    template <class F> int run_prefix(const PrefixKey &key, hipStream_t s, int &state, F &&enqueue);

  private:
    int fp64_launches(const OpView<double> &op, const double *b, double *x, const CgConfig &cfg, int k0, int count, int &nres,
                      hipStream_t s);
};

constexpr int kMaxBlindSweeps = 8;
constexpr int kMaxPrefixGraphs = 8;

}  // namespace mono
