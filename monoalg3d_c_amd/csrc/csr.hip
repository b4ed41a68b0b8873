// CSR conjugate gradient and the right-hand-side update as C ABI entry points
// (monoalg_hip.h): the linear-solver and update-monodomain plugins of the
// reference's dlopen ABI call these.
//
// mono_csr_* keeps one CSR operator resident on the device between solves, as
// the reference's init_gpu_conjugate_gradient keeps its cuSPARSE matrix in
// config->persistent_data (gpu_solvers_cublas_12.c:46-119): the matrix is
// uploaded once, a solve moves only b and x.  mono_cg_csr is the one-shot
// form (upload, solve, free).
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/monoalg_hip.h"
#include "cg.h"
#include "mono_common.h"

using namespace mono;

struct mono_csr {
    int n = 0;
    int64_t nnz = 0;
    int64_t *row_ptr = nullptr;
    int32_t *col = nullptr;
    double *val = nullptr;
    double *diag = nullptr;
    double *b = nullptr;      // right-hand side (uploaded, or formed by mono_csr_update_rhs)
    double *x = nullptr;      // warm start / solution
    double *vm = nullptr;     // mono_csr_update_rhs: uploaded Vm
    double *alpha = nullptr;  // mono_csr_set_alpha: per-row ALPHA
    int b_on_device = 0;      // b holds the last mono_csr_update_rhs result
    CgWorkspace ws;
    hipStream_t stream = nullptr;
};

// row_ptr must start at 0 and never decrease and every column must be a row
// index, so that nothing indexes col/val or x out of range.
static int check_csr(const char *who, int n, const int64_t *row_ptr, const int32_t *col) {
    if(row_ptr[0] != 0) {
        mono_set_error("%s: row_ptr[0] = %lld, expected 0", who, (long long)row_ptr[0]);
        return -1;
    }
    for(int i = 0; i < n; i++)
        if(row_ptr[i + 1] < row_ptr[i]) {
            mono_set_error("%s: row_ptr decreases at row %d", who, i);
            return -1;
        }
    const int64_t nnz = row_ptr[n];
    for(int64_t k = 0; k < nnz; k++)
        if(col[k] < 0 || col[k] >= n) {
            mono_set_error("%s: column index %d out of range", who, col[k]);
            return -1;
        }
    return 0;
}

// ---------------------------------------------------------------------------
// mono_cg_csr_device_noalloc: the CSR solve of mono_cg_csr on operands the
// caller already holds on the device, with every buffer of the solve in one
// fixed area of device memory that the library reserves at load time, so that
// a call allocates nothing.  Layout (doubles): diagonal, r, p, q, z (n each,
// n <= kNoAllocMaxRows), two partial-sum areas (kMaxBlocks * kMaxRed each),
// the scalar record, then the operand check's verdict.  Each device holds its
// own copy of the area; concurrent calls on one device share it and race.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxRows = MONO_CG_CSR_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocVecBytes = sizeof(double) * (size_t)kNoAllocMaxRows;
constexpr size_t kNoAllocPartBytes = sizeof(double) * (size_t)kMaxBlocks * kMaxRed;
constexpr size_t kNoAllocScOff = 5 * kNoAllocVecBytes + 2 * kNoAllocPartBytes;
constexpr size_t kNoAllocCheckOff = kNoAllocScOff + 256;
constexpr size_t kNoAllocBytes = kNoAllocCheckOff + 256;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static void *g_scratch_ptr[kNoAllocMaxDevices];  // device address of scratch, per device
static CgScalars g_noalloc_host_sc[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

// The device address of this device's scratch (taken once per device, with
// that device current; the array's name is not a device address in host code).
static unsigned char *scratch_on_current_device(const char *who) {
    int dev = -1;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) {
        mono_set_error("%s: no current HIP device", who);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if(!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !ptr) {
            mono_set_error("%s: hipGetSymbolAddress(scratch): %s", who, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return (unsigned char *)g_scratch_ptr[dev];
}

// Device-side operand check (the host cannot read device operands without a
// staging buffer): bit 0 row_ptr[0] != 0, bit 1 a decreasing row_ptr, bit 2
// row_ptr[n] < 0, bit 3 a column outside [0, n), ORed into verdict[0], which
// the caller clears first (with a kernel, not a memset).
__global__ void k_clear_verdict(int *verdict) {
    if(threadIdx.x == 0) verdict[0] = 0;
}
__global__ void k_csr_check_rows(int n, const int64_t *__restrict__ rp, int *verdict) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i > n) return;
    int bad = 0;
    if(i == 0 && rp[0] != 0) bad |= 1;
    if(i < n && rp[i + 1] < rp[i]) bad |= 2;
    if(i == n && rp[n] < 0) bad |= 4;
    if(bad) atomicOr(verdict, bad);
}
__global__ void k_csr_check_cols(int n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col, int *verdict) {
    const int64_t nnz = rp[n];
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
        if(col[k] < 0 || col[k] >= n) {
            atomicOr(verdict, 8);
            return;
        }
}

extern "C" {

int mono_cg_csr_device_noalloc(int n, const int64_t *row_ptr, const int32_t *col, const double *val, const double *b,
                               double *x, double tol, int max_its, int jacobi, int *iterations, double *error,
                               void *stream) {
    const char *who = "mono_cg_csr_device_noalloc";
    if(n <= 0 || !row_ptr || !col || !val || !b || !x) {
        mono_set_error("%s: bad arguments", who);
        return -1;
    }
    if((int64_t)n > kNoAllocMaxRows) {
        mono_set_error("%s: %d rows exceed the fixed workspace's maximum of %lld rows", who, n, (long long)kNoAllocMaxRows);
        return -1;
    }
    unsigned char *base = scratch_on_current_device(who);
    if(!base) return -1;
    hipStream_t s = (hipStream_t)stream;
    const int dev = [] {
        int d = 0;
        (void)hipGetDevice(&d);
        return d;
    }();
    // operand check before any kernel indexes col/val through row_ptr
    int *verdict = (int *)(base + kNoAllocCheckOff);
    k_clear_verdict<<<1, 64, 0, s>>>(verdict);
    k_csr_check_rows<<<(unsigned)((n + 1 + 255) / 256), 256, 0, s>>>(n, row_ptr, verdict);
    MONO_CHECK(hipGetLastError());
    int flags = 0;
    MONO_CHECK(hipMemcpyAsync(&flags, verdict, sizeof(int), hipMemcpyDeviceToHost, s));
    MONO_CHECK(hipStreamSynchronize(s));
    if(flags) {
        mono_set_error("%s: malformed row_ptr (%s%s%s)", who, flags & 1 ? "row_ptr[0] != 0 " : "",
                       flags & 2 ? "decreasing entries " : "", flags & 4 ? "row_ptr[n] < 0" : "");
        return -1;
    }
    k_csr_check_cols<<<1024, 256, 0, s>>>(n, row_ptr, col, verdict);
    MONO_CHECK(hipGetLastError());
    MONO_CHECK(hipMemcpyAsync(&flags, verdict, sizeof(int), hipMemcpyDeviceToHost, s));
    MONO_CHECK(hipStreamSynchronize(s));
    if(flags) {
        mono_set_error("%s: column index out of range", who);
        return -1;
    }
    double *v = (double *)base;
    CsrMatrix M;
    M.n = n;
    M.row_ptr = row_ptr;
    M.col = col;
    M.val = val;
    M.diag = v;
    MONO_CHECK(launch_csr_diag(M, v, s));
    CsrBuffers w;
    w.r = v + kNoAllocMaxRows;
    w.p = v + 2 * kNoAllocMaxRows;
    w.q = v + 3 * kNoAllocMaxRows;
    w.z = v + 4 * kNoAllocMaxRows;
    w.P0 = (double *)(base + 5 * kNoAllocVecBytes);
    w.P1 = (double *)(base + 5 * kNoAllocVecBytes + kNoAllocPartBytes);
    w.sc = (CgScalars *)(base + kNoAllocScOff);
    w.host_sc = &g_noalloc_host_sc[dev];
    CgConfig cfg;
    cfg.tol = tol;
    cfg.max_its = max_its;
    cfg.jacobi = jacobi;
    CgStats st{};
    if(csr_solve(M, b, x, cfg, &st, s, w)) return -1;
    MONO_CHECK(hipStreamSynchronize(s));
    if(iterations) *iterations = st.iterations;
    if(error) *error = st.error;
    return 0;
}

int mono_csr_destroy(mono_csr *A) {
    if(!A) return 0;
    int rc = 0;
    if(A->stream && hipStreamSynchronize(A->stream) != hipSuccess) rc = -1;
    if(A->ws.release()) rc = -1;
    void *bufs[] = {A->row_ptr, A->col, A->val, A->diag, A->b, A->x, A->vm, A->alpha};
    for(void *p : bufs)
        if(p && hipFree(p) != hipSuccess) rc = -1;
    if(A->stream && hipStreamDestroy(A->stream) != hipSuccess) rc = -1;
    delete A;
    return rc;
}

int mono_csr_create(int n, const int64_t *row_ptr, const int32_t *col, const double *val, mono_csr **out) {
    if(n <= 0 || !row_ptr || !col || !val || !out) {
        mono_set_error("mono_csr_create: bad arguments");
        return -1;
    }
    if(check_csr("mono_csr_create", n, row_ptr, col)) return -1;
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    mono_csr *A = new mono_csr();
    A->n = n;
    A->nnz = row_ptr[n];
    const size_t nz = (size_t)(A->nnz > 0 ? A->nnz : 1);
    auto fail = [&](hipError_t e) {
        mono_set_error("mono_csr_create: %s", hipGetErrorString(e));
        mono_csr_destroy(A);
        return -1;
    };
    hipError_t e = hipStreamCreate(&A->stream);
    if(e == hipSuccess) e = hipMalloc(&A->row_ptr, sizeof(int64_t) * (n + 1));
    if(e == hipSuccess) e = hipMalloc(&A->col, sizeof(int32_t) * nz);
    if(e == hipSuccess) e = hipMalloc(&A->val, sizeof(double) * nz);
    if(e == hipSuccess) e = hipMalloc(&A->diag, sizeof(double) * n);
    if(e == hipSuccess) e = hipMalloc(&A->b, sizeof(double) * n);
    if(e == hipSuccess) e = hipMalloc(&A->x, sizeof(double) * n);
    if(e == hipSuccess) e = hipMemcpyAsync(A->row_ptr, row_ptr, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, A->stream);
    if(e == hipSuccess && A->nnz > 0)
        e = hipMemcpyAsync(A->col, col, sizeof(int32_t) * A->nnz, hipMemcpyHostToDevice, A->stream);
    if(e == hipSuccess && A->nnz > 0)
        e = hipMemcpyAsync(A->val, val, sizeof(double) * A->nnz, hipMemcpyHostToDevice, A->stream);
    if(e != hipSuccess) return fail(e);
    CsrMatrix M;
    M.n = n;
    M.nnz = A->nnz;
    M.row_ptr = A->row_ptr;
    M.col = A->col;
    M.val = A->val;
    e = launch_csr_diag(M, A->diag, A->stream);
    if(e == hipSuccess) e = hipStreamSynchronize(A->stream);
    if(e != hipSuccess) return fail(e);
    if(A->ws.init(n, n, false)) {
        mono_csr_destroy(A);
        return -1;
    }
    *out = A;
    return 0;
}

int mono_csr_num_rows(const mono_csr *A) { return A ? A->n : -1; }

int mono_csr_set_alpha(mono_csr *A, const double *alpha) {
    if(!A || !alpha) {
        mono_set_error("mono_csr_set_alpha: bad arguments");
        return -1;
    }
    if(!A->alpha) MONO_CHECK(hipMalloc(&A->alpha, sizeof(double) * A->n));
    if(!A->vm) MONO_CHECK(hipMalloc(&A->vm, sizeof(double) * A->n));
    MONO_CHECK(hipMemcpyAsync(A->alpha, alpha, sizeof(double) * A->n, hipMemcpyHostToDevice, A->stream));
    MONO_CHECK(hipStreamSynchronize(A->stream));
    return 0;
}

int mono_csr_update_rhs(mono_csr *A, const double *vm) {
    if(!A || !vm) {
        mono_set_error("mono_csr_update_rhs: bad arguments");
        return -1;
    }
    if(!A->alpha) {
        mono_set_error("mono_csr_update_rhs: mono_csr_set_alpha has not been called");
        return -1;
    }
    MONO_CHECK(hipMemcpyAsync(A->vm, vm, sizeof(double) * A->n, hipMemcpyHostToDevice, A->stream));
    MONO_CHECK(launch_rhs(A->n, A->alpha, A->vm, A->b, A->stream));
    A->b_on_device = 1;
    return 0;
}

int mono_csr_get_rhs(mono_csr *A, double *b) {
    if(!A || !b) {
        mono_set_error("mono_csr_get_rhs: bad arguments");
        return -1;
    }
    MONO_CHECK(hipMemcpyAsync(b, A->b, sizeof(double) * A->n, hipMemcpyDeviceToHost, A->stream));
    MONO_CHECK(hipStreamSynchronize(A->stream));
    return 0;
}

int mono_csr_solve(mono_csr *A, const double *b, double *x, double tol, int max_its, int jacobi, int *iterations,
                   double *error) {
    if(!A || !x) {
        mono_set_error("mono_csr_solve: bad arguments");
        return -1;
    }
    if(!b && !A->b_on_device) {
        mono_set_error("mono_csr_solve: no right-hand side (b is NULL and mono_csr_update_rhs has not run)");
        return -1;
    }
    const int n = A->n;
    if(b) MONO_CHECK(hipMemcpyAsync(A->b, b, sizeof(double) * n, hipMemcpyHostToDevice, A->stream));
    A->b_on_device = 0;
    MONO_CHECK(hipMemcpyAsync(A->x, x, sizeof(double) * n, hipMemcpyHostToDevice, A->stream));
    CsrMatrix M;
    M.n = n;
    M.nnz = A->nnz;
    M.row_ptr = A->row_ptr;
    M.col = A->col;
    M.val = A->val;
    M.diag = A->diag;
    CgConfig cfg;
    cfg.tol = tol;
    cfg.max_its = max_its;
    cfg.jacobi = jacobi;
    CgStats st{};
    if(A->ws.solve_csr(M, A->b, A->x, cfg, &st, A->stream)) return -1;
    MONO_CHECK(hipMemcpyAsync(x, A->x, sizeof(double) * n, hipMemcpyDeviceToHost, A->stream));
    MONO_CHECK(hipStreamSynchronize(A->stream));
    if(iterations) *iterations = st.iterations;
    if(error) *error = st.error;
    return 0;
}

int mono_cg_csr(int n, const int64_t *row_ptr, const int32_t *col, const double *val, const double *b, double *x, double tol,
                int max_its, int jacobi, int *iterations, double *error) {
    if(n <= 0 || !row_ptr || !col || !val || !b || !x) {
        mono_set_error("mono_cg_csr: bad arguments");
        return -1;
    }
    if(check_csr("mono_cg_csr", n, row_ptr, col)) return -1;
    mono_csr *A = nullptr;
    const int rc = mono_csr_create(n, row_ptr, col, val, &A);
    if(rc) return rc;
    const int rs = mono_csr_solve(A, b, x, tol, max_its, jacobi, iterations, error);
    if(mono_csr_destroy(A) || rs) return -1;
    return 0;
}

int mono_update_rhs(int64_t n, const double *alpha, const double *vm, double *b) {
    if(n < 0 || !alpha || !vm || !b) {
        mono_set_error("mono_update_rhs: bad arguments");
        return -1;
    }
    if(mono_device_count() < 1) {
        mono_set_error("no HIP device available");
        return -2;
    }
    if(n == 0) return 0;
    double *d = nullptr;
    MONO_CHECK(hipMalloc(&d, sizeof(double) * 3 * n));
    hipError_t e = hipMemcpy(d, alpha, sizeof(double) * n, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(d + n, vm, sizeof(double) * n, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = launch_rhs(n, d, d + n, d + 2 * n, nullptr);
    if(e == hipSuccess) e = hipMemcpy(b, d + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost);
    const hipError_t ef = hipFree(d);
    if(e == hipSuccess) e = ef;
    if(e != hipSuccess) {
        mono_set_error("mono_update_rhs: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

}  // extern "C"
