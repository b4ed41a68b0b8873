// Finite-volume diffusion operator on a structured box (host interface).
//
// The operator replaces the per-cell `elements` rows that ASSEMBLY_MATRIX
// builds in the reference (matrix_assembly.c:334-514) and the CSR that
// INIT_LINEAR_SYSTEM uploads (grid.c:573-649).  Rows are first assembled on the
// device as 27 FP64 coefficients per cell; the operator is then stored in one
// of two forms:
//   - class table: when every row equals the row of its geometric class
//     (z layer x boundary position in x and y), only nz*9 rows are kept and the
//     kernels index them on the fly (homogeneous or layered tissue: zero bytes
//     of coefficient traffic per iteration);
//   - per-cell: otherwise (fibrosis, heterogeneous conductivity), the non-zero
//     slots are stored as structure-of-arrays [slot][cell].
// Both carry an FP64 and an FP32 copy (the FP32 one for the inner solver of
// the mixed-precision path).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mono {

constexpr int kMaxSlots = 27;

// Compile-time row patterns used by the tiled stencil kernels (diagonal first):
// 7 = faces, 19 = faces + in-plane edges + corners (fibres in the xy plane),
// 27 = the full neighbourhood.
template <int P> struct RowPattern;
template <> struct RowPattern<7> {
    static constexpr int n = 7;
    static constexpr int8_t o[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
};
template <> struct RowPattern<19> {
    static constexpr int n = 19;
    static constexpr int8_t o[19][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
                                        {1, 1, 0},   {-1, 1, 0},  {1, -1, 0},  {-1, -1, 0}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1},
                                        {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
};
template <> struct RowPattern<27> {
    static constexpr int n = 27;
    static constexpr int8_t o[27][3] = {{0, 0, 0},    {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
                                        {1, 1, 0},    {-1, 1, 0},  {1, -1, 0},  {-1, -1, 0}, {1, 0, 1},   {-1, 0, 1},  {1, 0, -1},
                                        {-1, 0, -1},  {0, 1, 1},   {0, -1, 1},  {0, 1, -1},  {0, -1, -1}, {-1, -1, -1}, {-1, -1, 1},
                                        {-1, 1, -1},  {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
};

// how the conductivity tensor of a cell is given
enum TensorSource : int {
    kTensorUniform = 0,   // one (sxx, syy, szz, sxy, sxz, syz) for all cells
    kTensorPerLayer = 1,  // one tensor per z layer: [nz][6]
    kTensorPerCell = 2,   // [n][6]
};

struct AssemblyParams {
    int nx, ny, nz;
    double dx, dy, dz;
    double alpha;
    int aniso;                // 0: isotropic 7-point rule, 1: anisotropic rule
    int tensor_source;        // TensorSource
    const double *tensor;     // device pointer, layout per tensor_source
    const double *g;          // optional per-cell conductivity factor (device)
    const uint8_t *active;    // optional per-cell 0/1 (device)
};

// Device view of an operator used by the kernels.  Interior and per-cell rows
// are stored in the slot order of a compile-time RowPattern (diagonal first),
// so the kernels address neighbours with immediate offsets; class rows keep all
// 27 slots in 27-point order (slot = (oz+1)*9 + (oy+1)*3 + (ox+1)).
template <typename T> struct OpView {
    int nx, ny, nz;
    int64_t n;
    int pat;                    // pattern of table_int / percell (7, 19 or 27)
    const T *table;             // class mode: [(z*9 + cy*3 + cx) * 27 + slot], else null
    const T *table_int;         // class mode: [z * pat + k] interior rows, else null
    const T *percell;           // per-cell mode: [k * n + cell], else null
    const uint8_t *codes;       // code mode: [cell] row code (7-point rows only), else null
    const T *ctab;              // code mode: [code * pat + k], kCodes rows, else null
    const uint8_t *active;      // optional
    // mask mode (27-point anisotropic rows formed in the kernel, see
    // aniso_row.h): [cell] neighbourhood mask, [z][6] layer tensors, and
    // dx, dy, dz, alpha; else null
    const uint32_t *nbm;
    const T *lt;
    T geo[4];
};

// Mask mode: for 27-point anisotropic rows on per-layer (or uniform) tensors
// without per-cell conductivity factors, a row is fixed by its layer and which
// of its 26 neighbours exist and are active.  The FP32 view then carries a
// 4-byte mask per cell and the inner solver's kernels form each row in FP32
// registers (aniso_row.h) instead of reading 27 stored coefficients; enabled
// only after the FP64 form of the same function reproduced every assembled
// row bit for bit.  MONO_MASK_ROWS=0 keeps the stored FP32 rows.
//
// Code mode: when every 7-point row is fixed by which of its six neighbours
// exist and are active (a uniform isotropic tensor on a box with inactive
// cells, e.g. many instances packed with gaps, or holes), the row of a cell is
// one of kCodes table rows, picked by a one-byte code: bit 0 the cell's own
// activity, bit k (k = 1..6) neighbour k of RowPattern<7>.  One byte per cell
// replaces 7 FP64 / FP32 coefficients.
constexpr int kCodes = 128;

struct Operator {
    int nx = 0, ny = 0, nz = 0;
    int64_t n = 0;
    int nslots = 0;        // distinct non-zero slots (union over all rows)
    int nint = 0;          // class mode: distinct slots of the interior rows
    int pat = 27;          // pattern of the interior rows (class mode) or of all rows (per-cell)
    bool class_mode = false;
    bool code_mode = false;
    bool mask32 = false;   // FP32 view in mask mode (the FP64 view keeps per-cell rows)
    uint32_t *nbm = nullptr;
    float *lt32 = nullptr;
    double *lt64 = nullptr;
    float geo32[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    double geo64[4] = {0.0, 0.0, 0.0, 0.0};
    uint8_t *codes = nullptr;
    double *ctab64 = nullptr;
    float *ctab32 = nullptr;
    double *table64 = nullptr, *tint64 = nullptr;
    float *table32 = nullptr, *tint32 = nullptr;
    double *percell64 = nullptr;
    float *percell32 = nullptr;
    const uint8_t *active = nullptr;  // not owned
    double alpha = 0.0;

    OpView<double> view64() const;
    OpView<float> view32() const;
    // the FP64 view the tissue solves read (the FP64 CG without Jacobi, and
    // the outer residuals of the mixed-precision solve): in mask mode, rows
    // formed from the masks in FP64 (the assembled rows up to rounding)
    // instead of the stored rows; else view64().  Its percell is null, so a
    // Jacobi solve (diag_of reads percell) takes view64()
    OpView<double> view64_solve() const;
    int release();
};

// Assemble rows on the device into coef27 ([27][n], FP64, caller-allocated).
hipError_t launch_assemble(const AssemblyParams &p, double *coef27, hipStream_t s);

// Build an Operator from assembled rows (coef27 is consumed but not freed).
// Returns 0 on success, -1 on a HIP error (message set).
int build_operator(const AssemblyParams &p, const double *coef27, Operator *op, hipStream_t s);

}  // namespace mono
