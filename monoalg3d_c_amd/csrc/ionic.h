// Host-side launch interface of the ionic kernels (ionic.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mono {

template <typename T> struct IonicLaunch {
    T *sv;                  // SoA state, neq rows of `pitch` elements
    size_t pitch;           // elements per row
    double *v;              // FP64 potential (read; the previous diffusion solution)
    double *b;              // FP64 right-hand side (written: alpha * V*), may be null
    int64_t n;              // cells
    const uint8_t *active;  // optional 0/1 per cell
    const double *stim_cell;  // optional per-cell stimulus (overrides the x-limit rule)
    const uint8_t *stim_mask; // optional per-cell 0/1: stimulated cells (replaces the x-limit rule)
    const float *scale;     // optional per-cell conductance scalings, 4 per cell
    // optional per-instance scalings: the instance of each cell (inst) and a
    // table of 4 factors per instance (inst_scale, ninst rows), staged in LDS
    // by each block; replaces `scale` when set
    const uint16_t *inst;
    const float *inst_scale;
    int ninst;
    int nx;                 // cells along x (for the centre of a cell)
    double dx;              // cell size along x (um)
    double stim_x_limit;    // cells with centre x < limit are stimulated
    double stim_amp;        // stimulus amplitude this step (0 outside the window)
    double alpha;           // ALPHA(beta, cm, dt_pde, dx, dy, dz)
    T dt;                   // ODE step
    int num_steps;          // ODE steps per PDE step
    int first_step;         // read V from sv row 0 instead of v (t = 0: no Vm mirror)
    int block;              // step kernels: threads per block (0: the built-in choice)
    int write_vm_row;       // also store V* in sv row 0
    // adaptive Euler (launch_ionic_adaptive): per-cell dt, time_new and
    // previous_dt in three FP64 rows of `pitch`; each cell advances from its
    // time_new to final_time
    double *aux;
    double final_time, min_dt, max_dt, abs_tol, rel_tol;
};

// Substep cap of one adaptive call per cell: a cell that would need more (a
// non-finite state, whose time never reaches final_time, or a step size held
// at a tiny min_dt for a whole call) stops with V = NaN, which the tissue's
// NaN check reports, instead of looping on.
constexpr int kMaxAdaptiveSubsteps = 1 << 20;

// largest per-instance table (4 floats per instance, staged in LDS: 64 KB)
constexpr int kMaxInstances = 4096;

hipError_t launch_ionic_step(int model, const IonicLaunch<double> &a, hipStream_t s);
hipError_t launch_ionic_step(int model, const IonicLaunch<float> &a, hipStream_t s);
hipError_t launch_ionic_adaptive(int model, const IonicLaunch<double> &a, hipStream_t s);
hipError_t launch_ionic_adaptive(int model, const IonicLaunch<float> &a, hipStream_t s);
hipError_t launch_ionic_init(int model, double *sv, size_t pitch, double *v, int64_t n, const double *ic, hipStream_t s);
hipError_t launch_ionic_init(int model, float *sv, size_t pitch, double *v, int64_t n, const float *ic, hipStream_t s);
void default_initial_state(int model, double *out);
double model_initial_v(int model);

}  // namespace mono
