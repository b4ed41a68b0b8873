// The tissue object behind mono_tissue* (monoalg_hip.h), shared by tissue.hip
// and the sensitivity code (adjoint.hip).
#pragma once

#include <stdint.h>

#include "../../include/monoalg_hip.h"
#include "cg.h"
#include "observer.h"
#include "operator.h"

struct mono_tissue {
    mono_tissue_desc d{};
    int64_t n = 0;
    int neq = 0;
    size_t pitch = 0;
    void *sv = nullptr;  // double* or float*
    double *v = nullptr;
    double *b = nullptr;
    uint8_t *active = nullptr;
    double *g = nullptr;
    double *tensor = nullptr;
    uint8_t *stim_mask = nullptr;
    float *ionic_scale = nullptr;
    uint16_t *inst = nullptr;      // per-cell instance (batches)
    float *inst_scale = nullptr;   // n_instances x 4 factors
    // batches with a relative CG tolerance: each instance's bounding box
    // (x0, x1, y0, y1, z0, z1 of its active cells), its ||b||^2 and the
    // solve's threshold rel^2 * min_i ||b_i||^2, set after every ionic step
    int *inst_box = nullptr;
    int4 *inst_items = nullptr;  // (instance, box row, first column) per partial sum
    int *inst_first = nullptr;   // first item of each instance, n_instances + 1
    double *inst_part = nullptr; // partial sums, one per item
    double *inst_bb = nullptr;   // ||b_i||^2 per instance
    int inst_nitems = 0;
    double *tol_dev = nullptr;
    double *aux = nullptr;  // adaptive ODE solver: dt, time_new, previous_dt rows of `pitch`
    mono::Operator op;
    mono::CgWorkspace cg;
    mono::CgConfig cgcfg;
    mono::Observer obs;
    mono::ApdTracker apd;
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    // Launch-shape tuner of the ionic step kernel: the first steps of the run
    // time each candidate block size twice (per-cell arithmetic does not depend
    // on it, so the run's results are the same whichever is picked), then the
    // fastest is kept.  Off for the adaptive solver and with MONO_TUNE=0.
    struct IonicTuner {
        int cand[3] = {64, 128, 256};
        float best_ms[3] = {0.f, 0.f, 0.f};
        int trial = -1;    // next trial (-1: tuning off or done)
        int pending = -1;  // candidate whose launch the events bracket
        int chosen = 0;    // block size in use (0: the built-in choice)
        hipEvent_t e0 = nullptr, e1 = nullptr;
    } tune;
    int timing = 0;
    int num_ode_steps = 1;
    double dt_ode = 0.02;
    double alpha = 0.0;
    double stim_start = 0.0;
    double cur_time = 0.0;
    mono_stats st{};
};
