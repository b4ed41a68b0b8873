/*
 * Shared state and helpers of the plugins over libmonoalg_hip.so.  Compiled
 * against the reference's headers (struct layouts and plugin macros).
 */
#ifndef MONO_PLUGIN_COMMON_H
#define MONO_PLUGIN_COMMON_H

#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "config/config_common.h"
#include "monoalg_hip.h"

struct cell_node;

/* The linear solver's persistent data: the device CSR of the active cells and
 * the host staging vectors.  b_on_device: this library's UPDATE_MONODOMAIN
 * formed b on the device for the next solve. */
struct mono_cg_data {
    mono_csr *A;
    uint32_t n;
    struct cell_node **cells;
    double *bvec, *xvec;
    double tol;
    int max_its;
    bool jacobi;
    bool initialised;
    bool b_on_device;
    bool alpha_set;
};

extern struct mono_cg_data mono_cg_state;

const char *mono_plugin_param(const struct config *config, const char *key);

static inline bool mono_plugin_is_true(const char *v) {
    /* IS_TRUE of the reference's config helpers (config_helpers.h:25) */
    return strcmp(v, "true") == 0 || strcmp(v, "yes") == 0 || strcmp(v, "1") == 0;
}

/* the reference's plugins exit on errors (log_error_and_exit) */
static inline void mono_plugin_die(const char *msg) {
    fprintf(stderr, "libmono_hip: %s\n", msg);
    exit(EXIT_FAILURE);
}

static inline void mono_plugin_fail(const char *what) {
    fprintf(stderr, "libmono_hip: %s failed: %s\n", what, mono_last_error());
    exit(EXIT_FAILURE);
}

#endif
