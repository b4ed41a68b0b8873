// The ionic models' source (monoalg3d_c_amd/csrc/ionic_models.h) compiled for
// the host, so the CPU tests can compare the device code's arithmetic with the
// oracle without a GPU (exp is the models' own function, host and device
// alike; log, pow and expm1 come from the host libm here and from the
// device's on the GPU, within an ulp of each other).
#include "../../monoalg3d_c_amd/csrc/ionic_models.h"

extern "C" {

// one FP64 step of n cells, states AoS (n x neq) in place; returns 0
int host_step(int model, double *y, const double *stim, int64_t n, double dt) {
    for(int64_t i = 0; i < n; i++) {
        double *c = y + i * (model == 0 ? 19 : (model == 1 ? 41 : 2));
        const double s = stim ? stim[i] : 0.0;
        double dv;
        if(model == 0) dv = mono::tt2006_step<double>(c, s, dt, 1.0, 1.0, 1.0);
        else if(model == 1) dv = mono::ord2011_step<double>(c, s, dt);
        else if(model == 2) dv = mono::ms2003_step<double>(c, s, dt, 1.0);
        else return -1;
        c[0] = c[0] + dv;
    }
    return 0;
}

double host_exp(double x) { return mono::exp_cw(x); }

}
