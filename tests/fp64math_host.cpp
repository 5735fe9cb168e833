// Host build of chimera_amd/csrc/fp64math.h for its accuracy tests (g++, no HIP).
#define __host__
#define __device__
#define __forceinline__ inline
#include "../chimera_amd/csrc/fp64math.h"

extern "C" void fp64_eval(int which, const double *x, double *y, long n) {
    for (long i = 0; i < n; ++i) {
        switch (which) {
            case 0: y[i] = chm::cos2pi(x[i], chm::kPolyCoefs); break;
            case 1: y[i] = chm::sin2pi(x[i], chm::kPolyCoefs); break;
            case 2: { double s, c; chm::sincos2pi(x[i], chm::kPolyCoefs, &s, &c); y[2 * i] = s; y[2 * i + 1] = c; } break;
            default: y[i] = chm::exp_fin(x[i], chm::kPolyCoefs); break;
        }
    }
}
