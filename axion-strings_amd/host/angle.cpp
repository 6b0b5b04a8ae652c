// angle.cpp -- host builds of the angle routines of csrc/axs_atan2.hpp: the
// double-double atan2 that the device uses for the string finder's margin
// cases (host and device evaluations of the same inputs agree bit for bit),
// and the |atan2| of the background sums (the same method; the device seeds
// its reciprocals with the hardware instruction, the host with a division).
#include <stdint.h>

#include "axs_atan2.hpp"
#include "axs_host.h"

extern "C" void axs_host_atan2_dd(int64_t n, const double *y, const double *x, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = axs::atan2_dd(y[i], x[i]);
}

extern "C" void axs_host_abs_atan2_fast(int64_t n, const double *y, const double *x, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = axs::abs_atan2_fast(y[i], x[i]);
}

extern "C" void axs_host_abs_atan2_f32in(int64_t n, const float *y, const float *x, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = axs::abs_atan2_f32in(y[i], x[i]);
}
