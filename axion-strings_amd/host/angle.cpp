// angle.cpp -- host build of the double-double atan2 that the device uses
// for the string finder's margin cases (csrc/axs_atan2.hpp), so that the
// host and device evaluations of the same inputs can be compared bit for bit.
#include <stdint.h>

#include "axs_atan2.hpp"
#include "axs_host.h"

extern "C" void axs_host_atan2_dd(int64_t n, const double *y, const double *x, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = axs::atan2_dd(y[i], x[i]);
}
