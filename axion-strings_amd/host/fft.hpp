// fft.hpp -- host complex FFT for the initial conditions (the reference
// uses FFTW 3.3.8's fftw_plan_dft_{2d,3d}, init.cpp:75-104, 145-174).
//
// Stockham autosort, decimation in frequency, radices 4, 2, 3, 5 and a
// generic odd radix; every axis of an N^d row-major array is transformed as
// a batch of interleaved lines (a block of neighbouring lines is gathered so
// that strided axes read contiguous runs), lines spread over host threads.
// Unnormalised; sign -1 is FFTW_FORWARD, +1 FFTW_BACKWARD.
#pragma once

#include <complex>
#include <cstdint>

namespace axs_host {

using cplx = std::complex<double>;

// in-place transform of an N^ndims array (last index fastest)
void fft_nd(cplx *data, int ndims, int N, int sign, int threads = 0);

// one line of length n (for tests)
void fft_1d(cplx *data, int n, int sign);

}  // namespace axs_host
