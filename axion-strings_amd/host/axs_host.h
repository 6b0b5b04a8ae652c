/*
 * axs_host.h -- C entry points of the host driver library (libaxs_host.so):
 * the parts of the reference's host code that sit around the GPU path
 * (parameter file, physics scalars, initial conditions, output files, the
 * run loop).  The driver executable axs_main is main() over these.
 */
#ifndef AXS_HOST_H
#define AXS_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parses a parameter file (read_parameters.cpp:182-238 rules) and writes the
 * values as one JSON object into buf.  0 on success; 1 on a parameter-file
 * error (message printed as the reference does); 2 if buf is too small.    */
int axs_host_params_json(const char *path, char *buf, int cap);

/* Initial fields on the host: mode 0 = k^-1 complex field (<|phi|^2> = 1,
 * phidot = 0), mode 1 = thermal spectrum of gaussian_thermal (init.cpp).
 * 0 on success.                                                          */
int axs_host_init_fields(int mode, int ndims, int N, float space_step, float lambda,
                         float T_initial, uint64_t seed, double *phi1, double *phi2,
                         double *phidot1, double *phidot2);

/* One Philox4x32-10 block (Salmon et al., SC'11): the counter-based
 * generator behind both initial-condition generators (host and device).  */
void axs_host_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* In-place complex FFT of an N^ndims (ndims 1..3) interleaved array. */
int axs_host_fft(double *data, int ndims, int N, int sign);

/* final_step of s_run.cpp:82 (float arithmetic as there). */
int axs_host_final_step(int N, float space_step, float time_step);

/* xi of s_run.cpp:111-116 (2D, num_cores) and :123-129 (3D, plaquettes). */
float axs_host_xi(int ndims, int64_t count, double tau, int N, float space_step);

/* atan2(y[i], x[i]) by the double-double evaluation the device uses for
 * the string finder's margin cases (correctly rounded but for inputs within
 * ~2^-100 of a rounding boundary), the same bits as the device build.      */
void axs_host_atan2_dd(int64_t n, const double *y, const double *x, double *out);

/* |atan2(y[i], x[i])| by the table method the background sums use on the
   device (within 6e-16 of the exact angle). */
void axs_host_abs_atan2_fast(int64_t n, const double *y, const double *x, double *out);

/* The same for float arguments (the FP32 form's phi), by the shortcut the
   FP32 background sums take: the same bits as axs_host_abs_atan2_fast. */
void axs_host_abs_atan2_f32in(int64_t n, const float *y, const float *x, double *out);

/* run_standard (s_run.cpp:26-232) with the parameters of `path`; returns 0
 * on success.  This is what axs_main runs.                                */
int axs_host_run(const char *path);

#ifdef __cplusplus
}
#endif
#endif
