/*
 * axs_oracle.h -- CPU oracle for the Axion-Strings uniform-lattice path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and in the
 * reference's own expression order, the arithmetic of the reference's
 * standard evolution path.  Only the test suite, the smoke check and the
 * benchmark's CPU-baseline leg load it, and only as the checker: the product
 * library never links, loads or calls it.
 *
 * Conventions (reference src/common/common.hpp:28, 81-119):
 *   2D: m = i + N*j             (i fastest)
 *   3D: m = (i*N + j)*N + k     (k fastest)
 * All field arrays are double (the reference's USE_DOUBLE_PRECISION data_t).
 */
#ifndef AXS_ORACLE_H
#define AXS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int   ndims;            /* 2 or 3                              */
    int   N;                /* points per side                     */
    float space_step;       /* dx, float as in parameters          */
    float time_step;        /* dt, float as in parameters          */
    float lambda;
    float T_initial;
    int   stencil_setting;  /* 0..3                                */
    int   thr;              /* string finder threshold (percent)   */
} orc_params;

/* opaque: the stencil matrix in CSR form (rows in COO insertion order) */
typedef struct orc_matrix orc_matrix;

int64_t orc_length(int ndims, int N);

/* build_coefficient_matrix, src/standard/s_integrate.cpp:5-149 */
orc_matrix *orc_matrix_create(int ndims, int N, int stencil_setting);
void        orc_matrix_destroy(orc_matrix *A);
int64_t     orc_matrix_nnz(const orc_matrix *A);
/* dense export of a small matrix (row-major L x L), for fixture checks */
int         orc_matrix_dense(const orc_matrix *A, double *out, int64_t cap);
/* y = alpha * A x, s_integrate.cpp:322-324 via mkl_wrapper.cpp:11-17 */
void        orc_spmv(const orc_matrix *A, double alpha, const double *x, double *y);

/* the inline reference stencils, src/standard/s_internal.hpp:26-35 */
double orc_laplacian2D(const double *phi, int i, int j, float dx, int N);
double orc_laplacian3D(const double *phi, int i, int j, int k, float dx, int N);

/* vvsl_field_rescaled, src/standard/s_integrate.cpp:305-348.
 * Arrays of length L; ker*_next are scratch.  *tau is advanced by dt.   */
void orc_vvsl_step(const orc_params *p, const orc_matrix *A,
                   double *phi1, double *phi2, double *phidot1, double *phidot2,
                   double *ker1_curr, double *ker2_curr,
                   double *ker1_next, double *ker2_next, double *tau);

/* n steps of the above; convenience loop for tests and timing */
void orc_vvsl_run(const orc_params *p, const orc_matrix *A, int64_t nsteps,
                  double *phi1, double *phi2, double *phidot1, double *phidot2,
                  double *ker1_curr, double *ker2_curr,
                  double *ker1_next, double *ker2_next, double *tau);

/* The same step as a fused CPU stencil stepper (a CPU baseline, not the
 * reference's structure): one drift sweep in place, then one sweep that
 * evaluates the Laplacian matrix-free in the matrix's row order, the
 * potential, the kick and the ker update -- two sweeps instead of five plus
 * two SpMVs, no matrix, no ker_next arrays.  Bit-identical to orc_vvsl_run
 * (same expression trees and summation order); ker must hold ker_curr.   */
void orc_vvsl_run_fused(const orc_params *p, int64_t nsteps,
                        double *phi1, double *phi2, double *phidot1, double *phidot2,
                        double *ker1, double *ker2, double *tau);

/* atan2(y, x) rounded correctly to double (the oracle's libm atan2) */
double orc_atan2_cr(double y, double x);
void orc_atan2_cr_array(int64_t n, const double *y, const double *x, double *out);

/* axion = atan2(phi1, phi2), src/standard/s_run.cpp:92-94 */
void orc_axion(int64_t L, const double *phi1, const double *phi2, double *axion);

/* cores2 / cores3, src/common/string_finding.cpp:3-86.
 * Writes (i,j) or (i,j,k) triples in ascending site order (the single-
 * thread order) into coords when coords != NULL and room allows; returns
 * the number of flagged sites.                                           */
int64_t orc_cores2(const double *axion, int N, int thr, int32_t *coords, int64_t cap);
int64_t orc_cores3(const double *axion, int N, int thr, int32_t *coords, int64_t cap);

/* num_cores, src/standard/s_run.cpp:101-109 (coords from orc_cores2) */
int64_t orc_num_cores2(const int32_t *coords, int64_t n);

/* xi, src/standard/s_run.cpp:111-131, float arithmetic as there */
float orc_xi2(int64_t num_cores, double tau, int N, float dx);
float orc_xi3(int64_t n_plaquettes, double tau, int N, float dx);

/* six background sums, src/standard/s_run.cpp:135-150 (unnormalised) */
void orc_background(int64_t L, const double *phi1, const double *phi2,
                    const double *phidot1, const double *phidot2, double out[6]);

/* energy components (kinetic, gradient, potential); no reference
 * counterpart -- defined so that its variation is the step's force.     */
void orc_energy(const orc_params *p, double tau,
                const double *phi1, const double *phi2,
                const double *phidot1, const double *phidot2, double out[3]);

/* in-place complex FFT over an N^ndims array of interleaved (re,im),
 * sign = -1 forward (FFTW_FORWARD), +1 backward; unnormalised.          */
int orc_fft(int ndims, int N, double *data, int sign);

/* compute_laplacian_fft, s_integrate.cpp:179-296 (ret = spectral
 * Laplacian of field, continuum -k^2 on every mode)                      */
int orc_laplacian_fft(int ndims, int N, float dx, const double *field, double *ret);

/* vvsl_hamiltonian_form, s_integrate.cpp:358-395 (v = pi = phidot a^2)   */
void orc_vvsl_hamiltonian_step(const orc_params *p, const orc_matrix *A,
                               double *phi1, double *phi2, double *pi1, double *pi2,
                               double *ker1_curr, double *ker2_curr,
                               double *ker1_next, double *ker2_next, double *tau);
/* vvsl_field_rescaled with compute_laplacian_fft (s_integrate.cpp:326-328) */
int orc_vvsl_spectral_step(const orc_params *p, int64_t length,
                           double *phi1, double *phi2, double *phidot1, double *phidot2,
                           double *ker1_curr, double *ker2_curr,
                           double *ker1_next, double *ker2_next, double *tau);
/* nsteps of the form: eom 0 field-rescaled / 1 Hamiltonian, laplacian 0
 * stencil matrix / 1 spectral (field-rescaled only); A may be NULL for the
 * spectral form                                                          */
int orc_vvsl_run_form(const orc_params *p, const orc_matrix *A, int eom, int laplacian,
                      int64_t nsteps, double *phi1, double *phi2, double *v1, double *v2,
                      double *ker1_curr, double *ker2_curr, double *ker1_next,
                      double *ker2_next, double *tau);

/* output_powerspec, src/common/pk_estimation.cpp:37-214.
 * data_imag may be NULL.  Outputs have N/2 entries.                     */
void orc_powerspec(int ndims, int N, float dx, const double *data_real,
                   const double *data_imag, double *ks, double *pk, int32_t *count);

/* log-spaced binning form: nbins bins over 1 <= |k| < N/2 (lattice units),
 * pk = mean |F|^2 per mode, ks = mean |k| * 2 pi / L; data_imag may be NULL. */
void orc_powerspec_log(int ndims, int N, float dx, const double *data_real,
                       const double *data_imag, int nbins, double *ks, double *pk,
                       int32_t *count);

#ifdef __cplusplus
}
#endif
#endif
