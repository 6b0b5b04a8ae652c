/*
 * axs_oracle.c -- CPU oracle for the Axion-Strings uniform-lattice path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared with.
 * Loaded by tests/, by the smoke check and by the benchmark's CPU-baseline
 * leg; never linked or called by the product library.
 *
 * Every function restates one routine of the reference (SeanBryan51/
 * Axion-Strings), keeping its expression trees, its float/double promotions
 * and its loop structure (separate OpenMP sweeps, a sparse-matrix Laplacian).
 * Where the reference's result depends on something its sources do not fix
 * (MKL's SpMV summation order, FFTW's rounding, OpenMP merge order) the
 * oracle fixes a canonical choice, stated at the function.
 *
 * Pinning (what checks each function against something other than itself):
 *   matrix/spmv      : reference fixture tests/fdm_matrix.cpp:47-64 (N=4),
 *                      the zero / plane-wave KATs of tests/laplacian.cpp and
 *                      tests/fdm_matrix.cpp, the discrete eigenvalue
 *                      -(4/dx^2) sum sin^2(pi n/N) in closed form.
 *   vvsl step        : exact vacuum solution phi = tau e^{i theta}; the lambda=0
 *                      plane wave against the closed-form leapfrog dispersion;
 *                      first step with ker=0 (phi_1 = phi_0 + dt phidot_0).
 *   cores2/cores3    : hand-built vortex / straight-string lattices, the
 *                      winding-number bound (every winding plaquette passes),
 *                      single raised corners and two-corner lattices that
 *                      fix each plaquette's corners (plaquette b's (i,j+1,k+1)),
 *                      the float threshold at thr = 0, 1, 5, 30, 100.
 *   atan2 (axion)    : mpmath's correctly rounded atan2 on random and edge
 *                      inputs; C99 Annex F special values.
 *   background       : closed forms on constant and antisymmetric fields.
 *   energy           : closed forms (uniform field, single plane wave) and
 *                      its variation equals the step's force (finite diff).
 *   fft / powerspec  : numpy.fft on random data; brute-force mode counts;
 *                      single-cosine field; Parseval.
 */
#include "axs_oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------- */
/* indexing: src/common/common.hpp:28 (periodic), 81-119 (offset/coordinate) */

static inline int64_t periodic(int64_t i, int64_t N)
{
    return (i >= 0) ? (i % N) : (N - ((-i) % N)) % N;
}

static inline int64_t offset2(int64_t i, int64_t j, int64_t N)
{
    return periodic(i, N) + N * periodic(j, N);
}

static inline int64_t offset3(int64_t i, int64_t j, int64_t k, int64_t N)
{
    return (periodic(i, N) * N + periodic(j, N)) * N + periodic(k, N);
}

int64_t orc_length(int ndims, int N)
{
    return (ndims == 3) ? (int64_t)N * N * N : (int64_t)N * N;
}

/* ---------------------------------------------------------------------- */
/* stencil matrix: build_coefficient_matrix, s_integrate.cpp:5-149        */

struct orc_matrix {
    int64_t  L;
    int64_t  nnz;
    int64_t *rowptr;
    int32_t *cols;
    double  *vals;
};

/* coefficient table, s_integrate.cpp:9-14; computed in float and stored as
 * double exactly as the reference's initialiser list does                 */
static int stencil_coefficients(int stencil, int ndims, double *c)
{
    float nd = (float)ndims;
    switch (stencil) {
    case 0:
        c[0] = (double)(-2.0f * nd); c[1] = (double)1.0f;
        return 2;
    case 1:
        c[0] = (double)(-5.0f / 2.0f * nd); c[1] = (double)(4.0f / 3.0f);
        c[2] = (double)(-1.0f / 12.0f);
        return 3;
    case 2:
        c[0] = (double)(-49.0f / 18.0f * nd); c[1] = (double)(3.0f / 2.0f);
        c[2] = (double)(-3.0f / 20.0f); c[3] = (double)(1.0f / 90.0f);
        return 4;
    case 3:
        c[0] = (double)(-205.0f / 72.0f * nd); c[1] = (double)(8.0f / 5.0f);
        c[2] = (double)(-1.0f / 5.0f); c[3] = (double)(8.0f / 315.0f);
        c[4] = (double)(-1.0f / 560.0f);
        return 5;
    default:
        return 0;
    }
}

/* Rows are stored in the reference's COO insertion order: the centre, then
 * for each distance l the +l/-l neighbours along the first coordinate, then
 * the second, then the third (s_integrate.cpp:43-75, 98-141).  In 2D the
 * first coordinate is the fast index i; in 3D it is the slow index.        */
orc_matrix *orc_matrix_create(int ndims, int N, int stencil)
{
    double c[5];
    int ncoef = stencil_coefficients(stencil, ndims, c);
    if (ncoef == 0 || (ndims != 2 && ndims != 3) || N < 1) return NULL;

    orc_matrix *A = (orc_matrix *)calloc(1, sizeof(orc_matrix));
    if (!A) return NULL;
    int64_t L = orc_length(ndims, N);
    int64_t per_row = 1 + 2 * ndims * (ncoef - 1);
    A->L = L;
    A->nnz = per_row * L;
    A->rowptr = (int64_t *)malloc((size_t)(L + 1) * sizeof(int64_t));
    A->cols = (int32_t *)malloc((size_t)A->nnz * sizeof(int32_t));
    A->vals = (double *)malloc((size_t)A->nnz * sizeof(double));
    if (!A->rowptr || !A->cols || !A->vals) {
        orc_matrix_destroy(A);
        return NULL;
    }

    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < L; r++) {
        int64_t e = r * per_row;
        A->rowptr[r] = e;
        if (ndims == 2) {
            int64_t x = r % N, y = r / N;            /* coordinate2 */
            A->cols[e] = (int32_t)offset2(x, y, N); A->vals[e++] = c[0];
            for (int l = 1; l < ncoef; l++) {
                A->cols[e] = (int32_t)offset2(x + l, y, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset2(x - l, y, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset2(x, y + l, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset2(x, y - l, N); A->vals[e++] = c[l];
            }
        } else {
            int64_t x = r / ((int64_t)N * N);          /* coordinate3 */
            int64_t layer = r - x * N * N;
            int64_t y = layer / N, z = layer % N;
            A->cols[e] = (int32_t)offset3(x, y, z, N); A->vals[e++] = c[0];
            for (int l = 1; l < ncoef; l++) {
                A->cols[e] = (int32_t)offset3(x + l, y, z, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset3(x - l, y, z, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset3(x, y + l, z, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset3(x, y - l, z, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset3(x, y, z + l, N); A->vals[e++] = c[l];
                A->cols[e] = (int32_t)offset3(x, y, z - l, N); A->vals[e++] = c[l];
            }
        }
    }
    A->rowptr[L] = A->nnz;
    return A;
}

void orc_matrix_destroy(orc_matrix *A)
{
    if (!A) return;
    free(A->rowptr);
    free(A->cols);
    free(A->vals);
    free(A);
}

int64_t orc_matrix_nnz(const orc_matrix *A) { return A ? A->nnz : 0; }

int orc_matrix_dense(const orc_matrix *A, double *out, int64_t cap)
{
    if (!A || cap < A->L * A->L) return -1;
    memset(out, 0, (size_t)(A->L * A->L) * sizeof(double));
    for (int64_t r = 0; r < A->L; r++)
        for (int64_t e = A->rowptr[r]; e < A->rowptr[r + 1]; e++)
            out[r * A->L + A->cols[e]] += A->vals[e];
    return 0;
}

/* y = alpha * (A x).  MKL's summation order is not fixed by the reference's
 * sources; the oracle sums each row in stored (COO insertion) order,
 * starting from the centre term, and scales by alpha last.               */
void orc_spmv(const orc_matrix *A, double alpha, const double *x, double *y)
{
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < A->L; r++) {
        int64_t e = A->rowptr[r];
        double sum = A->vals[e] * x[A->cols[e]];
        for (e = e + 1; e < A->rowptr[r + 1]; e++)
            sum = sum + A->vals[e] * x[A->cols[e]];
        y[r] = alpha * sum;
    }
}

/* s_internal.hpp:26-28 */
double orc_laplacian2D(const double *phi, int i, int j, float dx, int N)
{
    double ddx = (double)dx;
    return (phi[offset2(i + 1, j, N)] + phi[offset2(i, j + 1, N)] - 4.0f * phi[offset2(i, j, N)]
            + phi[offset2(i - 1, j, N)] + phi[offset2(i, j - 1, N)]) / (ddx * ddx);
}

/* s_internal.hpp:33-35 */
double orc_laplacian3D(const double *phi, int i, int j, int k, float dx, int N)
{
    double ddx = (double)dx;
    return (phi[offset3(i + 1, j, k, N)] + phi[offset3(i, j + 1, k, N)] + phi[offset3(i, j, k + 1, N)]
            - 6.0f * phi[offset3(i, j, k, N)] + phi[offset3(i - 1, j, k, N)]
            + phi[offset3(i, j - 1, k, N)] + phi[offset3(i, j, k - 1, N)]) / (ddx * ddx);
}

/* ---------------------------------------------------------------------- */
/* vvsl_field_rescaled, s_integrate.cpp:305-348.  Five sweeps, as there.   */

void orc_vvsl_step(const orc_params *p, const orc_matrix *A,
                   double *phi1, double *phi2, double *phidot1, double *phidot2,
                   double *ker1_curr, double *ker2_curr,
                   double *ker1_next, double *ker2_next, double *tau)
{
    int64_t length = A->L;
    float dt = p->time_step;                       /* :308 */
    float dx = p->space_step;                      /* :309 */
    double ddt = (double)dt;

    /* drift, :311-315 -- phi += dt * (phidot + 0.5f * ker * dt) */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        phi1[i] = phi1[i] + ddt * (phidot1[i] + 0.5 * ker1_curr[i] * ddt);
        phi2[i] = phi2[i] + ddt * (phidot2[i] + 0.5 * ker2_curr[i] * ddt);
    }

    *tau = *tau + ddt;                              /* :317 */
    double t = *tau;

    /* Laplacian, :322-324 -- alpha = 1.0f / pow_2(dx) */
    double alpha = (double)1.0f / ((double)dx * (double)dx);
    orc_spmv(A, alpha, phi1, ker1_next);
    orc_spmv(A, alpha, phi2, ker2_next);

    /* potential, :331-335.  Left-to-right:
     * ((((-1.0f / pow_2(tau)) * lambda) * phi) * (phi1^2 + phi2^2 - tau^2 + T0^2/3)) */
    double lam = (double)p->lambda;
    double T0 = (double)p->T_initial;
    double c = (double)(-1.0f) / (t * t);
    double tau2 = t * t;
    double T03 = (T0 * T0) / (double)3.0f;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        double s = phi1[i] * phi1[i] + phi2[i] * phi2[i] - tau2 + T03;
        ker1_next[i] = ker1_next[i] + c * lam * phi1[i] * s;
        ker2_next[i] = ker2_next[i] + c * lam * phi2[i] * s;
    }

    /* kick, :337-341 -- phidot += 0.5f * (kc + kn) * dt */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        phidot1[i] = phidot1[i] + 0.5 * (ker1_curr[i] + ker1_next[i]) * ddt;
        phidot2[i] = phidot2[i] + 0.5 * (ker2_curr[i] + ker2_next[i]) * ddt;
    }

    /* copy, :343-347 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        ker1_curr[i] = ker1_next[i];
        ker2_curr[i] = ker2_next[i];
    }
}

void orc_vvsl_run(const orc_params *p, const orc_matrix *A, int64_t nsteps,
                  double *phi1, double *phi2, double *phidot1, double *phidot2,
                  double *ker1_curr, double *ker2_curr,
                  double *ker1_next, double *ker2_next, double *tau)
{
    for (int64_t s = 0; s < nsteps; s++)
        orc_vvsl_step(p, A, phi1, phi2, phidot1, phidot2,
                      ker1_curr, ker2_curr, ker1_next, ker2_next, tau);
}

/* ---------------------------------------------------------------------- */
/* vvsl_field_rescaled as a fused stencil stepper (CPU baseline).  The row
 * sum runs in the matrix's COO order (centre; per distance l: +-first,
 * +-second, +-third coordinate) with the same coefficients, alpha last,
 * then the potential, kick and ker update of s_integrate.cpp:331-347 on
 * that row -- each site's values depend only on its own ker/phidot and the
 * drifted field, so the update is in place after the drift sweep.       */

static void fused_row(int64_t m, double x1, double x2, double s1, double s2, double alpha,
                      double c, double lam, double tau2, double T03, double ddt,
                      double *phidot1, double *phidot2, double *ker1, double *ker2)
{
    double k1 = alpha * s1, k2 = alpha * s2;
    double s = x1 * x1 + x2 * x2 - tau2 + T03;
    k1 = k1 + c * lam * x1 * s;
    k2 = k2 + c * lam * x2 * s;
    phidot1[m] = phidot1[m] + 0.5 * (ker1[m] + k1) * ddt;
    phidot2[m] = phidot2[m] + 0.5 * (ker2[m] + k2) * ddt;
    ker1[m] = k1;
    ker2[m] = k2;
}

void orc_vvsl_run_fused(const orc_params *p, int64_t nsteps,
                        double *phi1, double *phi2, double *phidot1, double *phidot2,
                        double *ker1, double *ker2, double *tau)
{
    double cf[5];
    const int ncoef = stencil_coefficients(p->stencil_setting, p->ndims, cf);
    const int64_t N = p->N, L = orc_length(p->ndims, p->N);
    const float dt = p->time_step, dx = p->space_step;
    const double ddt = (double)dt;
    const double alpha = (double)1.0f / ((double)dx * (double)dx);
    const double lam = (double)p->lambda, T0 = (double)p->T_initial;
    const double T03 = (T0 * T0) / (double)3.0f;
    for (int64_t step = 0; step < nsteps; step++) {
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < L; i++) {
            phi1[i] = phi1[i] + ddt * (phidot1[i] + 0.5 * ker1[i] * ddt);
            phi2[i] = phi2[i] + ddt * (phidot2[i] + 0.5 * ker2[i] * ddt);
        }
        *tau = *tau + ddt;
        const double t = *tau, c = (double)(-1.0f) / (t * t), tau2 = t * t;
        if (p->ndims == 2) {
            /* m = x + N y; first coordinate x (fast), second y */
            const int R = ncoef - 1;
            #pragma omp parallel for schedule(static)
            for (int64_t y = 0; y < N; y++) {
                const int64_t row = N * y;
                int64_t yb[4][2];    /* [l-1]: rows y+l, y-l */
                for (int l = 1; l <= R; l++) {
                    yb[l - 1][0] = N * periodic(y + l, N);
                    yb[l - 1][1] = N * periodic(y - l, N);
                }
                for (int64_t x = 0; x < N; x++) {
                    const int64_t m = x + row;
                    const int edge = x < R || x >= N - R;
                    double s1 = cf[0] * phi1[m], s2 = cf[0] * phi2[m];
                    for (int l = 1; l <= R; l++) {
                        const int64_t nb[4] = {edge ? row + periodic(x + l, N) : m + l,
                                               edge ? row + periodic(x - l, N) : m - l,
                                               yb[l - 1][0] + x, yb[l - 1][1] + x};
                        for (int q = 0; q < 4; q++) {
                            s1 = s1 + cf[l] * phi1[nb[q]];
                            s2 = s2 + cf[l] * phi2[nb[q]];
                        }
                    }
                    fused_row(m, phi1[m], phi2[m], s1, s2, alpha, c, lam, tau2, T03, ddt,
                              phidot1, phidot2, ker1, ker2);
                }
            }
        } else {
            /* m = (x N + y) N + z; first coordinate x (slow), then y, z.
             * Row bases of the x and y neighbours once per row; interior z
             * without wrapping. */
            const int R = ncoef - 1;
            #pragma omp parallel for collapse(2) schedule(static)
            for (int64_t x = 0; x < N; x++) {
                for (int64_t y = 0; y < N; y++) {
                    const int64_t row = (x * N + y) * N;
                    int64_t nb[4][4];    /* [l-1]: x+l, x-l, y+l, y-l row bases */
                    for (int l = 1; l <= R; l++) {
                        nb[l - 1][0] = (periodic(x + l, N) * N + y) * N;
                        nb[l - 1][1] = (periodic(x - l, N) * N + y) * N;
                        nb[l - 1][2] = (x * N + periodic(y + l, N)) * N;
                        nb[l - 1][3] = (x * N + periodic(y - l, N)) * N;
                    }
                    for (int64_t z = 0; z < N; z++) {
                        const int64_t m = row + z;
                        const int edge = z < R || z >= N - R;
                        double s1 = cf[0] * phi1[m], s2 = cf[0] * phi2[m];
                        for (int l = 1; l <= R; l++) {
                            const int64_t zp = edge ? row + periodic(z + l, N) : m + l;
                            const int64_t zm = edge ? row + periodic(z - l, N) : m - l;
                            const int64_t q[6] = {nb[l - 1][0] + z, nb[l - 1][1] + z,
                                                  nb[l - 1][2] + z, nb[l - 1][3] + z, zp, zm};
                            for (int e = 0; e < 6; e++) {
                                s1 = s1 + cf[l] * phi1[q[e]];
                                s2 = s2 + cf[l] * phi2[q[e]];
                            }
                        }
                        fused_row(m, phi1[m], phi2[m], s1, s2, alpha, c, lam, tau2, T03, ddt,
                                  phidot1, phidot2, ker1, ker2);
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* compute_laplacian_fft, s_integrate.cpp:179-296: the field times
 * (-1)^(i+j[+k]) (shift2D/3D, :153-177), forward complex FFT, each mode
 * times -k^2 with k[i] = (-N/2 + i) 2 pi / L on every axis (:192, :242),
 * backward FFT, the shift again, real part divided by the length.        */

int orc_laplacian_fft(int ndims, int N, float dx, const double *field, double *ret)
{
    if (ndims != 2 && ndims != 3) return -1;
    int64_t length = orc_length(ndims, N);
    float L = N * dx;                                              /* :183 */
    double *k = (double *)malloc((size_t)N * sizeof(double));
    double *data = (double *)malloc((size_t)length * 2 * sizeof(double));
    if (!k || !data) {
        free(k);
        free(data);
        return -1;
    }
    /* (- N / 2.0f + i) * 2.0f * M_PI / L: float, float, then double */
    for (int i = 0; i < N; i++) k[i] = (double)((-N / 2.0f + i) * 2.0f) * M_PI / (double)L;
    #pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < length; m++) {                         /* :205-208 */
        int64_t c0 = m % N, c1 = (m / N) % N, c2 = (ndims == 3) ? m / ((int64_t)N * N) : 0;
        double sgn = ((c0 + c1 + c2) & 1) ? -1.0 : 1.0;            /* pow(-1.0f, i+j+k) */
        data[2 * m] = field[m] * sgn;
        data[2 * m + 1] = 0.0 * sgn;
    }
    int rc = orc_fft(ndims, N, data, -1);                          /* FFTW_FORWARD */
    #pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < length; m++) {                         /* :214-220 */
        double k_sq;
        if (ndims == 2) {
            int64_t i = m % N, j = m / N;                          /* coordinate2 */
            k_sq = k[i] * k[i] + k[j] * k[j];
        } else {
            int64_t i = m / ((int64_t)N * N), j = (m / N) % N, kk = m % N;   /* coordinate3 */
            k_sq = k[i] * k[i] + k[j] * k[j] + k[kk] * k[kk];
        }
        data[2 * m] *= -k_sq;
        data[2 * m + 1] *= -k_sq;
    }
    if (rc == 0) rc = orc_fft(ndims, N, data, +1);                 /* FFTW_BACKWARD */
    #pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < length; m++) {                         /* :226-232 */
        int64_t c0 = m % N, c1 = (m / N) % N, c2 = (ndims == 3) ? m / ((int64_t)N * N) : 0;
        double sgn = ((c0 + c1 + c2) & 1) ? -1.0 : 1.0;
        ret[m] = (data[2 * m] * sgn) / (double)length;
    }
    free(k);
    free(data);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* vvsl_hamiltonian_form, s_integrate.cpp:358-395: tau advances first, the
 * drift uses dt / tau^2 and pi = phidot * a^2, the Laplacian carries
 * alpha = tau^2 / dx^2 and the potential -tau^2 lambda phi (|phi|^2 - 1 +
 * T0^2 / (3 tau^2)).                                                      */

void orc_vvsl_hamiltonian_step(const orc_params *p, const orc_matrix *A,
                               double *phi1, double *phi2, double *pi1, double *pi2,
                               double *ker1_curr, double *ker2_curr,
                               double *ker1_next, double *ker2_next, double *tau)
{
    int64_t length = A->L;
    float dt = p->time_step;                                       /* :361 */
    float dx = p->space_step;                                      /* :362 */
    double ddt = (double)dt;

    *tau = *tau + ddt;                                             /* :364 */
    double t = *tau;
    double t2 = t * t;

    /* :368-369 -- phi += dt / pow_2(tau) * (pi + 0.5f * dt * ker) */
    double cd = ddt / t2;
    double hdt = (double)(0.5f * dt);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        phi1[i] = phi1[i] + cd * (pi1[i] + hdt * ker1_curr[i]);
        phi2[i] = phi2[i] + cd * (pi2[i] + hdt * ker2_curr[i]);
    }

    /* :374-376 -- alpha = pow_2(tau) / pow_2(dx) */
    double alpha = t2 / ((double)dx * (double)dx);
    orc_spmv(A, alpha, phi1, ker1_next);
    orc_spmv(A, alpha, phi2, ker2_next);

    /* :379-380 -- ker += -pow_2(tau) * lambda * phi *
     *            (phi1^2 + phi2^2 - 1.0f + pow_2(T_initial) / (3.0f * pow_2(tau))) */
    double ch = (-t2) * (double)p->lambda;
    double T0 = (double)p->T_initial;
    double X = (T0 * T0) / ((double)3.0f * t2);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        double s = phi1[i] * phi1[i] + phi2[i] * phi2[i] - 1.0 + X;
        ker1_next[i] = ker1_next[i] + ch * phi1[i] * s;
        ker2_next[i] = ker2_next[i] + ch * phi2[i] * s;
    }

    /* :384-385, :389-392 -- kick and copy as in the field-rescaled form */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        pi1[i] = pi1[i] + 0.5 * (ker1_curr[i] + ker1_next[i]) * ddt;
        pi2[i] = pi2[i] + 0.5 * (ker2_curr[i] + ker2_next[i]) * ddt;
        ker1_curr[i] = ker1_next[i];
        ker2_curr[i] = ker2_next[i];
    }
}

/* vvsl_field_rescaled with compute_laplacian_fft in place of the sparse
 * matrix (the #else branch, s_integrate.cpp:326-328); everything else as
 * orc_vvsl_step.                                                          */
int orc_vvsl_spectral_step(const orc_params *p, int64_t length,
                           double *phi1, double *phi2, double *phidot1, double *phidot2,
                           double *ker1_curr, double *ker2_curr,
                           double *ker1_next, double *ker2_next, double *tau)
{
    float dt = p->time_step;
    double ddt = (double)dt;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        phi1[i] = phi1[i] + ddt * (phidot1[i] + 0.5 * ker1_curr[i] * ddt);
        phi2[i] = phi2[i] + ddt * (phidot2[i] + 0.5 * ker2_curr[i] * ddt);
    }
    *tau = *tau + ddt;
    double t = *tau;
    int rc = orc_laplacian_fft(p->ndims, p->N, p->space_step, phi1, ker1_next);
    if (rc == 0) rc = orc_laplacian_fft(p->ndims, p->N, p->space_step, phi2, ker2_next);
    if (rc != 0) return rc;
    double lam = (double)p->lambda;
    double T0 = (double)p->T_initial;
    double c = (double)(-1.0f) / (t * t);
    double tau2 = t * t;
    double T03 = (T0 * T0) / (double)3.0f;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        double s = phi1[i] * phi1[i] + phi2[i] * phi2[i] - tau2 + T03;
        ker1_next[i] = ker1_next[i] + c * lam * phi1[i] * s;
        ker2_next[i] = ker2_next[i] + c * lam * phi2[i] * s;
        phidot1[i] = phidot1[i] + 0.5 * (ker1_curr[i] + ker1_next[i]) * ddt;
        phidot2[i] = phidot2[i] + 0.5 * (ker2_curr[i] + ker2_next[i]) * ddt;
        ker1_curr[i] = ker1_next[i];
        ker2_curr[i] = ker2_next[i];
    }
    return 0;
}

int orc_vvsl_run_form(const orc_params *p, const orc_matrix *A, int eom, int laplacian,
                      int64_t nsteps, double *phi1, double *phi2, double *v1, double *v2,
                      double *ker1_curr, double *ker2_curr, double *ker1_next,
                      double *ker2_next, double *tau)
{
    int64_t length = orc_length(p->ndims, p->N);
    if (eom == 1 && laplacian != 0) return -1;     /* the reference has no such pair */
    for (int64_t s = 0; s < nsteps; s++) {
        if (eom == 1) {
            orc_vvsl_hamiltonian_step(p, A, phi1, phi2, v1, v2, ker1_curr, ker2_curr,
                                      ker1_next, ker2_next, tau);
        } else if (laplacian == 1) {
            int rc = orc_vvsl_spectral_step(p, length, phi1, phi2, v1, v2, ker1_curr,
                                            ker2_curr, ker1_next, ker2_next, tau);
            if (rc) return rc;
        } else {
            orc_vvsl_step(p, A, phi1, phi2, v1, v2, ker1_curr, ker2_curr, ker1_next,
                          ker2_next, tau);
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* strings                                                                 */

/* atan2 rounded correctly to double.  The reference calls libm's atan2
 * (s_run.cpp:94, 145, 189), whose last bit depends on the library (glibc
 * 2.35 misrounds about 0.1% of random arguments by one ulp; macOS's libm, the
 * reference's own build platform, is another implementation); the oracle
 * fixes the canonical value, the correctly rounded one.  Method: the x87
 * extended-precision atan2l (64-bit significand, error within an ulp of it)
 * rounded to double, unless it lies within 2^-60 (relative) of the midpoint
 * between two doubles; there libquadmath's atan2q (113-bit significand)
 * decides.  Zeros, infinities and NaN take libm's C99 Annex F values.    */
double orc_atan2_cr(double y, double x)
{
    if (isnan(x) || isnan(y) || isinf(x) || isinf(y) || x == 0.0 || y == 0.0)
        return atan2(y, x);
    long double l = atan2l((long double)y, (long double)x);
    double d = (double)l;
    if ((long double)d == l) return d;
    double nb = (l > (long double)d) ? nextafter(d, INFINITY) : nextafter(d, -INFINITY);
    long double mid = ((long double)d + (long double)nb) * 0.5L;      /* exact */
    if (fabsl(l - mid) > fabsl(mid) * 0x1p-60L) return d;
    return (double)atan2q((__float128)y, (__float128)x);
}

void orc_atan2_cr_array(int64_t n, const double *y, const double *x, double *out)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = orc_atan2_cr(y[i], x[i]);
}

void orc_axion(int64_t L, const double *phi1, const double *phi2, double *axion)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L; i++) axion[i] = orc_atan2_cr(phi1[i], phi2[i]);   /* s_run.cpp:94 */
}

/* std::min(a, b) = (b < a) ? b : a ; theta = min(|d|, 1 - |d|)          */
static inline double edge_theta(double a, double b)
{
    double d = fabs(b - a);
    double e = 1 - d;
    return (e < d) ? e : d;
}

/* string_finding.cpp:8 -- accept = 0.5f * (1.0f - thr/100.0f), in float  */
static inline double accept_threshold(int thr)
{
    float f = 0.5f * (1.0f - (float)thr / 100.0f);
    return (double)f;
}

/* cores2, string_finding.cpp:3-33.  The reference merges thread-local
 * lists, so its order depends on the thread count; the oracle emits the
 * single-thread order (ascending m).                                    */
int64_t orc_cores2(const double *field, int N, int thr, int32_t *coords, int64_t cap)
{
    int64_t length = (int64_t)N * N;
    double accept = accept_threshold(thr);
    unsigned char *flag = (unsigned char *)malloc((size_t)length);
    if (!flag) return -1;

    #pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < length; m++) {
        int64_t i = m % N, j = m / N;
        double norm1 = (field[offset2(i, j, N)] + M_PI) / (2 * M_PI);
        double norm2 = (field[offset2(i + 1, j, N)] + M_PI) / (2 * M_PI);
        double norm3 = (field[offset2(i + 1, j + 1, N)] + M_PI) / (2 * M_PI);
        double norm4 = (field[offset2(i, j + 1, N)] + M_PI) / (2 * M_PI);
        double theta1 = edge_theta(norm1, norm2);
        double theta2 = edge_theta(norm2, norm3);
        double theta3 = edge_theta(norm3, norm4);
        double theta_sum = theta1 + theta2 + theta3;
        flag[m] = theta_sum > accept;
    }

    int64_t n = 0;
    for (int64_t m = 0; m < length; m++) {
        if (!flag[m]) continue;
        if (coords && n < cap) {
            coords[2 * n] = (int32_t)(m % N);
            coords[2 * n + 1] = (int32_t)(m / N);
        }
        n++;
    }
    free(flag);
    return n;
}

/* cores3, string_finding.cpp:35-86, including plaquette "b" with corner
 * (i, j+1, k+1) (:56).  One entry per flagged site, ascending m.         */
int64_t orc_cores3(const double *field, int N, int thr, int32_t *coords, int64_t cap)
{
    int64_t NN = (int64_t)N * N;
    int64_t length = NN * N;
    double accept = accept_threshold(thr);
    double twopi = 2.0f * M_PI;
    unsigned char *flag = (unsigned char *)malloc((size_t)length);
    if (!flag) return -1;

    #pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < length; m++) {
        int64_t i = m / NN, layer = m - i * NN, j = layer / N, k = layer % N;
        double n1a = (field[offset3(i, j, k, N)] + M_PI) / twopi;
        double n2a = (field[offset3(i + 1, j, k, N)] + M_PI) / twopi;
        double n3a = (field[offset3(i + 1, j + 1, k, N)] + M_PI) / twopi;
        double n4a = (field[offset3(i, j + 1, k, N)] + M_PI) / twopi;

        double n1b = n1a;
        double n2b = n2a;
        double n3b = (field[offset3(i, j + 1, k + 1, N)] + M_PI) / twopi;
        double n4b = (field[offset3(i, j, k + 1, N)] + M_PI) / twopi;

        double n1c = n1a;
        double n2c = n4a;                        /* (i, j+1, k)   */
        double n3c = n3b;                        /* (i, j+1, k+1) */
        double n4c = n4b;                        /* (i, j, k+1)   */

        double sa = edge_theta(n1a, n2a) + edge_theta(n2a, n3a) + edge_theta(n3a, n4a);
        double sb = edge_theta(n1b, n2b) + edge_theta(n2b, n3b) + edge_theta(n3b, n4b);
        double sc = edge_theta(n1c, n2c) + edge_theta(n2c, n3c) + edge_theta(n3c, n4c);
        flag[m] = (sa > accept || sb > accept || sc > accept);
    }

    int64_t n = 0;
    for (int64_t m = 0; m < length; m++) {
        if (!flag[m]) continue;
        if (coords && n < cap) {
            int64_t i = m / NN, layer = m - i * NN;
            coords[3 * n] = (int32_t)i;
            coords[3 * n + 1] = (int32_t)(layer / N);
            coords[3 * n + 2] = (int32_t)(layer % N);
        }
        n++;
    }
    free(flag);
    return n;
}

/* s_run.cpp:101-109 */
int64_t orc_num_cores2(const int32_t *coords, int64_t n)
{
    int64_t count = 0;
    for (int64_t i = 0; i < n - 1; i++) {
        int diff_y = coords[2 * (i + 1) + 1] - coords[2 * i + 1];
        int diff_x = coords[2 * (i + 1)] - coords[2 * i];
        if (diff_y == 0 && diff_x == 1) count += 1;
        if (diff_y == 1 && diff_x == 0) count += 1;
    }
    return n - count;
}

/* physics.cpp:38-45 with tau held in double, returning float */
static float physical_time_f(double tau) { return (float)(0.5f * (tau * tau)); }
static float scale_factor_f(double tau) { return (float)tau; }

/* s_run.cpp:111-116 */
float orc_xi2(int64_t num_cores, double tau, int N, float dx)
{
    float t_phys = physical_time_f(tau);
    float a = scale_factor_f(tau);
    float physical_area = (float)(((double)(N * dx * a)) * ((double)(N * dx * a)));
    float xi = (float)((double)((int)num_cores * ((double)t_phys * (double)t_phys)) / physical_area);
    return xi;
}

/* s_run.cpp:123-131 */
float orc_xi3(int64_t n_plaquettes, double tau, int N, float dx)
{
    float total_length = (int)n_plaquettes * 2.0f / 3.0f;
    float t_phys = physical_time_f(tau);
    float a = scale_factor_f(tau);
    float physical_length = total_length * a;
    double side = (double)(N * dx * a);
    float physical_volume = (float)(side * side * side);
    float xi = (float)((double)physical_length * ((double)t_phys * (double)t_phys) / physical_volume);
    return xi;
}

/* ---------------------------------------------------------------------- */
/* background sums, s_run.cpp:135-150.  The reference uses an OpenMP
 * reduction (thread-dependent order); the oracle sums in site order.      */
void orc_background(int64_t L, const double *phi1, const double *phi2,
                    const double *phidot1, const double *phidot2, double out[6])
{
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    for (int64_t i = 0; i < L; i++) {
        s0 += phi1[i];
        s1 += phi2[i];
        s2 += phidot1[i];
        s3 += phidot2[i];
        s4 += fabs(orc_atan2_cr(phi1[i], phi2[i]));
        s5 += sqrt(phi1[i] * phi1[i] + phi2[i] * phi2[i]);
    }
    out[0] = s0; out[1] = s1; out[2] = s2; out[3] = s3; out[4] = s4; out[5] = s5;
}

/* ---------------------------------------------------------------------- */
/* energy (no reference counterpart).  With V = lambda/(4 tau^2) (|phi|^2 -
 * tau^2 + T0^2/3)^2 and the forward-difference gradient energy, the
 * Euler-Lagrange equation of sum_x dx^d [1/2 phidot^2 - 1/2 |grad phi|^2
 * - V] is exactly the step's phi'' = Lap phi - lambda tau^-2 phi (...).   */
void orc_energy(const orc_params *p, double tau,
                const double *phi1, const double *phi2,
                const double *phidot1, const double *phidot2, double out[3])
{
    int N = p->N, nd = p->ndims;
    int64_t L = orc_length(nd, N);
    double dx = (double)p->space_step;
    double vol = (nd == 3) ? dx * dx * dx : dx * dx;
    double lam = (double)p->lambda, T0 = (double)p->T_initial;
    double T03 = (T0 * T0) / 3.0;
    double tau2 = tau * tau;
    double ek = 0, eg = 0, ev = 0;
    #pragma omp parallel for schedule(static) reduction(+ : ek, eg, ev)
    for (int64_t m = 0; m < L; m++) {
        int64_t nb[3];
        if (nd == 2) {
            int64_t i = m % N, j = m / N;
            nb[0] = offset2(i + 1, j, N);
            nb[1] = offset2(i, j + 1, N);
        } else {
            int64_t NN = (int64_t)N * N;
            int64_t i = m / NN, layer = m - i * NN, j = layer / N, k = layer % N;
            nb[0] = offset3(i + 1, j, k, N);
            nb[1] = offset3(i, j + 1, k, N);
            nb[2] = offset3(i, j, k + 1, N);
        }
        double g = 0;
        for (int d = 0; d < nd; d++) {
            double d1 = phi1[nb[d]] - phi1[m];
            double d2 = phi2[nb[d]] - phi2[m];
            g += d1 * d1 + d2 * d2;
        }
        double r2 = phi1[m] * phi1[m] + phi2[m] * phi2[m];
        double s = r2 - tau2 + T03;
        ek += 0.5 * (phidot1[m] * phidot1[m] + phidot2[m] * phidot2[m]);
        eg += 0.5 * g / (dx * dx);
        ev += lam / (4.0 * tau2) * s * s;
    }
    out[0] = ek * vol;
    out[1] = eg * vol;
    out[2] = ev * vol;
}

/* ---------------------------------------------------------------------- */
/* FFT: recursive mixed-radix decimation in time on one line (radices 2, 3,
 * 5, generic), applied along every axis.  Stands in for FFTW 3.3.8's
 * fftw_plan_dft_{2d,3d} (pk_estimation.cpp:66, 144); unnormalised, sign
 * convention e^{sign * 2 pi i jk/n}.                                     */

static int smallest_factor(int n)
{
    if (n % 2 == 0) return 2;
    for (int f = 3; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

static void fft_rec(const double complex *x, double complex *tmp, int n, int stride,
                    double complex *out, const double complex *tw, int ntop)
{
    /* out[0..n) = DFT of x[0], x[stride], ... ; tw[t] = e^{sign 2 pi i t/ntop} */
    if (n == 1) {
        out[0] = x[0];
        return;
    }
    int p = smallest_factor(n);
    int m = n / p;
    int64_t scale = ntop / n;
    /* sub-transforms of the p decimated sequences into tmp */
    for (int r = 0; r < p; r++)
        fft_rec(x + (size_t)r * stride, out, m, stride * p, tmp + (size_t)r * m, tw, ntop);
    /* butterflies: X[k + q m] = sum_r w_n^{r (k + q m)} Y_r[k] */
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            int kk = k + q * m;
            double complex acc = tmp[k];
            for (int r = 1; r < p; r++) {
                int64_t e = ((int64_t)r * kk) % n;
                acc += tmp[(size_t)r * m + k] * tw[e * scale];
            }
            out[kk] = acc;
        }
    }
}

static void fft_line(double complex *line, int n, double complex *w1, double complex *w2,
                     const double complex *tw)
{
    memcpy(w1, line, (size_t)n * sizeof(double complex));
    fft_rec(w1, w2, n, 1, line, tw, n);
}

int orc_fft(int ndims, int N, double *data, int sign)
{
    if ((ndims != 2 && ndims != 3) || N < 1) return -1;
    double complex *d = (double complex *)data;
    int64_t L = orc_length(ndims, N);
    /* axis with stride s: lines start at every index whose coordinate on
     * that axis is zero                                                   */
    int64_t strides[3];
    int naxes = ndims;
    if (ndims == 2) { strides[0] = 1; strides[1] = N; }
    else { strides[0] = 1; strides[1] = N; strides[2] = (int64_t)N * N; }
    double complex *tw = (double complex *)malloc((size_t)N * sizeof(double complex));
    if (!tw) return -1;
    for (int t = 0; t < N; t++) {
        double ang = sign * 2.0 * M_PI * (double)t / (double)N;
        tw[t] = cos(ang) + I * sin(ang);
    }
    for (int a = 0; a < naxes; a++) {
        int64_t s = strides[a];
        int64_t nlines = L / N;
        #pragma omp parallel
        {
            double complex *line = (double complex *)malloc((size_t)N * sizeof(double complex));
            double complex *w1 = (double complex *)malloc((size_t)N * sizeof(double complex));
            double complex *w2 = (double complex *)malloc((size_t)N * sizeof(double complex));
            #pragma omp for schedule(static)
            for (int64_t l = 0; l < nlines; l++) {
                /* enumerate line starts: index with coordinate 0 on axis */
                int64_t lo = l % s, hi = l / s;
                int64_t base = hi * s * N + lo;
                for (int t = 0; t < N; t++) line[t] = d[base + (int64_t)t * s];
                fft_line(line, N, w1, w2, tw);
                for (int t = 0; t < N; t++) d[base + (int64_t)t * s] = line[t];
            }
            free(line); free(w1); free(w2);
        }
    }
    free(tw);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* output_powerspec, pk_estimation.cpp:37-214                              */

void orc_powerspec(int ndims, int N, float dx, const double *data_real,
                   const double *data_imag, double *ks, double *pk, int32_t *count)
{
    float L = N * dx;                                       /* :41 */
    int64_t length = orc_length(ndims, N);
    int n_bins = N / 2;                                     /* :45 */
    for (int b = 0; b < n_bins; b++) { ks[b] = 0; pk[b] = 0; count[b] = 0; }

    double *kv = (double *)malloc((size_t)N * sizeof(double));
    for (int i = 0; i < N; i++) kv[i] = (double)(-N / 2.0f + i);   /* :59, :137 */

    double complex *data = (double complex *)malloc((size_t)length * sizeof(double complex));
    /* copy (:68-72) and shift (:8-31): multiply by (-1)^(i+j[+k]) */
    for (int64_t m = 0; m < length; m++) {
        int64_t par;
        if (ndims == 2) par = (m % N) + (m / N);
        else {
            int64_t NN = (int64_t)N * N, i = m / NN, layer = m - i * NN;
            par = i + layer / N + layer % N;
        }
        double re = data_real[m];
        double im = data_imag ? data_imag[m] : 0.0;
        double sgn = (par & 1) ? -1.0 : 1.0;
        data[m] = (re * sgn) + I * (im * sgn);
    }
    orc_fft(ndims, N, (double *)data, -1);                 /* FFTW_FORWARD */

    /* binning, :88-101 / :166-179, in site order */
    for (int64_t m = 0; m < length; m++) {
        double kmag;
        if (ndims == 2) {
            int64_t i = m % N, j = m / N;
            kmag = sqrt(kv[i] * kv[i] + kv[j] * kv[j]);
        } else {
            int64_t NN = (int64_t)N * N, i = m / NN, layer = m - i * NN;
            int64_t j = layer / N, k = layer % N;
            kmag = sqrt(kv[i] * kv[i] + kv[j] * kv[j] + kv[k] * kv[k]);
        }
        if (kmag == 0.0f || kmag >= N / 2) continue;
        int ps_index = (int)floor(kmag) - 1;
        double re = creal(data[m]), im = cimag(data[m]);
        pk[ps_index] += re * re + im * im;
        ks[ps_index] += kmag;
        count[ps_index]++;
    }

    /* normalisation, :117-121 / :195-199 */
    for (int b = 0; b < n_bins; b++) {
        ks[b] /= count[b];
        if (ndims == 2) pk[b] /= 2.0f * M_PI * ks[b];
        else pk[b] /= 4.0f * M_PI * (ks[b] * ks[b]);
        ks[b] *= 2.0f * M_PI / L;
    }
    free(data);
    free(kv);
}

/* log-spaced bins: squared edges e2[b] = (N/2)^(2b/nbins), b = 0..nbins;
 * a mode with e2[b] <= |k|^2 < e2[b+1] falls in bin b.                   */
void orc_powerspec_log(int ndims, int N, float dx, const double *data_real,
                       const double *data_imag, int nbins, double *ks, double *pk,
                       int32_t *count)
{
    float L = N * dx;
    int64_t length = orc_length(ndims, N);
    double *e2 = (double *)malloc((size_t)(nbins + 1) * sizeof(double));
    for (int b = 0; b <= nbins; b++) e2[b] = pow((double)(N / 2), 2.0 * b / nbins);
    e2[nbins] = (double)(N / 2) * (double)(N / 2);
    for (int b = 0; b < nbins; b++) { ks[b] = 0; pk[b] = 0; count[b] = 0; }

    double complex *data = (double complex *)malloc((size_t)length * sizeof(double complex));
    for (int64_t m = 0; m < length; m++)
        data[m] = data_real[m] + I * (data_imag ? data_imag[m] : 0.0);
    orc_fft(ndims, N, (double *)data, -1);

    for (int64_t m = 0; m < length; m++) {
        int64_t c[3], nc = ndims;
        if (ndims == 2) { c[0] = m % N; c[1] = m / N; }
        else {
            int64_t NN = (int64_t)N * N, i = m / NN, layer = m - i * NN;
            c[0] = i; c[1] = layer / N; c[2] = layer % N;
        }
        int64_t k2 = 0;
        for (int d = 0; d < nc; d++) {
            int64_t f = (c[d] < N / 2) ? c[d] : c[d] - N;     /* wrapped frequency */
            k2 += f * f;
        }
        double kk2 = (double)k2;
        if (kk2 < e2[0] || kk2 >= e2[nbins]) continue;
        int b = 0;
        while (b + 1 < nbins && kk2 >= e2[b + 1]) b++;
        double re = creal(data[m]), im = cimag(data[m]);
        pk[b] += re * re + im * im;
        ks[b] += sqrt(kk2);
        count[b]++;
    }
    for (int b = 0; b < nbins; b++) {
        if (count[b] == 0) { ks[b] = 0; pk[b] = 0; continue; }
        pk[b] /= count[b];
        ks[b] = ks[b] / count[b] * (2.0 * M_PI / L);
    }
    free(data);
    free(e2);
}
