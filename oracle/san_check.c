/*
 * san_check.c -- runs every function of the CPU oracle (axs_oracle.c) on
 * small lattices, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle sanitize`): an out-of-bounds
 * access, a use after free, a leak or undefined arithmetic in the oracle
 * stops the run with the sanitizer's report.  TEST INFRASTRUCTURE ONLY.
 * Besides running clean, it checks a few facts that hold exactly: the
 * fused stepper equals the matrix stepper bit for bit, a vortex pair flags
 * sites, and the spectra's mode counts add up to the modes below N/2.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "axs_oracle.h"

static double *alloc(int64_t n)
{
    double *p = (double *)calloc((size_t)n, sizeof(double));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

/* deterministic pseudo-random values in [-1, 1) */
static double rnd(unsigned long long *s)
{
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(*s >> 11) * 0x1p-52 - 1.0;
}

static int failures = 0;

static void expect(int ok, const char *what)
{
    if (!ok) {
        fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

static void check_lattice(int nd, int N, int stencil)
{
    orc_params p = {nd, N, 0.5f, 0.1f, 1.0f, 4.0f, stencil, 1};
    const int64_t L = orc_length(nd, N);
    unsigned long long seed = 1234u + (unsigned)(nd * 100 + N + stencil);
    double *f[8], *g[8];
    for (int a = 0; a < 8; ++a) {
        f[a] = alloc(L);
        g[a] = alloc(L);
    }
    for (int64_t m = 0; m < L; ++m) {
        f[0][m] = rnd(&seed);
        f[1][m] = rnd(&seed);
        f[2][m] = 0.3 * rnd(&seed);
        f[3][m] = 0.3 * rnd(&seed);
    }
    for (int a = 0; a < 8; ++a) memcpy(g[a], f[a], (size_t)L * sizeof(double));

    orc_matrix *A = orc_matrix_create(nd, N, stencil);
    expect(A != NULL, "matrix");
    expect(orc_matrix_nnz(A) > 0, "matrix entries");
    orc_spmv(A, 1.0, f[0], f[6]);
    if (L <= 4096) {
        double *dense = alloc(L * L);
        expect(orc_matrix_dense(A, dense, L * L) == 0, "dense export");
        free(dense);
    }
    double tau = 1.0, tau2 = 1.0;
    orc_vvsl_run(&p, A, 3, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], &tau);
    orc_vvsl_run_fused(&p, 3, g[0], g[1], g[2], g[3], g[4], g[5], &tau2);
    expect(tau == tau2, "tau of the two steppers");
    for (int a = 0; a < 6; ++a)
        expect(memcmp(f[a], g[a], (size_t)L * sizeof(double)) == 0, "fused stepper bits");
    orc_vvsl_hamiltonian_step(&p, A, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], &tau2);
    if (stencil == 0) {
        double t3 = 1.0;
        expect(orc_vvsl_spectral_step(&p, L, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7],
                                      &t3) == 0,
               "spectral step");
        expect(orc_vvsl_run_form(&p, NULL, 0, 1, 2, g[0], g[1], g[2], g[3], g[4], g[5], g[6],
                                 g[7], &t3) == 0,
               "spectral run");
        expect(orc_laplacian_fft(nd, N, 0.5f, f[0], g[6]) == 0, "spectral Laplacian");
    }
    expect(orc_vvsl_run_form(&p, A, 1, 0, 2, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7],
                             &tau2) == 0,
           "Hamiltonian run");
    orc_matrix_destroy(A);

    if (nd == 2 || N <= 32)
        for (int64_t m = 0; m < L; ++m) (void)orc_laplacian2D(f[0], 1, 1, 0.5f, N);
    if (nd == 3) (void)orc_laplacian3D(f[0], 1, 2, 3, 0.5f, N);

    double *ax = alloc(L);
    orc_axion(L, f[0], f[1], ax);
    orc_atan2_cr_array(L, f[0], f[1], g[6]);
    for (int thr = 0; thr <= 30; thr += 15) {
        int64_t n = (nd == 2) ? orc_cores2(ax, N, thr, NULL, 0) : orc_cores3(ax, N, thr, NULL, 0);
        int32_t *coords = (int32_t *)calloc((size_t)(n + 1) * 3, sizeof(int32_t));
        int64_t n2 = (nd == 2) ? orc_cores2(ax, N, thr, coords, n)
                               : orc_cores3(ax, N, thr, coords, n);
        expect(n2 == n, "string count with a list");
        if (nd == 2) (void)orc_num_cores2(coords, n);
        /* a list cut short */
        if (n > 1)
            (void)((nd == 2) ? orc_cores2(ax, N, thr, coords, n / 2)
                             : orc_cores3(ax, N, thr, coords, n / 2));
        (void)((nd == 2) ? orc_xi2(n, tau, N, 0.5f) : orc_xi3(n, tau, N, 0.5f));
        free(coords);
    }
    free(ax);

    double bg[6], en[3];
    orc_background(L, f[0], f[1], f[2], f[3], bg);
    orc_energy(&p, tau, f[0], f[1], f[2], f[3], en);
    expect(isfinite(bg[4]) && isfinite(en[1]), "sums finite");

    const int nb = N / 2;
    double *ks = alloc(nb > 64 ? nb : 64), *pk = alloc(nb > 64 ? nb : 64);
    int32_t *cnt = (int32_t *)calloc((size_t)(nb > 64 ? nb : 64), sizeof(int32_t));
    orc_powerspec(nd, N, 0.5f, f[1], NULL, ks, pk, cnt);
    orc_powerspec(nd, N, 0.5f, f[0], f[1], ks, pk, cnt);
    orc_powerspec_log(nd, N, 0.5f, f[1], NULL, 20, ks, pk, cnt);
    int64_t modes = 0;
    for (int b = 0; b < 20; ++b) modes += cnt[b];
    int64_t inside = 0;                   /* modes with 1 <= |k|^2 < (N/2)^2 */
    const int64_t top = (int64_t)(N / 2) * (N / 2);
    for (int64_t m = 0; m < L; ++m) {
        int64_t c[3] = {m % N, (m / N) % N, m / ((int64_t)N * N)}, k2 = 0;
        for (int d = 0; d < nd; ++d) {
            const int64_t f_ = c[d] < N / 2 ? c[d] : c[d] - N;
            k2 += f_ * f_;
        }
        inside += (k2 >= 1 && k2 < top);
    }
    expect(modes == inside, "log-bin mode counts");
    free(ks);
    free(pk);
    free(cnt);

    double *cplx = alloc(2 * L);
    for (int64_t m = 0; m < L; ++m) cplx[2 * m] = f[0][m];
    expect(orc_fft(nd, N, cplx, -1) == 0 && orc_fft(nd, N, cplx, +1) == 0, "FFT round trip");
    for (int64_t m = 0; m < L; ++m)
        expect(fabs(cplx[2 * m] / (double)L - f[0][m]) < 1e-12, "FFT inverse");
    free(cplx);

    for (int a = 0; a < 8; ++a) {
        free(f[a]);
        free(g[a]);
    }
}

int main(void)
{
    for (int s = 0; s <= 3; ++s) {
        check_lattice(2, 16, s);
        check_lattice(3, 8, s);
    }
    check_lattice(2, 24, 0);
    check_lattice(3, 12, 1);
    (void)orc_atan2_cr(0.0, -0.0);
    (void)orc_atan2_cr(NAN, 1.0);
    (void)orc_atan2_cr(INFINITY, -INFINITY);
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("oracle sanitizer run: ok\n");
    return 0;
}
