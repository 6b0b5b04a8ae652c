// init.cpp -- initial conditions on the host (reference gaussian_thermal,
// src/common/init.cpp:33-200), plus the k^-1 field of the benchmark
// configurations.  Both fill the four field arrays that the GPU path and
// the CPU oracle then start from.
//
// Normals: the reference draws four arrays from MKL VSL (SFMT19937, inverse
// CDF), which is not available here; each Fourier mode instead takes its
// normals from a Philox4x32-10 counter (Salmon et al., SC'11) keyed by the
// seed and indexed by the mode's unshifted FFT index, with Box-Muller -- the
// same draws as the device-side generator of the library.
//
// Transform: the reference fills a shifted spectrum (index i <-> frequency
// i - N/2), runs FFTW_BACKWARD and multiplies by (-1)^(i+j+k)
// (init.cpp:97-115).  That equals an unshifted backward transform of the
// spectrum placed at its signed frequencies, which is what is done here.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "axs_host.h"
#include "fft.hpp"

namespace {

struct U4 {
    uint32_t x, y, z, w;
};

U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
        c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
               (uint32_t)p0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

void normal_pair(uint64_t seed, uint32_t stream, uint64_t idx, double &g1, double &g2)
{
    U4 r = philox4x32_10(U4{(uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u}, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)r.x << 32) | r.y;
    uint64_t b = ((uint64_t)r.z << 32) | r.w;
    double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;
    double u2 = (double)(b >> 11) * 0x1.0p-53;
    double rad = std::sqrt(-2.0 * std::log(u1));
    g1 = rad * std::cos(2.0 * M_PI * u2);
    g2 = rad * std::sin(2.0 * M_PI * u2);
}

}  // namespace

extern "C" void axs_host_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                                       uint32_t out[4])
{
    U4 r = philox4x32_10(U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

extern "C" int axs_host_init_fields(int mode, int ndims, int N, float space_step, float lambda,
                                    float T_initial, uint64_t seed, double *phi1, double *phi2,
                                    double *phidot1, double *phidot2)
{
    using axs_host::cplx;
    if ((ndims != 2 && ndims != 3) || N < 2 || !phi1 || !phi2 || !phidot1 || !phidot2) return 1;
    const int64_t L = (ndims == 3) ? (int64_t)N * N * N : (int64_t)N * N;
    std::vector<cplx> phik((size_t)L), velk;
    const float Lbox = space_step * N;                                 // init.cpp:36
    const double twopi_L = 2.0 * M_PI / (double)Lbox;
    if (mode == 1) velk.resize((size_t)L);

    // physics.cpp:19: m_eff^2 = lambda * (pow_2(T0) / 3.0f - 1.0f), evaluated
    // in double and stored in a float
    const float meff2 =
        (float)((double)lambda * ((double)T_initial * (double)T_initial / 3.0 - 1.0));
    const double kcut = M_PI / (4.0 * (double)space_step);
    bool any = false;
    for (int64_t h = 0; h < L; ++h) {
        int64_t rem = h, f2 = 0;
        for (int d = 0; d < ndims; ++d) {
            int c = (int)(rem % N);
            rem /= N;
            int f = (c < N / 2) ? c : c - N;
            f2 += (int64_t)f * f;
        }
        double k = std::sqrt((double)f2) * twopi_L;
        double g1, g2;
        normal_pair(seed, 0u, (uint64_t)h, g1, g2);
        if (mode == 0) {
            double amp = (f2 > 0 && k < kcut) ? 1.0 / k : 0.0;
            any = any || amp > 0.0;
            phik[h] = cplx(amp * g1, amp * g2);
        } else {
            // init.cpp:85-100: bose = 1/(exp(w/T0) - 1), amplitudes sqrt(bose/w), sqrt(bose w)
            double amp = 0.0, ampd = 0.0;
            if (f2 > 0) {
                double w = std::sqrt(k * k + (double)meff2);
                double bose = 1.0 / (std::exp(w / (double)T_initial) - 1.0);
                amp = std::sqrt(bose / w);
                ampd = std::sqrt(bose * w);
            }
            phik[h] = cplx(amp * g1, amp * g2);
            double g3, g4;
            normal_pair(seed, 1u, (uint64_t)h, g3, g4);
            velk[h] = cplx(ampd * g3, ampd * g4);
        }
    }
    if (mode == 0 && !any) return 2;    // no mode below pi/(4 dx) on this lattice

    axs_host::fft_nd(phik.data(), ndims, N, +1);
    if (mode == 1) axs_host::fft_nd(velk.data(), ndims, N, +1);

    if (mode == 0) {
        double s = 0.0;
        for (int64_t m = 0; m < L; ++m) s += std::norm(phik[m]);
        double scale = 1.0 / std::sqrt(s / (double)L);              // <|phi|^2> = 1
        for (int64_t m = 0; m < L; ++m) {
            phi1[m] = phik[m].real() * scale;
            phi2[m] = phik[m].imag() * scale;
            phidot1[m] = 0.0;
            phidot2[m] = 0.0;
        }
    } else {
        const double norm = (double)L;                                // init.cpp:112-115
        for (int64_t m = 0; m < L; ++m) {
            phi1[m] = phik[m].real() / norm;
            phi2[m] = phik[m].imag() / norm;
            phidot1[m] = velk[m].real() / norm;
            phidot2[m] = velk[m].imag() / norm;
        }
    }
    return 0;
}

extern "C" int axs_host_fft(double *data, int ndims, int N, int sign)
{
    if (!data || (ndims != 1 && ndims != 2 && ndims != 3) || N < 1) return 1;
    if (ndims == 1)
        axs_host::fft_1d((axs_host::cplx *)data, N, sign);
    else
        axs_host::fft_nd((axs_host::cplx *)data, ndims, N, sign);
    return 0;
}
