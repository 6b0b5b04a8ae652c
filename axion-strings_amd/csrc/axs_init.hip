// axs_init.hip -- initial conditions generated in device memory.
//
// Two spectra, both drawn mode by mode from a counter-based generator
// (Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) with Box-Muller normals, then inverse-transformed
// by rocFFT:
//
//  AXS_INIT_KINV     complex Gaussian random field, independent normal real
//                    and imaginary Fourier amplitudes times 1/|k| for
//                    0 < |k| < pi/(4 dx), rescaled to <|phi|^2> = 1;
//                    phidot = 0, ker = 0.
//  AXS_INIT_THERMAL  the thermal spectrum of gaussian_thermal (reference
//                    src/common/init.cpp:33-200): amplitudes sqrt(n_B/w) for
//                    phi and sqrt(n_B w) for phidot with w = sqrt(k^2 +
//                    m_eff^2), n_B = 1/(exp(w/T0) - 1), zero at k = 0, and
//                    the 1/N^d normalisation of init.cpp:112-115; ker = 0.
//
// Fields are produced in FP64 and rounded once for the FP32 form, so an FP64
// and an FP32 context initialised with the same seed hold the same field.
#include <rocfft/rocfft.h>

#include <math.h>

#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

// ---- Philox4x32-10 ------------------------------------------------------

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ void mulhilo(uint32_t a, uint32_t b, uint32_t &hi,
                                                 uint32_t &lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

__host__ __device__ __forceinline__ U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(M0, ctr.x, hi0, lo0);
        mulhilo(M1, ctr.z, hi1, lo1);
        ctr = U4{hi1 ^ ctr.y ^ k0, lo1, hi0 ^ ctr.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// two independent standard normals for (seed, stream, index)
__device__ __forceinline__ void normal_pair(uint64_t seed, uint32_t stream, uint64_t idx,
                                            double &g1, double &g2)
{
    U4 r = philox4x32_10(U4{(uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u},
                         (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)r.x << 32) | r.y;
    uint64_t b = ((uint64_t)r.z << 32) | r.w;
    double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;   // (0, 1]
    double u2 = (double)(b >> 11) * 0x1.0p-53;           // [0, 1)
    double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    g1 = rad * c;
    g2 = rad * s;
}

struct SpecParams {
    int nd, N;
    double twopi_over_L;   // 2 pi / (N dx)
    double kcut;           // pi / (4 dx)
    double meff2;          // lambda (T0^2/3 - 1), float as physics.cpp:19
    double T0;
    uint64_t seed;
    int mode;
};

// fills spectra for one unshifted mode index h (FFT order)
__global__ void __launch_bounds__(256)
fill_spectrum_kernel(uint32_t L, SpecParams sp, double2 *__restrict__ phik,
                     double2 *__restrict__ velk)
{
    uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= L) return;
    int N = sp.N;
    int64_t f2 = 0;
    uint32_t rem = h;
    for (int d = 0; d < sp.nd; ++d) {
        int c = (int)(rem % (uint32_t)N);
        rem /= (uint32_t)N;
        int f = (c < N / 2) ? c : c - N;
        f2 += (int64_t)f * f;
    }
    double k = sqrt((double)f2) * sp.twopi_over_L;
    double g1, g2;
    normal_pair(sp.seed, 0u, h, g1, g2);
    if (sp.mode == AXS_INIT_KINV) {
        double amp = (f2 > 0 && k < sp.kcut) ? 1.0 / k : 0.0;
        phik[h] = make_double2(amp * g1, amp * g2);
    } else {
        double amp = 0.0, ampd = 0.0;
        if (f2 > 0) {
            double w = sqrt(k * k + sp.meff2);
            double bose = 1.0 / (exp(w / sp.T0) - 1.0);
            amp = sqrt(bose / w);
            ampd = sqrt(bose * w);
        }
        phik[h] = make_double2(amp * g1, amp * g2);
        double g3, g4;
        normal_pair(sp.seed, 1u, h, g3, g4);
        velk[h] = make_double2(ampd * g3, ampd * g4);
    }
}

struct KahanI {
    double s, c;
};

__global__ void __launch_bounds__(256)
norm2_kernel(uint32_t L, const double2 *__restrict__ z, KahanI *__restrict__ part)
{
    __shared__ KahanI sh[4];
    KahanI k{0.0, 0.0};
    for (uint32_t m = blockIdx.x * 256 + threadIdx.x; m < L; m += gridDim.x * 256) {
        double2 v = z[m];
        double x = v.x * v.x + v.y * v.y;
        double t = k.s + x;
        k.c += (fabs(k.s) >= fabs(x)) ? (k.s - t) + x : (x - t) + k.s;
        k.s = t;
    }
    for (int o = 32; o > 0; o >>= 1) {
        double s2 = __shfl_down(k.s, o, 64), c2 = __shfl_down(k.c, o, 64);
        double t = k.s + s2;
        double bp = t - k.s;
        double e = (k.s - (t - bp)) + (s2 - bp);
        k = KahanI{t, k.c + c2 + e};
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        KahanI a = sh[0];
        for (int w = 1; w < 4; ++w) {
            double t = a.s + sh[w].s;
            double bp = t - a.s;
            double e = (a.s - (t - bp)) + (sh[w].s - bp);
            a = KahanI{t, a.c + sh[w].c + e};
        }
        part[blockIdx.x] = a;
    }
}

// scaled inverse transforms into the state layout; ker = 0 (and, for the
// FP32 form, u = phidot since ker = 0)
__global__ void __launch_bounds__(256)
store_init64_kernel(uint32_t L, const double2 *__restrict__ z, const double2 *__restrict__ zv,
                    double scale, dvec2 *__restrict__ phi, dvec2 *__restrict__ vel,
                    dvec2 *__restrict__ ker)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    double2 a = z[m];
    dvec2 p, v, k;
    p.x = a.x * scale;
    p.y = a.y * scale;
    v.x = 0.0;
    v.y = 0.0;
    if (zv) {
        double2 b = zv[m];
        v.x = b.x * scale;
        v.y = b.y * scale;
    }
    k.x = 0.0;
    k.y = 0.0;
    phi[m] = p;
    vel[m] = v;
    ker[m] = k;
}

__global__ void __launch_bounds__(256)
store_init32_kernel(uint32_t L, const double2 *__restrict__ z, const double2 *__restrict__ zv,
                    double scale, fvec4 *__restrict__ st, fvec2 *__restrict__ ker)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    double2 a = z[m];
    fvec4 r;
    r.x = (float)(a.x * scale);
    r.y = (float)(a.y * scale);
    r.z = 0.0f;
    r.w = 0.0f;
    if (zv) {
        double2 b = zv[m];
        r.z = (float)(b.x * scale);
        r.w = (float)(b.y * scale);
    }
    fvec2 k;
    k.x = 0.0f;
    k.y = 0.0f;
    st[m] = r;
    ker[m] = k;
}

// in-place complex inverse transform; plans and work buffers are cached per
// side in the context (ensembles initialise many members of a few sides)
struct InvPlan {
    int N = 0, nd = 0;
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    void *work = nullptr;
};

void destroy_inverse_plans(Context *c)
{
    for (void *v : c->inv_plans) {
        InvPlan *p = (InvPlan *)v;
        if (p->plan) rocfft_plan_destroy(p->plan);
        if (p->info) rocfft_execution_info_destroy(p->info);
        if (p->work) (void)hipFree(p->work);
        delete p;
    }
    c->inv_plans.clear();
}

static int inverse_fft(Context *c, int N, double2 *buf)
{
    InvPlan *P = nullptr;
    for (void *v : c->inv_plans)
        if (((InvPlan *)v)->N == N && ((InvPlan *)v)->nd == c->nd) P = (InvPlan *)v;
    if (!P) {
        P = new InvPlan();
        P->N = N;
        P->nd = c->nd;
        size_t lengths[3] = {(size_t)N, (size_t)N, (size_t)N};
        size_t wb = 0;
        bool ok = rocfft_setup() == rocfft_status_success &&
                  rocfft_plan_create(&P->plan, rocfft_placement_inplace,
                                     rocfft_transform_type_complex_inverse,
                                     rocfft_precision_double, (size_t)c->nd, lengths, 1,
                                     nullptr) == rocfft_status_success &&
                  rocfft_plan_get_work_buffer_size(P->plan, &wb) == rocfft_status_success &&
                  rocfft_execution_info_create(&P->info) == rocfft_status_success;
        if (ok && wb) {
            ok = hipMalloc(&P->work, wb) == hipSuccess &&
                 rocfft_execution_info_set_work_buffer(P->info, P->work, wb) ==
                     rocfft_status_success;
        }
        c->inv_plans.push_back(P);
        if (!ok) {
            set_error("rocfft inverse plan failed (N=%d)", N);
            return AXS_ERR_FFT;
        }
    }
    if (rocfft_execution_info_set_stream(P->info, c->stream) != rocfft_status_success) {
        set_error("rocfft_execution_info_set_stream failed");
        return AXS_ERR_FFT;
    }
    void *ib[1] = {buf};
    if (rocfft_execute(P->plan, ib, nullptr, P->info) != rocfft_status_success) {
        set_error("rocfft inverse transform failed (N=%d)", N);
        return AXS_ERR_FFT;
    }
    return AXS_OK;
}

int run_init(Context *c, Member &m, int mode, uint64_t seed)
{
    uint32_t L = (uint32_t)m.L;
    size_t zb = (size_t)L * sizeof(double2);
    double2 *phik = nullptr, *velk = nullptr;
    AXS_HIP_TRY(hipMalloc(&phik, zb));
    struct Guard {
        double2 **a, **b;
        ~Guard()
        {
            if (*a) (void)hipFree(*a);
            if (*b) (void)hipFree(*b);
        }
    } guard{&phik, &velk};
    if (mode == AXS_INIT_THERMAL) AXS_HIP_TRY(hipMalloc(&velk, zb));

    double dx = (double)c->p.space_step;
    float T0f = c->p.T_initial;
    SpecParams sp;
    sp.nd = c->nd;
    sp.N = m.N;
    sp.twopi_over_L = 2.0 * M_PI / ((double)m.N * dx);
    sp.kcut = M_PI / (4.0 * dx);
    // physics.cpp:19: lambda * (pow_2(T0) / 3.0f - 1.0f), double, stored as float
    sp.meff2 = (double)(float)((double)c->p.lambda * ((double)T0f * (double)T0f / 3.0 - 1.0));
    sp.T0 = (double)T0f;
    sp.seed = seed;
    sp.mode = mode;
    unsigned g = (L + 255) / 256;
    hipLaunchKernelGGL(fill_spectrum_kernel, dim3(g), dim3(256), 0, c->stream, L, sp, phik, velk);
    AXS_HIP_TRY(hipGetLastError());
    AXS_TRY(inverse_fft(c, m.N, phik));
    if (velk) AXS_TRY(inverse_fft(c, m.N, velk));

    double scale;
    if (mode == AXS_INIT_KINV) {
        const int nb = 1024;
        AXS_TRY(ensure_scratch(c, nb * sizeof(KahanI)));
        hipLaunchKernelGGL(norm2_kernel, dim3(nb), dim3(256), 0, c->stream, L, phik,
                           (KahanI *)c->scratch);
        AXS_HIP_TRY(hipGetLastError());
        std::vector<KahanI> part(nb);
        AXS_HIP_TRY(hipMemcpyAsync(part.data(), c->scratch, nb * sizeof(KahanI),
                                   hipMemcpyDeviceToHost, c->stream));
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
        double s = 0.0, comp = 0.0;
        for (const KahanI &k : part) {
            double t = s + k.s;
            comp += (fabs(s) >= fabs(k.s)) ? (s - t) + k.s : (k.s - t) + s;
            comp += k.c;
            s = t;
        }
        double mean = (s + comp) / (double)L;
        if (!(mean > 0.0)) {
            set_error("k^-1 spectrum has no mode below pi/(4 dx) for N=%d, dx=%g", m.N, dx);
            return AXS_ERR_INVALID_ARG;
        }
        scale = 1.0 / sqrt(mean);
    } else {
        scale = 1.0 / (double)L;                                      // init.cpp:112-115
    }

    int h = c->cur;
    if (c->prec == AXS_FP64) {
        hipLaunchKernelGGL(store_init64_kernel, dim3(g), dim3(256), 0, c->stream, L, phik, velk,
                           scale, (dvec2 *)m.state[h][PHI], (dvec2 *)m.state[h][VEL],
                           (dvec2 *)m.state[h][KER]);
    } else {
        hipLaunchKernelGGL(store_init32_kernel, dim3(g), dim3(256), 0, c->stream, L, phik, velk,
                           scale, (fvec4 *)m.state[h][PHI], (fvec2 *)m.aux_ker);
        m.ker_valid = true;      // zero, as after an upload without ker
        m.ker_uploaded = true;
    }
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    m.has_state = true;
    m.vel_valid = false;
    m.version++;
    return AXS_OK;
}

}  // namespace axs
