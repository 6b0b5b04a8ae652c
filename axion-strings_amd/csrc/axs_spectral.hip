// axs_spectral.hip -- the spectral Laplacian (reference compute_laplacian_fft,
// src/standard/s_integrate.cpp:179-296) and the field-rescaled step that uses
// it in place of the sparse matrix (the #else branch at :326-328), plus the
// Laplacian of the current field for either operator.
//
// The reference multiplies the field by (-1)^(i+j+k), takes a complex
// forward FFT, multiplies array index i by -k[i]^2 with
// k[i] = (-N/2 + i) 2 pi / L, transforms back, undoes the sign pattern and
// divides by the length: the shifts only relabel modes, so the operator is
// -k_f^2 on every frequency f in [-N/2, N/2).  Here the real field goes
// through a rocFFT real-to-complex transform (half spectrum), each stored
// mode is scaled by -k_f^2 with k_f taken from the same table as the
// reference (index f + N/2), and a complex-to-real transform brings it back;
// the result is divided by the length as there.  FFT rounding differs from
// FFTW's, so this path agrees with the oracle to rounding, not bit for bit.
#include <rocfft/rocfft.h>

#include <math.h>

#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

struct SpecPlan {
    int N = 0, nd = 0;
    rocfft_plan fwd = nullptr, bwd = nullptr;
    rocfft_execution_info fwd_info = nullptr, bwd_info = nullptr;
    void *work = nullptr;
    size_t work_bytes = 0;
    double *real = nullptr;     // one real component (transform input/output)
    double *spec = nullptr;     // half spectrum, interleaved complex
    double *lap = nullptr;      // two Laplacian components, L doubles each
    double *ktab = nullptr;     // k[i], i = 0..N-1 (s_integrate.cpp:192, :242)
    ~SpecPlan()
    {
        if (fwd) rocfft_plan_destroy(fwd);
        if (bwd) rocfft_plan_destroy(bwd);
        if (fwd_info) rocfft_execution_info_destroy(fwd_info);
        if (bwd_info) rocfft_execution_info_destroy(bwd_info);
        for (void *p : {(void *)work, (void *)real, (void *)spec, (void *)lap, (void *)ktab})
            if (p) (void)hipFree(p);
    }
};

static int spec_rocfft_ready = 0;

static int get_spec_plan(Context *c, int N, SpecPlan **out)
{
    for (void *q : c->spec_plans) {
        SpecPlan *p = (SpecPlan *)q;
        if (p->N == N && p->nd == c->nd) {
            *out = p;
            return AXS_OK;
        }
    }
    if (!spec_rocfft_ready) {
        if (rocfft_setup() != rocfft_status_success) {
            set_error("rocfft_setup failed");
            return AXS_ERR_FFT;
        }
        spec_rocfft_ready = 1;
    }
    SpecPlan *p = new SpecPlan();
    p->N = N;
    p->nd = c->nd;
    const size_t lengths[3] = {(size_t)N, (size_t)N, (size_t)N};   // fastest first
    const size_t L = (c->nd == 3) ? (size_t)N * N * N : (size_t)N * N;
    const size_t half = L / N * (N / 2 + 1);
    size_t wf = 0, wb = 0;
    if (rocfft_plan_create(&p->fwd, rocfft_placement_notinplace,
                           rocfft_transform_type_real_forward, rocfft_precision_double,
                           (size_t)c->nd, lengths, 1, nullptr) != rocfft_status_success ||
        rocfft_plan_create(&p->bwd, rocfft_placement_notinplace,
                           rocfft_transform_type_real_inverse, rocfft_precision_double,
                           (size_t)c->nd, lengths, 1, nullptr) != rocfft_status_success ||
        rocfft_plan_get_work_buffer_size(p->fwd, &wf) != rocfft_status_success ||
        rocfft_plan_get_work_buffer_size(p->bwd, &wb) != rocfft_status_success ||
        rocfft_execution_info_create(&p->fwd_info) != rocfft_status_success ||
        rocfft_execution_info_create(&p->bwd_info) != rocfft_status_success) {
        delete p;
        set_error("rocFFT plans for the spectral Laplacian failed (N=%d ndims=%d)", N, c->nd);
        return AXS_ERR_FFT;
    }
    p->work_bytes = wf > wb ? wf : wb;
    hipError_t e = hipMalloc(&p->real, L * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&p->spec, half * 2 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&p->lap, 2 * L * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&p->ktab, (size_t)N * sizeof(double));
    if (e == hipSuccess && p->work_bytes) e = hipMalloc(&p->work, p->work_bytes);
    if (e != hipSuccess) {
        delete p;
        set_error("spectral Laplacian buffers: %s", hipGetErrorString(e));
        return AXS_ERR_OUT_OF_MEMORY;
    }
    if (p->work_bytes &&
        (rocfft_execution_info_set_work_buffer(p->fwd_info, p->work, p->work_bytes) !=
             rocfft_status_success ||
         rocfft_execution_info_set_work_buffer(p->bwd_info, p->work, p->work_bytes) !=
             rocfft_status_success)) {
        delete p;
        set_error("rocfft_execution_info_set_work_buffer failed");
        return AXS_ERR_FFT;
    }
    // k[i] = (- N / 2.0f + i) * 2.0f * M_PI / L, L = N * dx in float (:183, :192)
    std::vector<double> k(N);
    const float Lf = N * c->p.space_step;
    for (int i = 0; i < N; ++i) k[i] = (double)((-N / 2.0f + i) * 2.0f) * M_PI / (double)Lf;
    e = hipMemcpy(p->ktab, k.data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete p;
        set_error("k table upload: %s", hipGetErrorString(e));
        return AXS_ERR_HIP;
    }
    c->spec_plans.push_back(p);
    *out = p;
    return AXS_OK;
}

void destroy_spec_plans(Context *c)
{
    for (void *q : c->spec_plans) delete (SpecPlan *)q;
    c->spec_plans.clear();
}

// one component of the field, widened to double
template <class P>
__global__ void __launch_bounds__(256)
component_kernel(uint32_t L, P phi, int comp, double *__restrict__ y)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) {
        double a, b;
        phi.get(m, a, b);
        y[m] = comp ? b : a;
    }
}

// half spectrum times -k_f^2; the halved axis is the fastest one (3D k,
// 2D i), its frequency is the index itself (0..N/2); a full axis holds f at
// index f (f >= 0) or f + N (f < 0).  The reference's table index is f + N/2.
template <int ND>
__global__ void __launch_bounds__(256)
minus_k2_kernel(int N, uint32_t nhalf, const double *__restrict__ k, double *__restrict__ spec)
{
    uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhalf) return;
    const uint32_t H = (uint32_t)(N / 2 + 1);
    const int fz = (int)(h % H);                    // fastest axis
    const uint32_t rest = h / H;
    auto tab = [&](int n) -> double {               // k of stored index n on a full axis
        int f = (n < N / 2) ? n : n - N;
        return k[f + N / 2];
    };
    const double kz = k[fz + N / 2 - ((fz == N / 2) ? N : 0)];
    double k_sq;
    if (ND == 2) {
        // m = i + N*j: i is the halved axis, j the full one; kx[i]^2 + ky[j]^2
        const double kj = tab((int)rest);
        k_sq = kz * kz + kj * kj;
    } else {
        // m = (i*N + j)*N + k: k halved; kx[i]^2 + ky[j]^2 + kz[k]^2
        const double ki = tab((int)(rest / (uint32_t)N));
        const double kj = tab((int)(rest % (uint32_t)N));
        k_sq = ki * ki + kj * kj + kz * kz;
    }
    const double s = -k_sq;
    spec[2 * h] = spec[2 * h] * s;
    spec[2 * h + 1] = spec[2 * h + 1] * s;
}

// ret = data / length (:230-231)
__global__ void __launch_bounds__(256)
scale_kernel(uint32_t L, const double *__restrict__ x, double len, double *__restrict__ y)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) y[m] = x[m] / len;
}

// spectral Laplacian of comp of the field into lap (device, L doubles)
template <class P>
static int spectral_component(Context *c, SpecPlan *sp, uint32_t L, const P &phi, int comp,
                              double *lap)
{
    const unsigned grid = (L + 255) / 256;
    hipLaunchKernelGGL(component_kernel<P>, dim3(grid), dim3(256), 0, c->stream, L, phi, comp,
                       sp->real);
    AXS_HIP_TRY(hipGetLastError());
    if (rocfft_execution_info_set_stream(sp->fwd_info, c->stream) != rocfft_status_success ||
        rocfft_execution_info_set_stream(sp->bwd_info, c->stream) != rocfft_status_success) {
        set_error("rocfft_execution_info_set_stream failed");
        return AXS_ERR_FFT;
    }
    void *in[1] = {sp->real}, *out[1] = {sp->spec};
    if (rocfft_execute(sp->fwd, in, out, sp->fwd_info) != rocfft_status_success) {
        set_error("rocfft forward transform failed");
        return AXS_ERR_FFT;
    }
    const uint32_t nhalf = L / (uint32_t)sp->N * (uint32_t)(sp->N / 2 + 1);
    auto kk = (c->nd == 2) ? minus_k2_kernel<2> : minus_k2_kernel<3>;
    hipLaunchKernelGGL(kk, dim3((nhalf + 255) / 256), dim3(256), 0, c->stream, sp->N, nhalf,
                       sp->ktab, sp->spec);
    AXS_HIP_TRY(hipGetLastError());
    void *bin[1] = {sp->spec}, *bout[1] = {sp->real};
    if (rocfft_execute(sp->bwd, bin, bout, sp->bwd_info) != rocfft_status_success) {
        set_error("rocfft inverse transform failed");
        return AXS_ERR_FFT;
    }
    hipLaunchKernelGGL(scale_kernel, dim3(grid), dim3(256), 0, c->stream, L, sp->real, (double)L,
                       lap);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

// drift of the field-rescaled form (s_integrate.cpp:313-314) into out.phi
__global__ void __launch_bounds__(256)
spectral_drift_kernel(uint32_t L, const dvec2 *__restrict__ phi, const dvec2 *__restrict__ vel,
                      const dvec2 *__restrict__ ker, dvec2 *__restrict__ xo, double dt)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    dvec2 p = phi[m], v = vel[m], k = ker[m], x;
    x.x = p.x + dt * (v.x + 0.5 * k.x * dt);
    x.y = p.y + dt * (v.y + 0.5 * k.y * dt);
    xo[m] = x;
}

// ker_next = Laplacian, then the potential (:331-335), kick (:337-341)
__global__ void __launch_bounds__(256)
spectral_finish_kernel(uint32_t L, const dvec2 *__restrict__ x, const double *__restrict__ lap1,
                       const double *__restrict__ lap2, const dvec2 *__restrict__ vel,
                       const dvec2 *__restrict__ ker, dvec2 *__restrict__ vo,
                       dvec2 *__restrict__ ko, StepScalars s)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    dvec2 xv = x[m];
    double S = xv.x * xv.x + xv.y * xv.y - s.tau2 + s.T03;
    dvec2 kn, vn, kc = ker[m], vc = vel[m];
    kn.x = lap1[m] + s.cl * xv.x * S;
    kn.y = lap2[m] + s.cl * xv.y * S;
    vn.x = vc.x + 0.5 * (kc.x + kn.x) * s.dt;
    vn.y = vc.y + 0.5 * (kc.y + kn.y) * s.dt;
    vo[m] = vn;
    ko[m] = kn;
}

int run_spectral_step(Context *c, Member &m, int src, int dst, const StepScalars &s)
{
    SpecPlan *sp = nullptr;
    AXS_TRY(get_spec_plan(c, m.N, &sp));
    const uint32_t L = (uint32_t)m.L;
    const unsigned grid = (L + 255) / 256;
    const dvec2 *phi = (const dvec2 *)m.state[src][PHI], *vel = (const dvec2 *)m.state[src][VEL],
                *ker = (const dvec2 *)m.state[src][KER];
    dvec2 *xo = (dvec2 *)m.state[dst][PHI];
    hipLaunchKernelGGL(spectral_drift_kernel, dim3(grid), dim3(256), 0, c->stream, L, phi, vel,
                       ker, xo, s.dt);
    AXS_HIP_TRY(hipGetLastError());
    PhiF64 x{xo};
    AXS_TRY(spectral_component(c, sp, L, x, 0, sp->lap));
    AXS_TRY(spectral_component(c, sp, L, x, 1, sp->lap + L));
    hipLaunchKernelGGL(spectral_finish_kernel, dim3(grid), dim3(256), 0, c->stream, L, xo,
                       sp->lap, sp->lap + L, vel, ker, (dvec2 *)m.state[dst][VEL],
                       (dvec2 *)m.state[dst][KER], s);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

// alpha * (stencil sum in COO order) of the current field (s_integrate.cpp:322-324)
template <int ND, class P>
__global__ void __launch_bounds__(256)
stencil_lap_kernel(int N, uint32_t L, P phi, StepScalars s, double *__restrict__ l1,
                   double *__restrict__ l2)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    Site<ND> site = decompose<ND>(m, N);
    double a, b;
    phi.get(m, a, b);
    double s1 = s.c[0] * a, s2 = s.c[0] * b;
    for (int l = 1; l < s.ncoef; ++l) {
        const double cl = s.c[l];
#pragma unroll
        for (int ax = 0; ax < ND; ++ax) {
            phi.get(shifted<ND>(site, ax, l, N), a, b);
            s1 = s1 + cl * a;
            s2 = s2 + cl * b;
            phi.get(shifted<ND>(site, ax, -l, N), a, b);
            s1 = s1 + cl * a;
            s2 = s2 + cl * b;
        }
    }
    l1[m] = s.alpha * s1;
    l2[m] = s.alpha * s2;
}

int run_laplacian(Context *c, Member &m, int mode, double *h1, double *h2)
{
    FieldView fv;
    AXS_TRY(current_fields(c, m, &fv));
    SpecPlan *sp = nullptr;
    AXS_TRY(get_spec_plan(c, m.N, &sp));     // its buffers hold the result either way
    const uint32_t L = (uint32_t)m.L;
    if (mode == AXS_LAPLACIAN_SPECTRAL) {
        if (c->prec == AXS_FP64) {
            PhiF64 x{(const dvec2 *)fv.phi};
            AXS_TRY(spectral_component(c, sp, L, x, 0, sp->lap));
            AXS_TRY(spectral_component(c, sp, L, x, 1, sp->lap + L));
        } else {
            PhiF32 x{(const fvec4 *)fv.phi};
            AXS_TRY(spectral_component(c, sp, L, x, 0, sp->lap));
            AXS_TRY(spectral_component(c, sp, L, x, 1, sp->lap + L));
        }
    } else {
        // the field-rescaled stencil operator: alpha = 1.0f / pow_2(dx)
        axs_params q = c->p;
        q.eom = AXS_EOM_FIELD_RESCALED;
        StepScalars s = make_scalars(q, c->tau);
        const unsigned grid = (L + 255) / 256;
        if (c->prec == AXS_FP64) {
            auto k = (c->nd == 2) ? stencil_lap_kernel<2, PhiF64> : stencil_lap_kernel<3, PhiF64>;
            hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, c->stream, m.N, L,
                               PhiF64{(const dvec2 *)fv.phi}, s, sp->lap, sp->lap + L);
        } else {
            auto k = (c->nd == 2) ? stencil_lap_kernel<2, PhiF32> : stencil_lap_kernel<3, PhiF32>;
            hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, c->stream, m.N, L,
                               PhiF32{(const fvec4 *)fv.phi}, s, sp->lap, sp->lap + L);
        }
        AXS_HIP_TRY(hipGetLastError());
    }
    if (h1)
        AXS_HIP_TRY(hipMemcpyAsync(h1, sp->lap, L * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
    if (h2)
        AXS_HIP_TRY(hipMemcpyAsync(h2, sp->lap + L, L * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    return AXS_OK;
}

}  // namespace axs
