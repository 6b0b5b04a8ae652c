// axs_step.hip -- the velocity-Verlet step of vvsl_field_rescaled
// (reference src/standard/s_integrate.cpp:305-348) as fused gfx950 kernels:
// the general one-site-per-work-item form used for 2D lattices, higher-order
// stencils and 3D sides that are not multiples of 64 (the 2.5D streaming
// kernel of axs_march.hip covers the rest).
//
// The reference makes five sweeps per step (drift, two sparse-matrix
// Laplacians, potential, kick, copy).  Here one kernel does the whole step:
// each work-item recomputes the drift of the 2*ND*r neighbours it needs
// (bit-identical to its owner's drift, since the expression is the same),
// applies the stencil directly (no matrix), adds the potential force, kicks,
// and writes the new state to the other half of a ping-pong pair, so the
// reference's ker copy becomes a pointer swap.
//
// FP64 form: every expression keeps the reference's tree and the library is
// built with -ffp-contract=off, so each result is the correctly rounded value
// of the same operation sequence as the CPU path -- bit-identical fields.
// FP32 form: phi and u = phidot + dt/2 ker are stored in FP32; one step is
//   phi' = phi + dt u;  ker' = F(phi', tau');  u' = u + dt ker'
// which is the same velocity-Verlet map written in half-step velocities.
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

struct In64 {
    const dvec2 *__restrict__ phi, *__restrict__ vel, *__restrict__ ker;
};
struct Out64 {
    dvec2 *__restrict__ phi, *__restrict__ vel, *__restrict__ ker;
};

// drift, s_integrate.cpp:313-314: phi += dt * (phidot + 0.5f * ker * dt);
// Hamiltonian form :368-369: phi += dt / pow_2(tau) * (pi + 0.5f * dt * ker)
template <bool HAM>
__device__ __forceinline__ dvec2 drift64(const In64 &a, uint32_t q, const StepScalars &s)
{
    dvec2 p = a.phi[q], v = a.vel[q], k = a.ker[q];
    dvec2 x;
    if (HAM) {
        x.x = p.x + s.dcoef * (v.x + s.hdt * k.x);
        x.y = p.y + s.dcoef * (v.y + s.hdt * k.y);
    } else {
        x.x = p.x + s.dt * (v.x + 0.5 * k.x * s.dt);
        x.y = p.y + s.dt * (v.y + 0.5 * k.y * s.dt);
    }
    return x;
}

// The Hamiltonian form differs only in the drift and in the scalars (its
// Laplacian, potential and kick have the same expression trees with
// alpha = tau^2/dx^2, cl = -tau^2 lambda, tau2 = 1, T03 = T0^2/(3 tau^2)).
template <int ND, bool HAM>
__device__ __forceinline__ void site_fp64(int N, uint32_t m, const In64 &in, const Out64 &out,
                                          const StepScalars &s)
{
    Site<ND> site = decompose<ND>(m, N);

    dvec2 x = drift64<HAM>(in, m, s);

    // Laplacian in COO row order (centre first), then alpha (s_integrate.cpp:322-324)
    double s1 = s.c[0] * x.x, s2 = s.c[0] * x.y;
    for (int l = 1; l < s.ncoef; ++l) {
        double cl = s.c[l];
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            dvec2 y = drift64<HAM>(in, shifted<ND>(site, a, l, N), s);
            s1 = s1 + cl * y.x;
            s2 = s2 + cl * y.y;
            y = drift64<HAM>(in, shifted<ND>(site, a, -l, N), s);
            s1 = s1 + cl * y.x;
            s2 = s2 + cl * y.y;
        }
    }
    double lap1 = s.alpha * s1, lap2 = s.alpha * s2;

    // potential, s_integrate.cpp:333-334
    double S = x.x * x.x + x.y * x.y - s.tau2 + s.T03;
    dvec2 kn;
    kn.x = lap1 + s.cl * x.x * S;
    kn.y = lap2 + s.cl * x.y * S;

    // kick, s_integrate.cpp:339-340
    dvec2 kc = in.ker[m], vc = in.vel[m];
    dvec2 vn;
    vn.x = vc.x + 0.5 * (kc.x + kn.x) * s.dt;
    vn.y = vc.y + 0.5 * (kc.y + kn.y) * s.dt;
    out.phi[m] = x;
    out.vel[m] = vn;
    out.ker[m] = kn;
}

template <int ND, bool HAM>
__global__ void __launch_bounds__(256)
step_fp64_kernel(int N, uint32_t L, In64 in, Out64 out, StepScalars s)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) site_fp64<ND, HAM>(N, m, in, out, s);
}

template <int ND>
__device__ __forceinline__ void site_fp32(int N, uint32_t m, const fvec4 *__restrict__ in,
                                          fvec4 *__restrict__ out, const StepScalarsF &s)
{
    Site<ND> site = decompose<ND>(m, N);

    fvec4 c = in[m];
    float x1 = F32Math::drift(c.x, c.z, s.dt);
    float x2 = F32Math::drift(c.y, c.w, s.dt);
    float s1 = -0.0f, s2 = -0.0f;     // the centre term joins the potential
    for (int l = 1; l < s.ncoef; ++l) {
        float cl = s.c[l];
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            fvec4 p = in[shifted<ND>(site, a, l, N)];
            fvec4 q = in[shifted<ND>(site, a, -l, N)];
            s1 = F32Math::acc(s1, cl, F32Math::drift(p.x, p.z, s.dt));
            s2 = F32Math::acc(s2, cl, F32Math::drift(p.y, p.w, s.dt));
            s1 = F32Math::acc(s1, cl, F32Math::drift(q.x, q.z, s.dt));
            s2 = F32Math::acc(s2, cl, F32Math::drift(q.y, q.w, s.dt));
        }
    }
    float t = F32Math::own(s.cl, F32Math::pot(x1, x2, s.pot0), s.ac0);
    float k1 = F32Math::force(s.alpha, s1, x1, t);
    float k2 = F32Math::force(s.alpha, s2, x2, t);
    fvec4 o;
    o.x = x1;
    o.y = x2;
    o.z = F32Math::kick(c.z, s.dt, k1);
    o.w = F32Math::kick(c.w, s.dt, k2);
    out[m] = o;
}

template <int ND>
__global__ void __launch_bounds__(256)
step_fp32_kernel(int N, uint32_t L, const fvec4 *__restrict__ in, fvec4 *__restrict__ out,
                 StepScalarsF s)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) site_fp32<ND>(N, m, in, out, s);
}

// Ensemble form for lattices the march kernel does not take (2D, higher-order
// stencils, other sides): one launch per step over the 256-site blocks of all
// members; each block finds its member by binary search over the members'
// first block.
template <int ND, bool F64, bool HAM>
__global__ void __launch_bounds__(256)
batch_site_kernel(const BatchMember *__restrict__ mem, int n_members, StepScalars s,
                  StepScalarsF sf)
{
    uint32_t b = blockIdx.x;
    int lo = 0, hi = n_members - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (mem[mid].item_begin <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const BatchMember &mb = mem[lo];
    uint32_t m = (b - mb.item_begin) * 256u + threadIdx.x;
    if (m >= mb.L) return;
    if (F64) {
        In64 in{(const dvec2 *)mb.in[PHI], (const dvec2 *)mb.in[VEL], (const dvec2 *)mb.in[KER]};
        Out64 out{(dvec2 *)mb.out[PHI], (dvec2 *)mb.out[VEL], (dvec2 *)mb.out[KER]};
        site_fp64<ND, HAM>(mb.N, m, in, out, s);
    } else {
        site_fp32<ND>(mb.N, m, (const fvec4 *)mb.in[PHI], (fvec4 *)mb.out[PHI], sf);
    }
}

// (Re)builds the ensemble tables for both ping-pong parities when the launch
// shape changed.  Not called during graph capture (it copies to the device).
int prepare_batch(Context *c)
{
    // the 2D multi-step kernel runs from a work list even for one lattice
    if (c->members.size() < 2)
        return (c->nd == 2) ? prepare_march2d_items(c, multi_depth(c)) : AXS_OK;
    // the spectral step runs member by member (rocFFT per lattice)
    if (c->p.laplacian != AXS_LAPLACIAN_STENCIL) return AXS_OK;
    // the Hamiltonian form's ensembles use the site kernel
    bool march = c->tune.kernel != AXS_KERNEL_SIMPLE && c->p.eom == AXS_EOM_FIELD_RESCALED;
    for (const Member &m : c->members) march = march && march3d_supported(c, m.N);
    const int rows = march_rows(c);
    int key = march ? rows * 100000 + ((c->prec == AXS_FP32) ? c->tune.ci32 : c->tune.ci64) : 0;
    if (key == c->batch_key && c->batch_tab[0])
        return prepare_multi_batch(c, multi_depth(c));
    std::vector<BatchMember> tab[2];
    uint32_t items = 0;
    for (const Member &m : c->members) {
        BatchMember b{};
        b.N = m.N;
        b.L = (uint32_t)m.L;
        b.item_begin = items;
        if (march) {
            b.nTj = m.N / (4 * rows);
            b.nTk = m.N / 64;
            b.CI = march_chunk(c, m.N);
            items += (uint32_t)b.nTj * b.nTk * (m.N / b.CI);
        } else {
            items += (uint32_t)((m.L + 255) / 256);
        }
        for (int h = 0; h < 2; ++h) {
            for (int a = 0; a < 3; ++a) {
                b.in[a] = m.state[h][a];
                b.out[a] = m.state[h ^ 1][a];
            }
            tab[h].push_back(b);
        }
    }
    for (int h = 0; h < 2; ++h) {
        if (!c->batch_tab[h]) {
            AXS_HIP_TRY(hipMalloc(&c->batch_tab[h], c->members.size() * sizeof(BatchMember)));
        }
        AXS_HIP_TRY(hipMemcpyAsync(c->batch_tab[h], tab[h].data(),
                                   tab[h].size() * sizeof(BatchMember), hipMemcpyHostToDevice,
                                   c->stream));
    }
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    c->batch_items = items;
    c->batch_key = key;
    c->batch_march = march;
    AXS_TRY(prepare_multi_batch(c, multi_depth(c)));
    return AXS_OK;
}

// FP32: ker = F(phi, tau) at the current state and phidot = u - dt/2 ker
template <int ND>
__global__ void __launch_bounds__(256)
canonical_fp32_kernel(int N, uint32_t L, const fvec4 *__restrict__ in, fvec2 *__restrict__ ker,
                      fvec2 *__restrict__ vel, StepScalarsF s, int recompute)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    fvec4 c = in[m];
    fvec2 k;
    if (recompute) {
        Site<ND> site = decompose<ND>(m, N);
        float s1 = -0.0f, s2 = -0.0f;
        for (int l = 1; l < s.ncoef; ++l) {
            float cl = s.c[l];
#pragma unroll
            for (int a = 0; a < ND; ++a) {
                fvec4 p = in[shifted<ND>(site, a, l, N)];
                fvec4 q = in[shifted<ND>(site, a, -l, N)];
                s1 = F32Math::acc(s1, cl, p.x);
                s2 = F32Math::acc(s2, cl, p.y);
                s1 = F32Math::acc(s1, cl, q.x);
                s2 = F32Math::acc(s2, cl, q.y);
            }
        }
        float t = F32Math::own(s.cl, F32Math::pot(c.x, c.y, s.pot0), s.ac0);
        k.x = F32Math::force(s.alpha, s1, c.x, t);
        k.y = F32Math::force(s.alpha, s2, c.y, t);
        ker[m] = k;
    } else {
        k = ker[m];
    }
    fvec2 v;
    v.x = c.z - s.half_dt * k.x;
    v.y = c.w - s.half_dt * k.y;
    vel[m] = v;
}

static inline unsigned grid_for(uint32_t L, unsigned block) { return (L + block - 1) / block; }

int launch_step(Context *c, const StepScalars &s)
{
    const int src = c->cur, dst = c->cur ^ 1;
    if (c->members.size() > 1 && c->batch_tab[src] && c->p.laplacian == AXS_LAPLACIAN_STENCIL) {
        // one launch for the whole ensemble (tables built by prepare_batch)
        if (c->batch_march) {
            AXS_TRY(launch_march3d_batch(c, src, s));
        } else {
            StepScalarsF f = to_float(s);
            const int nm = (int)c->members.size();
            dim3 grid(c->batch_items), block(256);
            if (c->prec == AXS_FP64) {
                auto k = (c->p.eom == AXS_EOM_HAMILTONIAN)
                             ? ((c->nd == 2) ? batch_site_kernel<2, true, true>
                                             : batch_site_kernel<3, true, true>)
                             : ((c->nd == 2) ? batch_site_kernel<2, true, false>
                                             : batch_site_kernel<3, true, false>);
                hipLaunchKernelGGL(k, grid, block, 0, c->stream, c->batch_tab[src], nm, s, f);
            } else {
                auto k = (c->nd == 2) ? batch_site_kernel<2, false, false>
                                      : batch_site_kernel<3, false, false>;
                hipLaunchKernelGGL(k, grid, block, 0, c->stream, c->batch_tab[src], nm, s, f);
            }
        }
        AXS_HIP_TRY(hipGetLastError());
        c->cur = dst;
        return AXS_OK;
    }
    for (Member &m : c->members) {
        uint32_t L = (uint32_t)m.L;
        unsigned grid = grid_for(L, 256);
        if (c->p.laplacian == AXS_LAPLACIAN_SPECTRAL) {
            AXS_TRY(run_spectral_step(c, m, src, dst, s));
            continue;
        }
        if (c->tune.kernel != AXS_KERNEL_SIMPLE && march3d_supported(c, m.N)) {
            AXS_TRY(launch_march3d(c, m, src, dst, s));
            continue;
        }
        if (c->prec == AXS_FP64) {
            In64 in{(const dvec2 *)m.state[src][PHI], (const dvec2 *)m.state[src][VEL],
                    (const dvec2 *)m.state[src][KER]};
            Out64 out{(dvec2 *)m.state[dst][PHI], (dvec2 *)m.state[dst][VEL],
                      (dvec2 *)m.state[dst][KER]};
            const bool ham = c->p.eom == AXS_EOM_HAMILTONIAN;
            auto k = (c->nd == 2) ? (ham ? step_fp64_kernel<2, true> : step_fp64_kernel<2, false>)
                                  : (ham ? step_fp64_kernel<3, true> : step_fp64_kernel<3, false>);
            hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, c->stream, m.N, L, in, out, s);
        } else {
            StepScalarsF f = to_float(s);
            const fvec4 *in = (const fvec4 *)m.state[src][PHI];
            fvec4 *out = (fvec4 *)m.state[dst][PHI];
            if (c->nd == 2)
                hipLaunchKernelGGL(step_fp32_kernel<2>, dim3(grid), dim3(256), 0, c->stream, m.N,
                                   L, in, out, f);
            else
                hipLaunchKernelGGL(step_fp32_kernel<3>, dim3(grid), dim3(256), 0, c->stream, m.N,
                                   L, in, out, f);
        }
    }
    AXS_HIP_TRY(hipGetLastError());
    c->cur = dst;
    return AXS_OK;
}

int launch_fp32_canonical(Context *c, Member &m, const StepScalars &s, bool recompute_ker)
{
    StepScalarsF f = to_float(s);
    uint32_t L = (uint32_t)m.L;
    unsigned grid = grid_for(L, 256);
    auto kern = (c->nd == 2) ? canonical_fp32_kernel<2> : canonical_fp32_kernel<3>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, c->stream, m.N, L,
                       (const fvec4 *)m.state[c->cur][PHI], (fvec2 *)m.aux_ker,
                       (fvec2 *)m.aux_vel, f, recompute_ker ? 1 : 0);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

}  // namespace axs
