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
// The FP32 form's read-back of phidot and ker (canonical_fp32_kernel) is
// here too.
#include <algorithm>

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
        for (int ax = 0; ax < ND; ++ax) {
            // the FP32 form's axis order: 3D 0, 1, 2; 2D 1 (j), then 0 (i),
            // so that the strip kernels start each sum from a register
            // operand and take both lane-shifted i neighbours as DPP operands
            const int a = (ND == 2) ? 1 - ax : ax;
            fvec4 p = in[shifted<ND>(site, a, l, N)];
            fvec4 q = in[shifted<ND>(site, a, -l, N)];
            s1 = F32Math::acc(s1, cl, F32Math::drift(p.x, p.z, s.dt));
            s2 = F32Math::acc(s2, cl, F32Math::drift(p.y, p.w, s.dt));
            s1 = F32Math::acc(s1, cl, F32Math::drift(q.x, q.z, s.dt));
            s2 = F32Math::acc(s2, cl, F32Math::drift(q.y, q.w, s.dt));
        }
    }
    float tdt = F32Math::own(s.dtcl, F32Math::pot(x1, x2, s.pot0), s.dtac0);
    fvec4 o;
    o.x = x1;
    o.y = x2;
    o.z = F32Math::kick_dt(c.z, s.dta, s1, x1, tdt);
    o.w = F32Math::kick_dt(c.w, s.dta, s2, x2, tdt);
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

// FP32: ker = F(phi, tau) at the current state and phidot = u - dt/2 ker.
// A work-item owns one site q of the plane orthogonal to the slowest axis
// (3D: i, 2D: j) and marches along that axis over a chunk of CI planes,
// keeping phi of planes sl-R .. sl+R in a register ring: every plane is
// read from HBM once per chunk (the in-plane neighbours, whose offsets are
// fixed along the march, are cache hits).  The neighbour terms are added in
// the one-step kernels' order (distance l, then axis, +l before -l) with
// the same F32Math expressions, so ker is theirs bit for bit.
template <int ND, int R>
__global__ void __launch_bounds__(256)
canonical_fp32_kernel(int N, int CI, uint32_t ncols, const fvec4 *__restrict__ in,
                      fvec2 *__restrict__ ker, fvec2 *__restrict__ vel, StepScalarsF s,
                      int recompute)
{
    // XCD-aware order: hardware deals work-groups round-robin to the 8 XCDs;
    // give each XCD a contiguous run of columns, so that the rows j +- l a
    // work-group reads are being read by work-groups on the same L2
    uint32_t b = blockIdx.x;
    if ((gridDim.x & 7u) == 0) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
    const uint32_t col = b * 256u + threadIdx.x;
    if (col >= ncols) return;
    const uint32_t plane = (ND == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    const uint32_t q = col % plane;
    const int s0 = (int)(col / plane) * CI, s1 = (s0 + CI < N) ? s0 + CI : N;
    if (!recompute) {
        for (int sl = s0; sl < s1; ++sl) {
            const uint32_t m = (uint32_t)sl * plane + q;
            const fvec4 c = in[m];
            const fvec2 k = ker[m];
            fvec2 v;
            v.x = c.z - s.half_dt * k.x;
            v.y = c.w - s.half_dt * k.y;
            vel[m] = v;
        }
        return;
    }
    auto wrap = [N](int x) {               // periodic index, |x| within a few N
        while (x < 0) x += N;
        while (x >= N) x -= N;
        return x;
    };
    // in-plane neighbour offsets at distance l (+l, -l) along the in-plane
    // axes: 3D axis 1 (j) and axis 2 (k); 2D axis 0 (i)
    uint32_t ap[R], am[R], bp[R], bm[R];
    if (ND == 3) {
        const int j = (int)(q / (uint32_t)N), k = (int)(q - (uint32_t)j * (uint32_t)N);
#pragma unroll
        for (int l = 1; l <= R; ++l) {
            ap[l - 1] = (uint32_t)wrap(j + l) * (uint32_t)N + (uint32_t)k;
            am[l - 1] = (uint32_t)wrap(j - l) * (uint32_t)N + (uint32_t)k;
            bp[l - 1] = (uint32_t)j * (uint32_t)N + (uint32_t)wrap(k + l);
            bm[l - 1] = (uint32_t)j * (uint32_t)N + (uint32_t)wrap(k - l);
        }
    } else {
#pragma unroll
        for (int l = 1; l <= R; ++l) {
            ap[l - 1] = (uint32_t)wrap((int)q + l);
            am[l - 1] = (uint32_t)wrap((int)q - l);
            bp[l - 1] = bm[l - 1] = 0;
        }
    }
    // phi (x, y) of a site as one 8-byte load from the (phi, u) array
    const fvec2 *__restrict__ in2 = reinterpret_cast<const fvec2 *>(in);
    auto xy = [&](uint32_t m) { return in2[2 * (size_t)m]; };
    // one plane's operands: centre (phi, u) and the in-plane neighbours;
    // the next plane's are loaded while the current one is computed
    struct Plane {
        fvec4 c;
        fvec2 a[2 * R], b[2 * R];
    };
    auto load = [&](int sl, Plane &P) {
        const uint32_t base = (uint32_t)sl * plane;
        P.c = in[base + q];
#pragma unroll
        for (int l = 0; l < R; ++l) {
            P.a[2 * l] = xy(base + ap[l]);
            P.a[2 * l + 1] = xy(base + am[l]);
            if (ND == 3) {
                P.b[2 * l] = xy(base + bp[l]);
                P.b[2 * l + 1] = xy(base + bm[l]);
            }
        }
    };
    fvec2 ring[2 * R + 1];                 // phi of slow index sl - R + t
#pragma unroll
    for (int t = 0; t < 2 * R + 1; ++t) ring[t] = xy((uint32_t)wrap(s0 - R + t) * plane + q);
    Plane cur, nxt;
    load(s0, cur);
    for (int sl = s0; sl < s1; ++sl) {
        const uint32_t base = (uint32_t)sl * plane;
        fvec2 ring_next = ring[2 * R];
        if (sl + 1 < s1) {
            load(sl + 1, nxt);
            ring_next = xy((uint32_t)wrap(sl + 1 + R) * plane + q);
        }
        const fvec4 c = cur.c;
        float s1x = -0.0f, s2x = -0.0f;
#pragma unroll
        for (int l = 1; l <= R; ++l) {
            const float cl = s.c[l];
            const fvec2 p1 = cur.a[2 * l - 2], q1 = cur.a[2 * l - 1];
            if (ND == 3) {
                // axis 0: the slow axis, from the ring; then j, then k
                s1x = F32Math::acc(s1x, cl, ring[R + l].x);
                s2x = F32Math::acc(s2x, cl, ring[R + l].y);
                s1x = F32Math::acc(s1x, cl, ring[R - l].x);
                s2x = F32Math::acc(s2x, cl, ring[R - l].y);
                s1x = F32Math::acc(s1x, cl, p1.x);
                s2x = F32Math::acc(s2x, cl, p1.y);
                s1x = F32Math::acc(s1x, cl, q1.x);
                s2x = F32Math::acc(s2x, cl, q1.y);
                const fvec2 p2 = cur.b[2 * l - 2], q2 = cur.b[2 * l - 1];
                s1x = F32Math::acc(s1x, cl, p2.x);
                s2x = F32Math::acc(s2x, cl, p2.y);
                s1x = F32Math::acc(s1x, cl, q2.x);
                s2x = F32Math::acc(s2x, cl, q2.y);
            } else {
                // the FP32 form's 2D order: axis 1 (the slow axis j, from
                // the ring), then axis 0 (in-plane i)
                s1x = F32Math::acc(s1x, cl, ring[R + l].x);
                s2x = F32Math::acc(s2x, cl, ring[R + l].y);
                s1x = F32Math::acc(s1x, cl, ring[R - l].x);
                s2x = F32Math::acc(s2x, cl, ring[R - l].y);
                s1x = F32Math::acc(s1x, cl, p1.x);
                s2x = F32Math::acc(s2x, cl, p1.y);
                s1x = F32Math::acc(s1x, cl, q1.x);
                s2x = F32Math::acc(s2x, cl, q1.y);
            }
        }
        const float t = F32Math::own(s.cl, F32Math::pot(c.x, c.y, s.pot0), s.ac0);
        fvec2 k;
        k.x = F32Math::force(s.alpha, s1x, c.x, t);
        k.y = F32Math::force(s.alpha, s2x, c.y, t);
        ker[base + q] = k;
        fvec2 v;
        v.x = c.z - s.half_dt * k.x;
        v.y = c.w - s.half_dt * k.y;
        vel[base + q] = v;
#pragma unroll
        for (int t2 = 0; t2 < 2 * R; ++t2) ring[t2] = ring[t2 + 1];
        ring[2 * R] = ring_next;
        cur = nxt;
    }
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

template <int ND>
static void launch_canonical_nd(Context *c, int R, unsigned grid, int N, int CI, uint32_t ncols,
                                const fvec4 *in, fvec2 *ker, fvec2 *vel, const StepScalarsF &f,
                                int recompute)
{
    auto k = (R == 1)   ? canonical_fp32_kernel<ND, 1>
             : (R == 2) ? canonical_fp32_kernel<ND, 2>
             : (R == 3) ? canonical_fp32_kernel<ND, 3>
                        : canonical_fp32_kernel<ND, 4>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, c->stream, N, CI, ncols, in, ker, vel, f,
                       recompute);
}

int launch_fp32_canonical(Context *c, Member &m, const StepScalars &s, bool recompute_ker)
{
    StepScalarsF f = to_float(s);
    const int N = m.N, R = f.ncoef - 1;
    if (R < 1 || R > 4) {
        set_error("stencil radius %d is not supported", R);
        return AXS_ERR_UNSUPPORTED;
    }
    const uint32_t plane = (c->nd == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    // chunks along the slow axis: as few as still give a work-item to every
    // lane of four work-groups per CU
    const uint64_t want = (uint64_t)c->num_cu * 4 * 256;
    const uint32_t nch = (uint32_t)std::min<uint64_t>((uint64_t)N, (want + plane - 1) / plane);
    // at most 64 planes per chunk: work-groups of one round then march in
    // step, so the rows j +- l they share are still in L2
    const int CI = std::min(64, (int)((N + nch - 1) / nch));
    const uint32_t ncols = plane * (uint32_t)((N + CI - 1) / CI);
    const unsigned grid = grid_for(ncols, 256);
    const fvec4 *in = (const fvec4 *)m.state[c->cur][PHI];
    if (c->nd == 2)
        launch_canonical_nd<2>(c, R, grid, N, CI, ncols, in, (fvec2 *)m.aux_ker,
                               (fvec2 *)m.aux_vel, f, recompute_ker ? 1 : 0);
    else
        launch_canonical_nd<3>(c, R, grid, N, CI, ncols, in, (fvec2 *)m.aux_ker,
                               (fvec2 *)m.aux_vel, f, recompute_ker ? 1 : 0);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

}  // namespace axs
