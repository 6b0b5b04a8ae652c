// axs_march.hip -- 3D step kernel with 2.5D streaming for gfx950.
//
// A workgroup of four 64-wide waves owns a TJ x 64 tile of the (j, k) plane
// (k is the contiguous axis: one wave covers one 1 KB row segment of the
// interleaved per-site state)
// and marches along i over a chunk of CI planes.  Each work-item keeps the
// drifted field of its RJ sites at planes i-1, i, i+1 in registers; plane i
// with a one-site halo in j and k sits in LDS (double-buffered, one barrier
// per plane), so every field value is read from HBM once per chunk, plus the
// tile halo (mostly L2 hits) and two extra planes per chunk.
//
// Work-group -> tile order is XCD-aware: the hardware deals work-groups to
// the eight XCDs round-robin, so the logical tile index is rebuilt so that
// each XCD walks a contiguous run of tiles, k fastest: the work-groups
// resident on an XCD then cover whole 4 KB rows, and neighbouring tiles --
// which read each other's edge sites as halo -- share an L2.  (With j
// fastest the one-site k-halo columns miss in L2 and each drags in a full
// cache line: measured at twice the algorithmic read traffic.)
//
// The per-site arithmetic is supplied by a traits class:
//   ExactF64  the reference's velocity-Verlet (s_integrate.cpp:305-348) in
//             FP64 with its expression trees: drift from (phi, phidot, ker),
//             Laplacian summed in COO row order (centre, +-i, +-j, +-k),
//             then alpha, potential, kick.  Bit-identical to the oracle.
//   HalfF32   FP32 phi and half-step velocity u:  x = phi + dt u,
//             ker = F(x), u' = u + dt ker.
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kMK = 64;          // tile width along k = wave width
constexpr int kMW = 4;           // waves per work-group
constexpr int kLDK = kMK + 2;    // LDS row stride (halo column each side)

template <bool HAM>
struct ExactF64T {
    using T = double;
    using S = StepScalars;
    struct In {
        const dvec2 *__restrict__ phi, *__restrict__ vel, *__restrict__ ker;
    };
    struct Out {
        dvec2 *__restrict__ phi, *__restrict__ vel, *__restrict__ ker;
    };
    struct Raw {
        dvec2 p, v, k;
    };
    using Own = Raw;
    // Laplacian start: the centre term c0 x (s_integrate.cpp:322-324, COO row order)
    static __device__ __forceinline__ double centre(const S &s, double x) { return s.c[0] * x; }
    static __device__ __forceinline__ In make_in(void *const *p)
    {
        return In{(const dvec2 *)p[0], (const dvec2 *)p[1], (const dvec2 *)p[2]};
    }
    static __device__ __forceinline__ Out make_out(void *const *p)
    {
        return Out{(dvec2 *)p[0], (dvec2 *)p[1], (dvec2 *)p[2]};
    }
    static __device__ __forceinline__ void load(const In &a, uint32_t q, Raw &r)
    {
        r.p = a.phi[q];
        r.v = a.vel[q];
        r.k = a.ker[q];
    }
    // drift, s_integrate.cpp:313-314; Hamiltonian form :368-369
    static __device__ __forceinline__ void drift(const Raw &r, const S &s, double &x1, double &x2,
                                                 Own &o)
    {
        o = r;
        if (HAM) {
            x1 = r.p.x + s.dcoef * (r.v.x + s.hdt * r.k.x);
            x2 = r.p.y + s.dcoef * (r.v.y + s.hdt * r.k.y);
        } else {
            x1 = r.p.x + s.dt * (r.v.x + 0.5 * r.k.x * s.dt);
            x2 = r.p.y + s.dt * (r.v.y + 0.5 * r.k.y * s.dt);
        }
    }
    // s_integrate.cpp:322-340 for one site, sums already in COO order
    template <bool NT>
    static __device__ __forceinline__ void update(const Out &o, uint32_t q, const S &s, double x1,
                                                  double x2, double sum1, double sum2,
                                                  const Own &w)
    {
        double lap1 = s.alpha * sum1, lap2 = s.alpha * sum2;
        double S_ = x1 * x1 + x2 * x2 - s.tau2 + s.T03;
        dvec2 kn, vn, xn;
        kn.x = lap1 + s.cl * x1 * S_;
        kn.y = lap2 + s.cl * x2 * S_;
        vn.x = w.v.x + 0.5 * (w.k.x + kn.x) * s.dt;
        vn.y = w.v.y + 0.5 * (w.k.y + kn.y) * s.dt;
        xn.x = x1;
        xn.y = x2;
        if (NT) {
            store_nt(o.phi + q, xn);
            store_nt(o.vel + q, vn);
            store_nt(o.ker + q, kn);
        } else {
            o.phi[q] = xn;
            o.vel[q] = vn;
            o.ker[q] = kn;
        }
    }
};

using ExactF64 = ExactF64T<false>;

struct HalfF32 {
    using T = float;
    using S = StepScalarsF;
    struct In {
        const fvec4 *__restrict__ st;
    };
    struct Out {
        fvec4 *__restrict__ st;
    };
    using Raw = fvec4;
    struct Own {
        float u1, u2;
    };
    // the neighbour sum alone: the centre term joins the potential (F32Math)
    static __device__ __forceinline__ float centre(const S &, float) { return -0.0f; }
    static __device__ __forceinline__ In make_in(void *const *p) { return In{(const fvec4 *)p[0]}; }
    static __device__ __forceinline__ Out make_out(void *const *p) { return Out{(fvec4 *)p[0]}; }
    static __device__ __forceinline__ void load(const In &a, uint32_t q, Raw &r) { r = a.st[q]; }
    static __device__ __forceinline__ void drift(const Raw &r, const S &s, float &x1, float &x2,
                                                 Own &o)
    {
        o.u1 = r.z;
        o.u2 = r.w;
        x1 = F32Math::drift(r.x, r.z, s.dt);
        x2 = F32Math::drift(r.y, r.w, s.dt);
    }
    template <bool NT>
    static __device__ __forceinline__ void update(const Out &o, uint32_t q, const S &s, float x1,
                                                  float x2, float sum1, float sum2, const Own &w)
    {
        float tdt = F32Math::own(s.dtcl, F32Math::pot(x1, x2, s.pot0), s.dtac0);
        fvec4 r;
        r.x = x1;
        r.y = x2;
        r.z = F32Math::kick_dt(w.u1, s.dta, sum1, x1, tdt);
        r.w = F32Math::kick_dt(w.u2, s.dta, sum2, x2, tdt);
        if (NT)
            store_nt(o.st + q, r);
        else
            o.st[q] = r;
    }
};

struct MarchGeom {
    int N;        // lattice side
    int CI;       // planes per chunk
    int nTj, nTk; // tiles along j and k
    int remap;    // XCD-aware work-group order (xcd_order)
    int kfast;    // 1: tiles ordered k fastest (k-neighbours co-resident)
};

// PF: prefetch plane i+2 while plane i is computed (two planes of raw loads
// in flight per work-item).  NT: non-temporal stores of the new state.
// One work-group marches tile (tj, tk) of the lattice of side N over planes
// [ch*CI, ch*CI + CI).
template <class Tr, int RJ>
struct MarchLds {
    static constexpr int TJ = kMW * RJ;
    static constexpr int PL = (TJ + 2) * kLDK;    // one component of one LDS plane
    typename Tr::T v[2][2][PL];
};

template <class Tr, int RJ, bool PF, bool NT>
__device__ __forceinline__ void march_tile(const int N, const int CI, const int tj, const int tk,
                                           const int ch, const typename Tr::In &in,
                                           const typename Tr::Out &out,
                                           const typename Tr::S &s, MarchLds<Tr, RJ> &sh)
{
    using T = typename Tr::T;
    using Raw = typename Tr::Raw;
    using Own = typename Tr::Own;
    constexpr int TJ = kMW * RJ;
    auto &lds = sh.v;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t NN = (uint32_t)N * (uint32_t)N;
    const int j0 = tj * TJ, k0 = tk * kMK, i0 = ch * CI;
    const int k = k0 + lane;

    // halo role of this work-item: waves 0/1 the rows j0-1 / j0+TJ, wave 2
    // the columns k0-1 (lanes 0..TJ-1) and k0+64 (lanes 32..32+TJ-1)
    bool has_halo = false;
    uint32_t hrow = 0;        // j*N + k of the halo site within a plane
    int hoff = 0;             // LDS offset of the halo site
    if (wave == 0) {
        int j = (j0 == 0) ? N - 1 : j0 - 1;
        has_halo = true;
        hrow = (uint32_t)j * N + k;
        hoff = lane + 1;
    } else if (wave == 1) {
        int j = (j0 + TJ == N) ? 0 : j0 + TJ;
        has_halo = true;
        hrow = (uint32_t)j * N + k;
        hoff = (TJ + 1) * kLDK + lane + 1;
    } else if (wave == 2) {
        if (lane < TJ) {
            int kk = (k0 == 0) ? N - 1 : k0 - 1;
            has_halo = true;
            hrow = (uint32_t)(j0 + lane) * N + kk;
            hoff = (lane + 1) * kLDK;
        } else if (lane >= 32 && lane < 32 + TJ) {
            int kk = (k0 + kMK == N) ? 0 : k0 + kMK;
            has_halo = true;
            hrow = (uint32_t)(j0 + lane - 32) * N + kk;
            hoff = (lane - 32 + 1) * kLDK + kMK + 1;
        }
    }

    auto plane = [&](int i) -> uint32_t {
        int w = (i < 0) ? i + N : ((i >= N) ? i - N : i);
        return (uint32_t)w * NN;
    };
    auto own_q = [&](uint32_t po, int r) -> uint32_t {
        return po + (uint32_t)(j0 + wave + kMW * r) * N + k;
    };
    auto load_plane = [&](int i, Raw (&r)[RJ], Raw &h) {
        uint32_t po = plane(i);
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr) Tr::load(in, own_q(po, rr), r[rr]);
        if (has_halo) Tr::load(in, po + hrow, h);
    };
    // drift a loaded plane into registers (own) and LDS buffer `buf`
    auto drift_plane = [&](const Raw (&r)[RJ], const Raw &h, T (&x)[RJ][2], Own (&o)[RJ],
                           int buf) {
        T *W1 = lds[buf][0];
        T *W2 = lds[buf][1];
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr) {
            Tr::drift(r[rr], s, x[rr][0], x[rr][1], o[rr]);
            int off = (wave + kMW * rr + 1) * kLDK + lane + 1;
            W1[off] = x[rr][0];
            W2[off] = x[rr][1];
        }
        if (has_halo) {
            T h1, h2;
            Own dummy;
            Tr::drift(h, s, h1, h2, dummy);
            W1[hoff] = h1;
            W2[hoff] = h2;
        }
    };

    T xp[RJ][2], xc[RJ][2], xn[RJ][2];
    Own oc[RJ], on[RJ];
    Raw r1[RJ], r2[RJ], h1r, h2r;

    // prologue: plane i0-1 (own sites), plane i0 (own + halo) into LDS[0],
    // plane i0+1 requested
    {
        Raw rp[RJ];
        uint32_t po = plane(i0 - 1);
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr) Tr::load(in, own_q(po, rr), rp[rr]);
        load_plane(i0, r1, h1r);
        Own dummy[RJ];
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr) Tr::drift(rp[rr], s, xp[rr][0], xp[rr][1], dummy[rr]);
        drift_plane(r1, h1r, xc, oc, 0);
    }
    if (PF) load_plane(i0 + 1, r1, h1r);
    __syncthreads();

    for (int it = 0; it < CI; ++it) {
        const int i = i0 + it;
        const int buf = it & 1;
        if (PF) {
            if (it + 1 < CI) load_plane(i + 2, r2, h2r);       // in flight during plane i
            drift_plane(r1, h1r, xn, on, buf ^ 1);            // plane i+1 (loaded earlier)
        } else {
            load_plane(i + 1, r1, h1r);
            drift_plane(r1, h1r, xn, on, buf ^ 1);
        }

        // plane i: Laplacian in COO order, potential, kick
        uint32_t pc = plane(i);
        const T *L1 = lds[buf][0];
        const T *L2 = lds[buf][1];
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr) {
            int o = (wave + kMW * rr + 1) * kLDK + lane + 1;
            T s1 = Tr::centre(s, xc[rr][0]);
            T s2 = Tr::centre(s, xc[rr][1]);
            s1 = s1 + xn[rr][0];
            s2 = s2 + xn[rr][1];      // (i+1, j, k)
            s1 = s1 + xp[rr][0];
            s2 = s2 + xp[rr][1];      // (i-1, j, k)
            s1 = s1 + L1[o + kLDK];
            s2 = s2 + L2[o + kLDK];   // (i, j+1, k)
            s1 = s1 + L1[o - kLDK];
            s2 = s2 + L2[o - kLDK];   // (i, j-1, k)
            s1 = s1 + L1[o + 1];
            s2 = s2 + L2[o + 1];      // (i, j, k+1)
            s1 = s1 + L1[o - 1];
            s2 = s2 + L2[o - 1];      // (i, j, k-1)
            Tr::template update<NT>(out, own_q(pc, rr), s, xc[rr][0], xc[rr][1], s1, s2, oc[rr]);
        }
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr) {
            xp[rr][0] = xc[rr][0];
            xp[rr][1] = xc[rr][1];
            xc[rr][0] = xn[rr][0];
            xc[rr][1] = xn[rr][1];
            oc[rr] = on[rr];
            if (PF) r1[rr] = r2[rr];
        }
        if (PF) h1r = h2r;
    }
}

template <class Tr, int RJ, bool PF, bool NT>
__global__ void __launch_bounds__(kMW * 64)
march3d_kernel(MarchGeom g, typename Tr::In in, typename Tr::Out out, typename Tr::S s)
{
    __shared__ MarchLds<Tr, RJ> sh;
    uint32_t b = blockIdx.x;
    if (g.remap) b = xcd_order(b, gridDim.x);
    int tj, tk;
    if (g.kfast) {
        tk = (int)(b % (uint32_t)g.nTk);
        b /= (uint32_t)g.nTk;
        tj = (int)(b % (uint32_t)g.nTj);
        b /= (uint32_t)g.nTj;
    } else {
        tj = (int)(b % (uint32_t)g.nTj);
        b /= (uint32_t)g.nTj;
        tk = (int)(b % (uint32_t)g.nTk);
        b /= (uint32_t)g.nTk;
    }
    march_tile<Tr, RJ, PF, NT>(g.N, g.CI, tj, tk, (int)b, in, out, s, sh);
}

// Ensemble form: one persistent launch per step for every member.  Work
// items are the (member, tile, chunk) triples of all members, members in
// order, tiles k fastest; each work-group takes items blockIdx.x,
// blockIdx.x + gridDim.x, ...  All members share tau and hence the scalars.
template <class Tr, int RJ, bool PF, bool NT>
__global__ void __launch_bounds__(kMW * 64)
march3d_batch_kernel(const BatchMember *__restrict__ mem, int n_members, uint32_t n_items,
                     typename Tr::S s)
{
    __shared__ MarchLds<Tr, RJ> sh;
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        // member owning the item: last r with mem[r].item_begin <= item
        int lo = 0, hi = n_members - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (mem[mid].item_begin <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchMember &m = mem[lo];
        uint32_t t = item - m.item_begin;
        int tk = (int)(t % (uint32_t)m.nTk);
        t /= (uint32_t)m.nTk;
        int tj = (int)(t % (uint32_t)m.nTj);
        int ch = (int)(t / (uint32_t)m.nTj);
        march_tile<Tr, RJ, PF, NT>(m.N, m.CI, tj, tk, ch, Tr::make_in(m.in), Tr::make_out(m.out),
                                   s, sh);
    }
}

// ---------------------------------------------------------------------------

static int pick_ci(int N, int tiles_per_plane)
{
    // enough work-groups to fill 256 CUs several times over, planes per
    // chunk in [16, 128] and dividing N
    int want_chunks = (2048 + tiles_per_plane - 1) / tiles_per_plane;
    int ci = N / (want_chunks > 0 ? want_chunks : 1);
    if (ci > 128) ci = 128;
    if (ci < 16) ci = 16;
    while (ci > 1 && N % ci != 0) --ci;
    return ci;
}

bool march3d_supported(const Context *c, int N)
{
    return c->nd == 3 && c->p.stencil_setting == 0 && c->p.laplacian == AXS_LAPLACIAN_STENCIL &&
           N % kMK == 0 && N >= kMK && N % (kMW * march_rows(c)) == 0;
}

int march_rows(const Context *c) { return (c->prec == AXS_FP32) ? c->tune.rj32 : c->tune.rj64; }

int march_chunk(const Context *c, int N)
{
    int ci = (c->prec == AXS_FP32) ? c->tune.ci32 : c->tune.ci64;
    if (ci > 0 && N % ci == 0) return ci;
    int TJ = kMW * march_rows(c);
    return pick_ci(N, (N / TJ) * (N / kMK));
}

template <class Tr, int RJ, bool PF, bool NT>
static int launch_batch(Context *c, int src, const typename Tr::S &s)
{
    auto kern = march3d_batch_kernel<Tr, RJ, PF, NT>;
    int per_cu = 0;
    AXS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kMW * 64, 0));
    uint32_t cap = (uint32_t)c->num_cu * (uint32_t)(per_cu > 0 ? per_cu : 1);
    uint32_t grid = c->batch_items < cap ? c->batch_items : cap;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kMW * 64), 0, c->stream, c->batch_tab[src],
                       (int)c->members.size(), c->batch_items, s);
    return AXS_OK;
}

int launch_march3d_batch(Context *c, int src, const StepScalars &s)
{
    const int f = c->tune.flags & (kTunePrefetch | kTuneNT);
    const int rj = march_rows(c);
    if (c->prec == AXS_FP64) {
        if (rj == 2) {
            switch (f) {
            case 0: return launch_batch<ExactF64, 2, false, false>(c, src, s);
            case kTunePrefetch: return launch_batch<ExactF64, 2, true, false>(c, src, s);
            case kTuneNT: return launch_batch<ExactF64, 2, false, true>(c, src, s);
            default: return launch_batch<ExactF64, 2, true, true>(c, src, s);
            }
        }
        switch (f) {
        case 0: return launch_batch<ExactF64, 4, false, false>(c, src, s);
        case kTunePrefetch: return launch_batch<ExactF64, 4, true, false>(c, src, s);
        case kTuneNT: return launch_batch<ExactF64, 4, false, true>(c, src, s);
        default: return launch_batch<ExactF64, 4, true, true>(c, src, s);
        }
    }
    StepScalarsF fs = to_float(s);
    if (rj == 8) {
        switch (f) {
        case 0: return launch_batch<HalfF32, 8, false, false>(c, src, fs);
        case kTunePrefetch: return launch_batch<HalfF32, 8, true, false>(c, src, fs);
        case kTuneNT: return launch_batch<HalfF32, 8, false, true>(c, src, fs);
        default: return launch_batch<HalfF32, 8, true, true>(c, src, fs);
        }
    }
    switch (f) {
    case 0: return launch_batch<HalfF32, 4, false, false>(c, src, fs);
    case kTunePrefetch: return launch_batch<HalfF32, 4, true, false>(c, src, fs);
    case kTuneNT: return launch_batch<HalfF32, 4, false, true>(c, src, fs);
    default: return launch_batch<HalfF32, 4, true, true>(c, src, fs);
    }
}

template <class Tr, int RJ, bool PF, bool NT>
static int launch_march(Context *c, int N, typename Tr::In in, typename Tr::Out out,
                        const typename Tr::S &s, int ci)
{
    constexpr int TJ = kMW * RJ;
    if (N % TJ != 0) return AXS_ERR_UNSUPPORTED;
    MarchGeom g;
    g.N = N;
    g.nTj = N / TJ;
    g.nTk = N / kMK;
    g.CI = ci > 0 ? ci : pick_ci(N, g.nTj * g.nTk);
    if (N % g.CI != 0) return AXS_ERR_UNSUPPORTED;
    uint32_t nblocks = (uint32_t)g.nTj * g.nTk * (N / g.CI);
    g.remap = (nblocks % 8 == 0) ? 1 : 0;
    g.kfast = (c->tune.flags & kTuneJFast) ? 0 : 1;
    hipLaunchKernelGGL((march3d_kernel<Tr, RJ, PF, NT>), dim3(nblocks), dim3(kMW * 64), 0,
                       c->stream, g, in, out, s);
    return AXS_OK;
}

template <class Tr>
static int launch_march_f64(Context *c, Member &m, int src, int dst, const StepScalars &s)
{
    const int N = m.N;
    typename Tr::In in{(const dvec2 *)m.state[src][PHI], (const dvec2 *)m.state[src][VEL],
                       (const dvec2 *)m.state[src][KER]};
    typename Tr::Out out{(dvec2 *)m.state[dst][PHI], (dvec2 *)m.state[dst][VEL],
                         (dvec2 *)m.state[dst][KER]};
    const int f = c->tune.flags & (kTunePrefetch | kTuneNT);
    const int ci = c->tune.ci64;
    if (c->tune.rj64 == 2) {
        switch (f) {
        case 0: return launch_march<Tr, 2, false, false>(c, N, in, out, s, ci);
        case kTunePrefetch: return launch_march<Tr, 2, true, false>(c, N, in, out, s, ci);
        case kTuneNT: return launch_march<Tr, 2, false, true>(c, N, in, out, s, ci);
        default: return launch_march<Tr, 2, true, true>(c, N, in, out, s, ci);
        }
    }
    switch (f) {
    case 0: return launch_march<Tr, 4, false, false>(c, N, in, out, s, ci);
    case kTunePrefetch: return launch_march<Tr, 4, true, false>(c, N, in, out, s, ci);
    case kTuneNT: return launch_march<Tr, 4, false, true>(c, N, in, out, s, ci);
    default: return launch_march<Tr, 4, true, true>(c, N, in, out, s, ci);
    }
}

int launch_march3d(Context *c, Member &m, int src, int dst, const StepScalars &s)
{
    const int N = m.N;
    if (c->prec == AXS_FP64) {
        if (c->p.eom == AXS_EOM_HAMILTONIAN)
            return launch_march_f64<ExactF64T<true>>(c, m, src, dst, s);
        return launch_march_f64<ExactF64>(c, m, src, dst, s);
    }
    StepScalarsF f = to_float(s);
    HalfF32::In in{(const fvec4 *)m.state[src][PHI]};
    HalfF32::Out out{(fvec4 *)m.state[dst][PHI]};
    const int fl = c->tune.flags & (kTunePrefetch | kTuneNT);
    if (c->tune.rj32 == 8 && N % 32 == 0) {
        switch (fl) {
        case 0: return launch_march<HalfF32, 8, false, false>(c, N, in, out, f, c->tune.ci32);
        case kTunePrefetch: return launch_march<HalfF32, 8, true, false>(c, N, in, out, f, c->tune.ci32);
        case kTuneNT: return launch_march<HalfF32, 8, false, true>(c, N, in, out, f, c->tune.ci32);
        default: return launch_march<HalfF32, 8, true, true>(c, N, in, out, f, c->tune.ci32);
        }
    }
    switch (fl) {
    case 0: return launch_march<HalfF32, 4, false, false>(c, N, in, out, f, c->tune.ci32);
    case kTunePrefetch: return launch_march<HalfF32, 4, true, false>(c, N, in, out, f, c->tune.ci32);
    case kTuneNT: return launch_march<HalfF32, 4, false, true>(c, N, in, out, f, c->tune.ci32);
    default: return launch_march<HalfF32, 4, true, true>(c, N, in, out, f, c->tune.ci32);
    }
}

}  // namespace axs
