// axs_reduce.hip -- whole-lattice reductions in compensated FP64:
//   background sums of s_run.cpp:135-150 (six quantities);
//   energy components (kinetic, gradient, potential).
// The FP64 form sums its stored fields (background_kernel, energy_kernel);
// the FP32 form gets both sets from one pass over its (phi, u) state that
// rebuilds phidot per site (sample_fp32_kernel), kept per state and tau.
//
// Each work-item keeps Neumaier (sum, compensation) pairs over a fixed
// subset of sites (background: a strided set; energy: columns along the
// slowest axis, marched so that the slow-axis neighbour stays in
// registers); waves combine pairs with cross-lane shuffles, the block
// writes one pair per quantity, and a single-block pass folds the block
// pairs in block order.  The partition depends only on the lattice size,
// so results are identical from run to run.
#include <algorithm>
#include <mutex>

#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 2048;

// Fixed device workspace of axs_member_energy_noalloc: the per-block
// partial sums (at most kRedMaxBlocks blocks of nine Neumaier pairs, 288 KiB)
// and the folded sums.  One copy per GPU, held until the process ends.  Every
// call writes each partial it later reads (one per block) and the folded
// sums, so nothing left by an earlier call is read and nothing needs
// clearing.  Calls that overlap on one GPU share it and race.
constexpr size_t kNoAllocBytes = 512 * 1024;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kMaxDevices] = {};

// once per device, with the device current: the device address of that
// device's copy of `scratch` (host code uses only this address)
int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d outside the %d devices the fixed workspace table holds", device,
                  kMaxDevices);
        return AXS_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return AXS_OK;
    void *ptr = nullptr;
    AXS_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
        set_error("hipGetSymbolAddress returned no address for the fixed workspace");
        return AXS_ERR_HIP;
    }
    g_noalloc_ptr[device] = ptr;
    return AXS_OK;
}

int noalloc_workspace(int device, Workspace *ws)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (device < 0 || device >= kMaxDevices || !g_noalloc_ptr[device]) {
        set_error("the fixed workspace of device %d is not set up", device);
        return AXS_ERR_STATE;
    }
    ws->p = g_noalloc_ptr[device];
    ws->bytes = kNoAllocBytes;
    return AXS_OK;
}

// `need` bytes of device workspace: the caller's (checked against its size)
// or the context's scratch (grown on demand)
static int workspace_for(Context *c, const Workspace *ws, size_t need, void **out)
{
    if (ws) {
        if (need > ws->bytes) {
            set_error("the reduction needs %zu bytes of workspace; the fixed workspace holds %zu",
                      need, ws->bytes);
            return AXS_ERR_CAPACITY;
        }
        *out = ws->p;
        return AXS_OK;
    }
    AXS_TRY(ensure_scratch(c, need));
    *out = c->scratch;
    return AXS_OK;
}

struct Kahan {
    double s, c;
};

__device__ __forceinline__ void kadd(Kahan &k, double x)
{
    double t = k.s + x;
    if (fabs(k.s) >= fabs(x))
        k.c += (k.s - t) + x;
    else
        k.c += (x - t) + k.s;
    k.s = t;
}

__device__ __forceinline__ Kahan kmerge(Kahan a, Kahan b)
{
    // two-sum of the leading parts, compensations added
    double t = a.s + b.s;
    double bp = t - a.s;
    double e = (a.s - (t - bp)) + (b.s - bp);
    return Kahan{t, a.c + b.c + e};
}

__device__ __forceinline__ Kahan shfl_down_k(Kahan k, int o)
{
    return Kahan{__shfl_down(k.s, o, 64), __shfl_down(k.c, o, 64)};
}

// block reduction of Q quantities; result in lane 0 of wave 0
template <int Q>
__device__ __forceinline__ void block_reduce(Kahan (&k)[Q], Kahan (*sh)[kRedBlock / 64])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        for (int o = 32; o > 0; o >>= 1) k[q] = kmerge(k[q], shfl_down_k(k[q], o));
        if (lane == 0) sh[q][wave] = k[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            Kahan a = sh[q][0];
            for (int w = 1; w < kRedBlock / 64; ++w) a = kmerge(a, sh[q][w]);
            k[q] = a;
        }
    }
}

template <class P, class V>
__global__ void __launch_bounds__(kRedBlock)
background_kernel(uint32_t L, P phi, V vel, Kahan *__restrict__ part)
{
    __shared__ Kahan sh[6][kRedBlock / 64];
    Kahan k[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) k[q] = Kahan{0.0, 0.0};
    uint32_t stride = gridDim.x * kRedBlock;
    for (uint32_t m = blockIdx.x * kRedBlock + threadIdx.x; m < L; m += stride) {
        double a, b, va, vb;
        phi.get(m, a, b);
        vel.get(m, va, vb);
        kadd(k[0], a);
        kadd(k[1], b);
        kadd(k[2], va);
        kadd(k[3], vb);
        kadd(k[4], fabs(atan2(a, b)));            // s_run.cpp:145
        kadd(k[5], sqrt(a * a + b * b));          // s_run.cpp:146
    }
    block_reduce<6>(k, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) part[blockIdx.x * 6 + q] = k[q];
}

// Energy densities summed over the lattice.  A work-item owns one site q
// of the plane orthogonal to the slowest axis (3D: i, 2D: j) and marches
// along that axis over a chunk of CI planes: the forward neighbour along
// the slow axis is the next iteration's centre, kept in registers, so phi
// is read from HBM once per site (the in-plane neighbours +1 are cache
// hits).  The gradient terms are added in axis order (3D: i, j, k; 2D: i,
// j), as the reference's axis order of neighbours.
template <int ND, class P, class V>
__global__ void __launch_bounds__(kRedBlock)
energy_kernel(int N, int CI, uint32_t ncols, P phi, V vel, double tau2, double T03,
              double lam_4tau2, Kahan *__restrict__ part)
{
    __shared__ Kahan sh[3][kRedBlock / 64];
    Kahan k[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) k[q] = Kahan{0.0, 0.0};
    const uint32_t plane = (ND == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    const uint32_t stride = gridDim.x * kRedBlock;
    // XCD-aware order (work-groups go round-robin to the 8 XCDs): each XCD
    // takes a contiguous run of columns, so row j + 1 is read on the same L2
    uint32_t blk = blockIdx.x;
    if ((gridDim.x & 7u) == 0) blk = (blk & 7u) * (gridDim.x >> 3) + (blk >> 3);
    for (uint32_t col = blk * kRedBlock + threadIdx.x; col < ncols; col += stride) {
        const uint32_t q = col % plane, ch = col / plane;
        // in-plane forward neighbours: 3D (j, k+1) and (j+1, k); 2D i+1
        uint32_t qk, qj = 0;
        if (ND == 3) {
            const uint32_t j = q / (uint32_t)N, kk = q - j * (uint32_t)N;
            qk = j * (uint32_t)N + (kk + 1 == (uint32_t)N ? 0u : kk + 1);
            qj = (j + 1 == (uint32_t)N ? 0u : j + 1) * (uint32_t)N + kk;
        } else {
            qk = (q + 1 == (uint32_t)N) ? 0u : q + 1;
        }
        const int s0 = (int)ch * CI, s1 = (s0 + CI < N) ? s0 + CI : N;
        double a, b;
        phi.get((uint32_t)s0 * plane + q, a, b);
        for (int sl = s0; sl < s1; ++sl) {
            const uint32_t base = (uint32_t)sl * plane;
            const uint32_t nbase = (sl + 1 == N) ? 0u : base + plane;
            double sa, sb, ka, kb, va, vb;
            phi.get(nbase + q, sa, sb);              // slow-axis neighbour, next centre
            phi.get(base + qk, ka, kb);
            vel.get(base + q, va, vb);
            double g = 0.0;
            if (ND == 3) {
                double ja, jb;
                phi.get(base + qj, ja, jb);
                double d1 = sa - a, d2 = sb - b;
                g += d1 * d1 + d2 * d2;
                d1 = ja - a, d2 = jb - b;
                g += d1 * d1 + d2 * d2;
                d1 = ka - a, d2 = kb - b;
                g += d1 * d1 + d2 * d2;
            } else {
                double d1 = ka - a, d2 = kb - b;
                g += d1 * d1 + d2 * d2;
                d1 = sa - a, d2 = sb - b;
                g += d1 * d1 + d2 * d2;
            }
            const double S = a * a + b * b - tau2 + T03;
            kadd(k[0], 0.5 * (va * va + vb * vb));
            kadd(k[1], 0.5 * g);
            kadd(k[2], lam_4tau2 * S * S);
            a = sa;
            b = sb;
        }
    }
    block_reduce<3>(k, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[blockIdx.x * 3 + q] = k[q];
}

// FP32 form: the six background sums and the three energy components in one
// pass over the stored (phi, u) state, nothing written but the block sums.
// phidot = u - dt/2 ker is rebuilt per site with canonical_fp32_kernel's
// float expressions -- ker recomputed from phi with the step kernels'
// neighbour order, or read from kin while that holds the current state's
// -- so the sums run over the fields a read-back gives.  A work-item
// marches columns along the slowest axis (as energy_kernel), keeping phi of
// planes sl-R .. sl+R in a register ring and loading the next plane's
// operands while the current one is summed.  A column's sites (at most 64
// planes) are summed in plain FP64 -- within 63 ulp of the sum of their
// magnitudes -- and each column sum enters the work-item's compensated sums:
// the per-site compensated additions were a quarter of the kernel's vector
// instructions.  In 3D with N a multiple of 64 a work-group's 256 columns
// are 4 rows (j) of 64 columns (k), so the j +- 1 neighbours of its middle
// rows are its own rows' cache lines.
constexpr int kSampQ = 9;     // background (6), then kinetic, gradient, potential

template <int ND, int R>
__global__ void __launch_bounds__(kRedBlock)
sample_fp32_kernel(int N, int CI, uint32_t ncols, const fvec4 *__restrict__ in,
                   const fvec2 *__restrict__ kin, StepScalarsF s, int recompute, double tau2,
                   double T03, double lam_4tau2, Kahan *__restrict__ part)
{
    __shared__ Kahan sh[kSampQ][kRedBlock / 64];
    Kahan k[kSampQ];
#pragma unroll
    for (int q = 0; q < kSampQ; ++q) k[q] = Kahan{0.0, 0.0};
    const uint32_t plane = (ND == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    const uint32_t stride = gridDim.x * kRedBlock;
    uint32_t blk = blockIdx.x;
    if ((gridDim.x & 7u) == 0) blk = (blk & 7u) * (gridDim.x >> 3) + (blk >> 3);
    auto wrap = [N](int x) {               // periodic index, |x| within a few N
        while (x < 0) x += N;
        while (x >= N) x -= N;
        return x;
    };
    // phi (x, y) of a site as one 8-byte load from the (phi, u) array
    const fvec2 *__restrict__ in2 = reinterpret_cast<const fvec2 *>(in);
    auto xy = [&](uint32_t m) { return in2[2 * (size_t)m]; };
    struct Plane {
        fvec4 c;
        fvec2 a[2 * R], b[2 * R];
    };
    const bool tiled = ND == 3 && (N % 64) == 0;
    for (uint32_t col = blk * kRedBlock + threadIdx.x; col < ncols; col += stride) {
        uint32_t q = col % plane;
        if (tiled) {            // 4 x 64 tiles: band of 4 rows, then 64 columns
            const uint32_t tile = q >> 8, w = q & 255u, tpr = (uint32_t)N >> 6;
            q = ((tile / tpr) * 4u + (w >> 6)) * (uint32_t)N + (tile % tpr) * 64u + (w & 63u);
        }
        const int s0 = (int)(col / plane) * CI, s1 = (s0 + CI < N) ? s0 + CI : N;
        double acc[kSampQ];     // plain sums over this column's sites
#pragma unroll
        for (int t = 0; t < kSampQ; ++t) acc[t] = 0.0;
        // in-plane neighbours at distance l (+l, -l): 3D axis 1 (j) and
        // axis 2 (k); 2D axis 0 (i).  Index 0 is the forward neighbour.
        uint32_t ap[R], am[R], bp[R], bm[R];
        if (ND == 3) {
            const int j = (int)(q / (uint32_t)N), kk = (int)(q - (uint32_t)j * (uint32_t)N);
#pragma unroll
            for (int l = 1; l <= R; ++l) {
                ap[l - 1] = (uint32_t)wrap(j + l) * (uint32_t)N + (uint32_t)kk;
                am[l - 1] = (uint32_t)wrap(j - l) * (uint32_t)N + (uint32_t)kk;
                bp[l - 1] = (uint32_t)j * (uint32_t)N + (uint32_t)wrap(kk + l);
                bm[l - 1] = (uint32_t)j * (uint32_t)N + (uint32_t)wrap(kk - l);
            }
        } else {
#pragma unroll
            for (int l = 1; l <= R; ++l) {
                ap[l - 1] = (uint32_t)wrap((int)q + l);
                am[l - 1] = (uint32_t)wrap((int)q - l);
                bp[l - 1] = bm[l - 1] = 0;
            }
        }
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
        fvec2 ring[2 * R + 1];             // phi of slow index sl - R + t
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
            fvec2 kf;
            if (recompute) {
                float s1x = -0.0f, s2x = -0.0f;
#pragma unroll
                for (int l = 1; l <= R; ++l) {
                    const float cl = s.c[l];
                    const fvec2 p1 = cur.a[2 * l - 2], q1 = cur.a[2 * l - 1];
                    if (ND == 3) {
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
                        s1x = F32Math::acc(s1x, cl, p1.x);
                        s2x = F32Math::acc(s2x, cl, p1.y);
                        s1x = F32Math::acc(s1x, cl, q1.x);
                        s2x = F32Math::acc(s2x, cl, q1.y);
                        s1x = F32Math::acc(s1x, cl, ring[R + l].x);
                        s2x = F32Math::acc(s2x, cl, ring[R + l].y);
                        s1x = F32Math::acc(s1x, cl, ring[R - l].x);
                        s2x = F32Math::acc(s2x, cl, ring[R - l].y);
                    }
                }
                const float t = F32Math::own(s.cl, F32Math::pot(c.x, c.y, s.pot0), s.ac0);
                kf.x = F32Math::force(s.alpha, s1x, c.x, t);
                kf.y = F32Math::force(s.alpha, s2x, c.y, t);
            } else {
                kf = kin[base + q];
            }
            const float vx = c.z - s.half_dt * kf.x, vy = c.w - s.half_dt * kf.y;
            const double a = c.x, b = c.y, va = vx, vb = vy;
            acc[0] += a;                                 // s_run.cpp:141-146
            acc[1] += b;
            acc[2] += va;
            acc[3] += vb;
            acc[4] += fabs(atan2(a, b));
            acc[5] += sqrt(a * a + b * b);
            // forward differences in energy_kernel's axis order
            const fvec2 fs = ring[R + 1], fa = cur.a[0];
            double g = 0.0, d1, d2;
            if (ND == 3) {
                const fvec2 fb = cur.b[0];
                d1 = (double)fs.x - a, d2 = (double)fs.y - b;
                g += d1 * d1 + d2 * d2;
                d1 = (double)fa.x - a, d2 = (double)fa.y - b;
                g += d1 * d1 + d2 * d2;
                d1 = (double)fb.x - a, d2 = (double)fb.y - b;
                g += d1 * d1 + d2 * d2;
            } else {
                d1 = (double)fa.x - a, d2 = (double)fa.y - b;
                g += d1 * d1 + d2 * d2;
                d1 = (double)fs.x - a, d2 = (double)fs.y - b;
                g += d1 * d1 + d2 * d2;
            }
            const double S = a * a + b * b - tau2 + T03;
            acc[6] += 0.5 * (va * va + vb * vb);
            acc[7] += 0.5 * g;
            acc[8] += lam_4tau2 * S * S;
#pragma unroll
            for (int t2 = 0; t2 < 2 * R; ++t2) ring[t2] = ring[t2 + 1];
            ring[2 * R] = ring_next;
            cur = nxt;
        }
#pragma unroll
        for (int t = 0; t < kSampQ; ++t) kadd(k[t], acc[t]);
    }
    block_reduce<kSampQ>(k, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < kSampQ; ++q) part[blockIdx.x * kSampQ + q] = k[q];
}

__global__ void fold_kernel(const Kahan *__restrict__ part, int nblocks, int Q,
                            double *__restrict__ out)
{
    int q = threadIdx.x;
    if (q >= Q) return;
    Kahan a = part[q];
    for (int b = 1; b < nblocks; ++b) a = kmerge(a, part[b * Q + q]);
    out[q] = a.s + a.c;
}

static int red_blocks(uint32_t L)
{
    uint32_t need = (L + kRedBlock - 1) / kRedBlock;
    return (int)(need < (uint32_t)kRedMaxBlocks ? need : (uint32_t)kRedMaxBlocks);
}

// FP32 form: the nine sums of sample_fp32_kernel for the member's current
// state, computed once per (state, tau) and kept for the other call
template <int ND>
static auto sample_kernel_for(int R)
{
    return (R == 1)   ? sample_fp32_kernel<ND, 1>
           : (R == 2) ? sample_fp32_kernel<ND, 2>
           : (R == 3) ? sample_fp32_kernel<ND, 3>
                      : sample_fp32_kernel<ND, 4>;
}

static int fp32_sums(Context *c, Member &m, double out[kSampQ], const Workspace *ws = nullptr)
{
    if (!m.has_state) {
        set_error("member has no state: call axs_set_member_state first");
        return AXS_ERR_STATE;
    }
    if (m.samp_valid && m.samp_version == m.version && m.samp_tau == c->tau) {
        std::copy(m.samp, m.samp + kSampQ, out);
        return AXS_OK;
    }
    const StepScalarsF f = to_float(make_scalars(c->p, c->tau));
    const int N = m.N, R = f.ncoef - 1;
    if (R < 1 || R > 4) {
        set_error("stencil radius %d is not supported", R);
        return AXS_ERR_UNSUPPORTED;
    }
    const uint32_t plane = (c->nd == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    // slow-axis chunks as in run_energy
    const uint64_t want = (uint64_t)c->num_cu * 4 * kRedBlock;
    const uint32_t nch = (uint32_t)std::min<uint64_t>((uint64_t)N, (want + plane - 1) / plane);
    const int CI = std::min(64, (int)((N + nch - 1) / nch));
    const uint32_t ncols = plane * (uint32_t)((N + CI - 1) / CI);
    const int nb = red_blocks(ncols);
    const size_t part_b = (size_t)nb * kSampQ * sizeof(Kahan);
    void *wsp = nullptr;
    AXS_TRY(workspace_for(c, ws, part_b + kSampQ * sizeof(double), &wsp));
    Kahan *part = (Kahan *)wsp;
    double *dout = (double *)((char *)wsp + part_b);
    const double lam = (double)c->p.lambda, T0 = (double)c->p.T_initial;
    const double tau2 = c->tau * c->tau, T03 = (T0 * T0) / 3.0, lam_4tau2 = lam / (4.0 * tau2);
    auto k = (c->nd == 2) ? sample_kernel_for<2>(R) : sample_kernel_for<3>(R);
    hipLaunchKernelGGL(k, dim3(nb), dim3(kRedBlock), 0, c->stream, N, CI, ncols,
                       (const fvec4 *)m.state[c->cur][PHI], (const fvec2 *)m.aux_ker, f,
                       ker_current(c, m) ? 0 : 1, tau2, T03, lam_4tau2, part);
    hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(64), 0, c->stream, part, nb, kSampQ, dout);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(m.samp, dout, kSampQ * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    m.samp_valid = true;
    m.samp_version = m.version;
    m.samp_tau = c->tau;
    std::copy(m.samp, m.samp + kSampQ, out);
    return AXS_OK;
}

static void energy_scale(const Context *c, const double raw[3], double out[3])
{
    const double dx = (double)c->p.space_step;
    const double vol = (c->nd == 3) ? dx * dx * dx : dx * dx;
    out[0] = raw[0] * vol;
    out[1] = raw[1] / (dx * dx) * vol;
    out[2] = raw[2] * vol;
}

int run_background(Context *c, Member &m, double out[6])
{
    if (c->prec == AXS_FP32) {
        double sums[kSampQ];
        AXS_TRY(fp32_sums(c, m, sums));
        std::copy(sums, sums + 6, out);
        return AXS_OK;
    }
    FieldView fv;
    AXS_TRY(current_fields(c, m, &fv));
    uint32_t L = (uint32_t)m.L;
    int nb = red_blocks(L);
    size_t part_b = (size_t)nb * 6 * sizeof(Kahan);
    AXS_TRY(ensure_scratch(c, part_b + 64));
    Kahan *part = (Kahan *)c->scratch;
    double *dout = (double *)((char *)c->scratch + part_b);
    if (c->prec == AXS_FP64)
        hipLaunchKernelGGL((background_kernel<PhiF64, VelF64>), dim3(nb), dim3(kRedBlock), 0,
                           c->stream, L, PhiF64{(const dvec2 *)fv.phi},
                           VelF64{(const dvec2 *)fv.vel}, part);
    else
        hipLaunchKernelGGL((background_kernel<PhiF32, VelF32>), dim3(nb), dim3(kRedBlock), 0,
                           c->stream, L, PhiF32{(const fvec4 *)fv.phi},
                           VelF32{(const fvec2 *)fv.vel}, part);
    hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(64), 0, c->stream, part, nb, 6, dout);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(out, dout, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    return AXS_OK;
}

int run_energy(Context *c, Member &m, double out[3], const Workspace *ws)
{
    if (c->prec == AXS_FP32) {
        double sums[kSampQ];
        AXS_TRY(fp32_sums(c, m, sums, ws));
        energy_scale(c, sums + 6, out);
        return AXS_OK;
    }
    FieldView fv;
    AXS_TRY(current_fields(c, m, &fv));
    const int N = m.N;
    const uint32_t plane = (c->nd == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    // chunks along the slow axis: as few as still give a work-item to every
    // lane of four work-groups per CU
    const uint64_t want = (uint64_t)c->num_cu * 4 * kRedBlock;
    const uint32_t nch = (uint32_t)std::min<uint64_t>((uint64_t)N, (want + plane - 1) / plane);
    // at most 64 planes per chunk: work-items of one round march in step,
    // so the row j + 1 they share is still in L2
    const int CI = std::min(64, (int)((N + nch - 1) / nch));
    const uint32_t ncols = plane * (uint32_t)((N + CI - 1) / CI);
    int nb = red_blocks(ncols);
    size_t part_b = (size_t)nb * 3 * sizeof(Kahan);
    void *wsp = nullptr;
    AXS_TRY(workspace_for(c, ws, part_b + 64, &wsp));
    Kahan *part = (Kahan *)wsp;
    double *dout = (double *)((char *)wsp + part_b);
    double lam = (double)c->p.lambda, T0 = (double)c->p.T_initial;
    double tau2 = c->tau * c->tau;
    double T03 = (T0 * T0) / 3.0;
    double lam_4tau2 = lam / (4.0 * tau2);
    if (c->prec == AXS_FP64) {
        auto k = (c->nd == 2) ? energy_kernel<2, PhiF64, VelF64> : energy_kernel<3, PhiF64, VelF64>;
        hipLaunchKernelGGL(k, dim3(nb), dim3(kRedBlock), 0, c->stream, N, CI, ncols,
                           PhiF64{(const dvec2 *)fv.phi}, VelF64{(const dvec2 *)fv.vel}, tau2, T03,
                           lam_4tau2, part);
    } else {
        auto k = (c->nd == 2) ? energy_kernel<2, PhiF32, VelF32> : energy_kernel<3, PhiF32, VelF32>;
        hipLaunchKernelGGL(k, dim3(nb), dim3(kRedBlock), 0, c->stream, N, CI, ncols,
                           PhiF32{(const fvec4 *)fv.phi}, VelF32{(const fvec2 *)fv.vel}, tau2, T03,
                           lam_4tau2, part);
    }
    hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(64), 0, c->stream, part, nb, 3, dout);
    AXS_HIP_TRY(hipGetLastError());
    double raw[3];
    AXS_HIP_TRY(hipMemcpyAsync(raw, dout, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    energy_scale(c, raw, out);
    return AXS_OK;
}

}  // namespace axs
