// axs_reduce.hip -- whole-lattice reductions in compensated FP64:
//   background sums of s_run.cpp:135-150 (six quantities);
//   energy components (kinetic, gradient, potential).
// Both sets come from one pass over the member's state, kept per state and
// tau: sample_fp64_kernel over the FP64 form's stored (phi, phidot) pairs,
// sample_fp32_kernel over the FP32 form's (phi, u) state, rebuilding phidot
// per site.
//
// A work-item marches columns along the slowest axis, so the slow-axis
// neighbour stays in registers; it sums a column's sites in plain FP64 and
// adds each column sum to its Neumaier (sum, compensation) pairs; waves
// combine pairs with cross-lane shuffles, the block writes one pair per
// quantity, and a single-block pass folds the block pairs in block order.
// The partition depends only on the lattice size, so results are identical
// from run to run.
#include <algorithm>
#include <mutex>

#include "axs_atan2.hpp"
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

// the arctangent table of abs_atan2_fast, staged in LDS by the block
__device__ __forceinline__ void stage_atan_table(double (*tab)[2])
{
    for (int i = threadIdx.x; i < 257; i += blockDim.x) {
        tab[i][0] = kAtanTable[i][0];
        tab[i][1] = kAtanTable[i][1];
    }
    __syncthreads();
}

// sqrt(a^2 + b^2) for float values a, b: the sum of squares (a^2 exact in
// double, one rounding) lies in {0} and [2^-298, 2^257], so the hardware
// reciprocal square root needs no range scaling; one Goldschmidt step from
// it and one residual correction (y + h (s - y^2)) give the root within
// an ulp, and s = 0 gives 0.
__device__ __forceinline__ double norm_f32in(double a, double b)
{
    const double s = fma(a, a, b * b);
    const double r = __builtin_amdgcn_rsq(s);
    double y = s * r, h = 0.5 * r;
    const double e = fma(-h, y, 0.5);
    y = fma(y, e, y);
    h = fma(h, e, h);
    y = fma(fma(-y, y, s), h, y);
    return (s > 0.0) ? y : 0.0;
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

template <int ND, int R, bool UNI>
__global__ void __launch_bounds__(kRedBlock) __attribute__((amdgpu_waves_per_eu(R == 1 ? 4 : 1)))
sample_fp32_kernel(int N, int CI, uint32_t ncols, const fvec4 *__restrict__ in,
                   const fvec2 *__restrict__ kin, StepScalarsF s, int recompute, double tau2,
                   double T03, double lam_4tau2, Kahan *__restrict__ part)
{
    __shared__ Kahan sh[kSampQ][kRedBlock / 64];
    __shared__ double atab[257][2];
    stage_atan_table(atab);
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
    // Slices of the slow axis by byte address: the slice base is one value
    // per wave when UNI (a slice holds a multiple of 64 sites, so a wave's
    // 64 columns share their slices) and the per-site offsets are 32-bit
    const char *__restrict__ inb = reinterpret_cast<const char *>(in);
    const char *__restrict__ kinb = reinterpret_cast<const char *>(kin);
    auto slice = [&](int sl) { return inb + (size_t)(uint32_t)sl * plane * sizeof(fvec4); };
    auto ld4 = [](const char *sp, uint32_t o) { return *reinterpret_cast<const fvec4 *>(sp + o); };
    auto ld2 = [](const char *sp, uint32_t o) { return *reinterpret_cast<const fvec2 *>(sp + o); };
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
        int s0 = (int)(col / plane) * CI;
        if (UNI) s0 = __builtin_amdgcn_readfirstlane(s0);
        const int s1 = (s0 + CI < N) ? s0 + CI : N;
        double acc[kSampQ];     // plain sums over this column's sites
#pragma unroll
        for (int t = 0; t < kSampQ; ++t) acc[t] = 0.0;
        // byte offsets within a slice: the site (oc) and its in-plane
        // neighbours at distance l (+l, -l): 3D axis 1 (j) and axis 2 (k);
        // 2D axis 0 (i).  Index 0 is the forward neighbour.
        const uint32_t oc = q * (uint32_t)sizeof(fvec4);
        uint32_t ap[R], am[R], bp[R], bm[R];
        if (ND == 3) {
            const int j = (int)(q / (uint32_t)N), kk = (int)(q - (uint32_t)j * (uint32_t)N);
#pragma unroll
            for (int l = 1; l <= R; ++l) {
                ap[l - 1] = ((uint32_t)wrap(j + l) * (uint32_t)N + (uint32_t)kk) * 16u;
                am[l - 1] = ((uint32_t)wrap(j - l) * (uint32_t)N + (uint32_t)kk) * 16u;
                bp[l - 1] = ((uint32_t)j * (uint32_t)N + (uint32_t)wrap(kk + l)) * 16u;
                bm[l - 1] = ((uint32_t)j * (uint32_t)N + (uint32_t)wrap(kk - l)) * 16u;
            }
        } else {
#pragma unroll
            for (int l = 1; l <= R; ++l) {
                ap[l - 1] = (uint32_t)wrap((int)q + l) * 16u;
                am[l - 1] = (uint32_t)wrap((int)q - l) * 16u;
                bp[l - 1] = bm[l - 1] = 0;
            }
        }
        auto load = [&](int sl, Plane &P) {
            const char *sp = slice(sl);
            P.c = ld4(sp, oc);
#pragma unroll
            for (int l = 0; l < R; ++l) {
                P.a[2 * l] = ld2(sp, ap[l]);
                P.a[2 * l + 1] = ld2(sp, am[l]);
                if (ND == 3) {
                    P.b[2 * l] = ld2(sp, bp[l]);
                    P.b[2 * l + 1] = ld2(sp, bm[l]);
                }
            }
        };
        fvec2 ring[2 * R + 1];             // phi of slow index sl - R + t
#pragma unroll
        for (int t = 0; t < 2 * R + 1; ++t) ring[t] = ld2(slice(wrap(s0 - R + t)), oc);
        Plane cur, nxt;
        load(s0, cur);
        for (int sl = s0; sl < s1; ++sl) {
            fvec2 ring_next = ring[2 * R];
            if (sl + 1 < s1) {
                load(sl + 1, nxt);
                ring_next = ld2(slice(wrap(sl + 1 + R)), oc);
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
                        // the FP32 form's 2D order: j (the slow axis, from
                        // the ring), then i
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
                kf.x = F32Math::force(s.alpha, s1x, c.x, t);
                kf.y = F32Math::force(s.alpha, s2x, c.y, t);
            } else {
                kf = *reinterpret_cast<const fvec2 *>(kinb + (size_t)(uint32_t)sl * plane * sizeof(fvec2) +
                                                      q * (uint32_t)sizeof(fvec2));
            }
            const float vx = c.z - s.half_dt * kf.x, vy = c.w - s.half_dt * kf.y;
            const double a = c.x, b = c.y, va = vx, vb = vy;
            acc[0] += a;                                 // s_run.cpp:141-146
            acc[1] += b;
            acc[2] += va;
            acc[3] += vb;
            acc[4] += abs_atan2_f32in(a, b, atab);
            acc[5] += norm_f32in(a, b);
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

// FP64 form: the six background sums and the three energy components in one
// pass over the stored (phi, phidot) pairs.  A work-item marches a column
// along the slowest axis (3D: i, 2D: j) over a chunk of CI planes: the
// forward neighbour along that axis is the next iteration's centre, kept in
// registers, so phi is read from HBM once per site, and the in-plane forward
// neighbours (3D: j + 1, k + 1; 2D: i + 1) are cache hits; the next plane's
// loads are issued while the current one is summed.  The gradient terms are
// added in the reference's axis order (3D: i, j, k; 2D: i, j).  Columns,
// tiles and sums as in sample_fp32_kernel: plain FP64 sums over a column's
// sites, each entering the work-item's compensated sums.
template <int ND, bool UNI>
__global__ void __launch_bounds__(kRedBlock) __attribute__((amdgpu_waves_per_eu(4)))
sample_fp64_kernel(int N, int CI, uint32_t ncols, const dvec2 *__restrict__ phi,
                   const dvec2 *__restrict__ vel, double tau2, double T03, double lam_4tau2,
                   Kahan *__restrict__ part)
{
    __shared__ Kahan sh[kSampQ][kRedBlock / 64];
    __shared__ double atab[257][2];
    stage_atan_table(atab);
    Kahan k[kSampQ];
#pragma unroll
    for (int q = 0; q < kSampQ; ++q) k[q] = Kahan{0.0, 0.0};
    const uint32_t plane = (ND == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    const uint32_t stride = gridDim.x * kRedBlock;
    uint32_t blk = blockIdx.x;
    if ((gridDim.x & 7u) == 0) blk = (blk & 7u) * (gridDim.x >> 3) + (blk >> 3);
    const char *__restrict__ pb = reinterpret_cast<const char *>(phi);
    const char *__restrict__ vb = reinterpret_cast<const char *>(vel);
    const size_t slice_b = (size_t)plane * sizeof(dvec2);
    auto ld = [](const char *sp, uint32_t o) { return *reinterpret_cast<const dvec2 *>(sp + o); };
    const bool tiled = ND == 3 && (N % 64) == 0;
    for (uint32_t col = blk * kRedBlock + threadIdx.x; col < ncols; col += stride) {
        uint32_t q = col % plane;
        if (tiled) {            // 4 x 64 tiles: band of 4 rows, then 64 columns
            const uint32_t tile = q >> 8, w = q & 255u, tpr = (uint32_t)N >> 6;
            q = ((tile / tpr) * 4u + (w >> 6)) * (uint32_t)N + (tile % tpr) * 64u + (w & 63u);
        }
        int s0 = (int)(col / plane) * CI;
        if (UNI) s0 = __builtin_amdgcn_readfirstlane(s0);
        const int s1 = (s0 + CI < N) ? s0 + CI : N;
        // byte offsets within a slice: the site and its in-plane forward
        // neighbours (3D: (j + 1, k) and (j, k + 1); 2D: i + 1)
        const uint32_t oc = q * (uint32_t)sizeof(dvec2);
        uint32_t oj = 0, ok;
        if (ND == 3) {
            const uint32_t j = q / (uint32_t)N, kk = q - j * (uint32_t)N;
            ok = (j * (uint32_t)N + (kk + 1 == (uint32_t)N ? 0u : kk + 1)) * 16u;
            oj = ((j + 1 == (uint32_t)N ? 0u : j + 1) * (uint32_t)N + kk) * 16u;
        } else {
            ok = (q + 1 == (uint32_t)N ? 0u : q + 1) * 16u;
        }
        double acc[kSampQ];
#pragma unroll
        for (int t = 0; t < kSampQ; ++t) acc[t] = 0.0;
        const char *sp = pb + (size_t)(uint32_t)s0 * slice_b;
        dvec2 c = ld(sp, oc);
        int nsl = (s0 + 1 == N) ? 0 : s0 + 1;
        dvec2 nx = ld(pb + (size_t)(uint32_t)nsl * slice_b, oc);   // slow-axis neighbour
        dvec2 nj = ND == 3 ? ld(sp, oj) : dvec2{0.0, 0.0};
        dvec2 nk = ld(sp, ok);
        dvec2 v = ld(vb + (size_t)(uint32_t)s0 * slice_b, oc);
        for (int sl = s0; sl < s1; ++sl) {
            dvec2 nx2 = nx, nj2 = nj, nk2 = nk, v2 = v;
            if (sl + 1 < s1) {            // the next plane's operands
                const int sl1 = sl + 1, n2 = (sl1 + 1 == N) ? 0 : sl1 + 1;
                const char *sp1 = pb + (size_t)(uint32_t)sl1 * slice_b;
                nx2 = ld(pb + (size_t)(uint32_t)n2 * slice_b, oc);
                if (ND == 3) nj2 = ld(sp1, oj);
                nk2 = ld(sp1, ok);
                v2 = ld(vb + (size_t)(uint32_t)sl1 * slice_b, oc);
            }
            const double a = c.x, b = c.y, va = v.x, vv = v.y;
            acc[0] += a;                                 // s_run.cpp:141-146
            acc[1] += b;
            acc[2] += va;
            acc[3] += vv;
            acc[4] += abs_atan2_fast(a, b, atab);
            acc[5] += sqrt(a * a + b * b);
            double g = 0.0, d1, d2;
            if (ND == 3) {
                d1 = nx.x - a, d2 = nx.y - b;
                g += d1 * d1 + d2 * d2;
                d1 = nj.x - a, d2 = nj.y - b;
                g += d1 * d1 + d2 * d2;
                d1 = nk.x - a, d2 = nk.y - b;
                g += d1 * d1 + d2 * d2;
            } else {
                d1 = nk.x - a, d2 = nk.y - b;
                g += d1 * d1 + d2 * d2;
                d1 = nx.x - a, d2 = nx.y - b;
                g += d1 * d1 + d2 * d2;
            }
            const double S = a * a + b * b - tau2 + T03;
            acc[6] += 0.5 * (va * va + vv * vv);
            acc[7] += 0.5 * g;
            acc[8] += lam_4tau2 * S * S;
            c = nx;
            nx = nx2;
            nj = nj2;
            nk = nk2;
            v = v2;
        }
#pragma unroll
        for (int t = 0; t < kSampQ; ++t) kadd(k[t], acc[t]);
    }
    block_reduce<kSampQ>(k, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < kSampQ; ++q) part[blockIdx.x * kSampQ + q] = k[q];
}

// quantity blockIdx.x of the block pairs part[b * Q + q]: work-item t
// merges blocks t, t + 256, ... in order, then a fixed pairwise tree over
// the work-items (deterministic for a given block count)
__global__ void __launch_bounds__(256)
fold_kernel(const Kahan *__restrict__ part, int nblocks, int Q, double *__restrict__ out)
{
    __shared__ Kahan sa[256];
    const int q = blockIdx.x;
    Kahan a{0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += 256) a = kmerge(a, part[b * Q + q]);
    sa[threadIdx.x] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sa[threadIdx.x] = kmerge(sa[threadIdx.x], sa[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sa[0].s + sa[0].c;
}

// The index of the first site whose phi1 or phi2 is NaN (the reference's
// debug scan, s_run.cpp:4-12): each work-item's first hit in its strided
// (ascending) set of sites, the wave's minimum, and one atomic minimum per
// wave into *first -- the minimum of a set does not depend on the order, so
// the result is the lowest such index whatever the schedule.
template <class V>
__global__ void __launch_bounds__(kRedBlock)
nan_first_kernel(uint32_t L, const V *__restrict__ phi, unsigned long long *__restrict__ first)
{
    unsigned long long best = ~0ull;
    const uint32_t stride = gridDim.x * kRedBlock;
    for (uint32_t m = blockIdx.x * kRedBlock + threadIdx.x; m < L; m += stride) {
        const V w = phi[m];
        if (isnan(w.x) || isnan(w.y)) {
            best = m;
            break;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_down(best, o, 64);
        best = t < best ? t : best;
    }
    if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(first, best);
}

static int red_blocks(uint32_t L)
{
    uint32_t need = (L + kRedBlock - 1) / kRedBlock;
    return (int)(need < (uint32_t)kRedMaxBlocks ? need : (uint32_t)kRedMaxBlocks);
}

// FP32 form: the nine sums of sample_fp32_kernel for the member's current
// state, computed once per (state, tau) and kept for the other call
template <int ND, bool UNI>
static auto sample_kernel_for(int R)
{
    return (R == 1)   ? sample_fp32_kernel<ND, 1, UNI>
           : (R == 2) ? sample_fp32_kernel<ND, 2, UNI>
           : (R == 3) ? sample_fp32_kernel<ND, 3, UNI>
                      : sample_fp32_kernel<ND, 4, UNI>;
}

// The nine sums of the member's current state (background, then the energy
// components), computed once per (state, tau) and kept for the other call:
// sample_fp64_kernel over the FP64 form's stored pairs, sample_fp32_kernel
// over the FP32 form's (phi, u) state.
static int sample_sums(Context *c, Member &m, double out[kSampQ], const Workspace *ws = nullptr)
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
    if (c->prec == AXS_FP32 && (R < 1 || R > 4)) {
        set_error("stencil radius %d is not supported", R);
        return AXS_ERR_UNSUPPORTED;
    }
    const uint32_t plane = (c->nd == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    // chunks along the slow axis: as few as still give a work-item to every
    // lane of four work-groups per CU; at most 64 planes per chunk, so that
    // the work-items of one round march in step and the rows they share are
    // still in L2
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
    const bool uni = plane % 64 == 0;       // every wave's columns share their slices
    if (c->prec == AXS_FP64) {
        auto k = (c->nd == 2) ? (uni ? sample_fp64_kernel<2, true> : sample_fp64_kernel<2, false>)
                              : (uni ? sample_fp64_kernel<3, true> : sample_fp64_kernel<3, false>);
        hipLaunchKernelGGL(k, dim3(nb), dim3(kRedBlock), 0, c->stream, N, CI, ncols,
                           (const dvec2 *)m.state[c->cur][PHI], (const dvec2 *)m.state[c->cur][VEL],
                           tau2, T03, lam_4tau2, part);
    } else {
        auto k = (c->nd == 2)
                     ? (uni ? sample_kernel_for<2, true>(R) : sample_kernel_for<2, false>(R))
                     : (uni ? sample_kernel_for<3, true>(R) : sample_kernel_for<3, false>(R));
        hipLaunchKernelGGL(k, dim3(nb), dim3(kRedBlock), 0, c->stream, N, CI, ncols,
                           (const fvec4 *)m.state[c->cur][PHI], (const fvec2 *)m.aux_ker, f,
                           ker_current(c, m) ? 0 : 1, tau2, T03, lam_4tau2, part);
    }
    hipLaunchKernelGGL(fold_kernel, dim3(kSampQ), dim3(256), 0, c->stream, part, nb, kSampQ, dout);
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

int run_nan_check(Context *c, Member &m, int64_t *first)
{
    if (!m.has_state) {
        set_error("member has no state: call axs_set_member_state first");
        return AXS_ERR_STATE;
    }
    AXS_TRY(ensure_scratch(c, 256));
    unsigned long long *d = (unsigned long long *)c->scratch;
    AXS_HIP_TRY(hipMemsetAsync(d, 0xff, sizeof *d, c->stream));
    const uint32_t L = (uint32_t)m.L;
    const int nb = red_blocks(L);
    if (c->prec == AXS_FP64)
        hipLaunchKernelGGL(nan_first_kernel<dvec2>, dim3(nb), dim3(kRedBlock), 0, c->stream, L,
                           (const dvec2 *)m.state[c->cur][PHI], d);
    else
        hipLaunchKernelGGL(nan_first_kernel<fvec4>, dim3(nb), dim3(kRedBlock), 0, c->stream, L,
                           (const fvec4 *)m.state[c->cur][PHI], d);
    AXS_HIP_TRY(hipGetLastError());
    unsigned long long h = 0;
    AXS_HIP_TRY(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    *first = (h == ~0ull) ? -1 : (int64_t)h;
    return AXS_OK;
}

int run_background(Context *c, Member &m, double out[6])
{
    double sums[kSampQ];
    AXS_TRY(sample_sums(c, m, sums));
    std::copy(sums, sums + 6, out);
    return AXS_OK;
}

int run_energy(Context *c, Member &m, double out[3], const Workspace *ws)
{
    double sums[kSampQ];
    AXS_TRY(sample_sums(c, m, sums, ws));
    energy_scale(c, sums + 6, out);
    return AXS_OK;
}

}  // namespace axs
