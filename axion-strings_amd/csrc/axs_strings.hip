// axs_strings.hip -- plaquette string finder (reference
// src/common/string_finding.cpp:3-86 on axion = atan2(phi1, phi2),
// src/standard/s_run.cpp:92-94) with ordered compaction.
//
// The reference appends flagged sites to thread-local vectors and merges
// them, so its list order depends on the OpenMP thread count; here the list
// is always in ascending site order (the single-thread order), which also
// fixes the 2D core count of s_run.cpp:101-109.
//
// Each site's angle (atan2(phi1, phi2) + pi) / (2 pi) is computed in FP64
// once -- the reference's axion array (s_run.cpp:92-94) scaled as in
// string_finding.cpp:14-16 -- rather than once per plaquette corner, and
// the plaquette test leaves one 64-bit ballot word (a bitmask of flagged
// sites) per 64 sites in site order.  3D: one pass over phi
// (strings_march_kernel) keeps the angles of two planes of a tile in LDS,
// then a small kernel counts the flags per block.  2D: pass 0 writes the
// angles to a scratch array, pass 1 evaluates the plaquettes from it and
// counts.  Then one work-group scans the block counts, and the writer
// re-reads the ballot words and writes the coordinates at their ordered
// positions.  The search is skipped when the same member's unchanged state
// was searched by the previous call.
#include <algorithm>

#include "axs_atan2.hpp"
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kStrBlock = 256;
constexpr int kStrIters = 8;                    // sites per thread per block
constexpr int kStrSpan = kStrBlock * kStrIters; // sites per block

// std::min(|b-a|, 1-|b-a|), string_finding.cpp:22-24
__device__ __forceinline__ double edge_theta(double a, double b)
{
    double d = fabs(b - a);
    double e = 1 - d;
    return (e < d) ? e : d;
}

// the three plaquettes of string_finding.cpp:49-62 at site (i, j, k) from
// the scaled angles of (i,j,k), (i+1,j,k), (i+1,j+1,k), (i,j+1,k),
// (i,j+1,k+1), (i,j,k+1); plaquette b uses corner (i, j+1, k+1)
__device__ __forceinline__ bool plaquettes3(double n0, double ni, double nij, double nj,
                                            double njk, double nk, double accept)
{
    double sa = edge_theta(n0, ni) + edge_theta(ni, nij) + edge_theta(nij, nj);
    double sb = edge_theta(n0, ni) + edge_theta(ni, njk) + edge_theta(njk, nk);
    double sc = edge_theta(n0, nj) + edge_theta(nj, njk) + edge_theta(njk, nk);
    return sa > accept || sb > accept || sc > accept;
}

// The reference's axion (s_run.cpp:94) scaled to [0, 1]
// (string_finding.cpp:14-16).  scaled_angle: libm's atan2 rounded
// correctly (the double-double evaluation of axs_atan2.hpp, bit for bit the
// value the host build and the oracle give); it decides every flag that
// depends on the last bits of an angle.  scaled_angle_fast: the device
// library's atan2 (within a few ulp of it), for the 2D angle pass whose
// plaquettes near the threshold are re-evaluated with scaled_angle.
__device__ __forceinline__ double scaled_angle(double a1, double a2)
{
    double a = atan2_dd(a1, a2);
    return (a + M_PI) / (2 * M_PI);
}

__device__ __forceinline__ double scaled_angle_fast(double a1, double a2)
{
    double a = atan2(a1, a2);
    return (a + M_PI) / (2 * M_PI);
}

template <class P>
__global__ void __launch_bounds__(256)
angle_kernel(uint32_t L, P phi, double *__restrict__ ang)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) {
        double a1, a2;
        phi.get(m, a1, a2);
        ang[m] = scaled_angle_fast(a1, a2);
    }
}

// the reference's axion array (s_run.cpp:92-94): atan2(phi1, phi2) of every
// site, rounded correctly
template <class P>
__global__ void __launch_bounds__(256)
axion_kernel(uint32_t L, P phi, double *__restrict__ out)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) {
        double a1, a2;
        phi.get(m, a1, a2);
        out[m] = atan2_dd(a1, a2);
    }
}

// A three-edge sum from the fast angles (each within a few 1e-16 of the
// correctly rounded one, so the sum within ~1e-15) decides the flag unless
// it lies within this margin of the threshold; there the four corners'
// angles are evaluated again with scaled_angle.
constexpr double kAngleMargin = 0x1p-30;

// 2D plaquette of site m (string_finding.cpp:17-20): corners (i,j),
// (i+1,j), (i+1,j+1), (i,j+1), three edges
template <class P>
__device__ __forceinline__ bool plaquette_flag2(const double *__restrict__ ang, const P &phi,
                                                uint32_t m, int N, double accept)
{
    const Site<2> s = decompose<2>(m, N);
    const uint32_t q2 = shifted<2>(s, 0, 1, N);
    const Site<2> s2 = decompose<2>(q2, N);
    const uint32_t q3 = shifted<2>(s2, 1, 1, N);
    const uint32_t q4 = shifted<2>(s, 1, 1, N);
    double n1 = ang[m], n2 = ang[q2], n3 = ang[q3], n4 = ang[q4];
    double sum = edge_theta(n1, n2) + edge_theta(n2, n3) + edge_theta(n3, n4);
    if (fabs(sum - accept) < kAngleMargin) {
        auto exact = [&](uint32_t q) {
            double a1, a2;
            phi.get(q, a1, a2);
            return scaled_angle(a1, a2);
        };
        n1 = exact(m);
        n2 = exact(q2);
        n3 = exact(q3);
        n4 = exact(q4);
        sum = edge_theta(n1, n2) + edge_theta(n2, n3) + edge_theta(n3, n4);
    }
    return sum > accept;
}

// 2D plaquette test of block `blk` (kStrSpan sites) of one lattice from
// its angles: one ballot word per wave and pass, the block's count in bcount
template <class P>
__device__ __forceinline__ void flag_block(int N, uint32_t L, const double *__restrict__ ang,
                                           const P &phi, double accept,
                                           uint64_t *__restrict__ words,
                                           uint32_t *__restrict__ bcount, uint32_t blk)
{
    __shared__ uint32_t wave_tot[kStrBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t total = 0;
    uint32_t base = blk * (uint32_t)kStrSpan;
    for (int it = 0; it < kStrIters; ++it) {
        uint32_t m = base + it * kStrBlock + threadIdx.x;
        bool f = (m < L) && plaquette_flag2(ang, phi, m, N, accept);
        uint64_t w = __ballot(f);
        uint32_t wi = (base + it * kStrBlock) / 64 + wave;
        if (lane == 0 && (uint64_t)wi * 64 < L) words[wi] = w;
        total += (uint32_t)__popcll(w);
    }
    if (lane == 0) wave_tot[wave] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kStrBlock / 64; ++w) t += wave_tot[w];
        bcount[blk] = t;
    }
}

template <class P>
__global__ void __launch_bounds__(kStrBlock)
strings_flag_kernel(int N, uint32_t L, const double *__restrict__ ang, P phi, double accept,
                    uint64_t *__restrict__ words, uint32_t *__restrict__ bcount)
{
    flag_block(N, L, ang, phi, accept, words, bcount, blockIdx.x);
}

// Ensemble-wide 2D finder: member r owns blocks [first_block, first_block +
// ceil(L / kStrSpan)) of one concatenated index space (its words start at
// word first_block * kStrSpan / 64, its angles at site_base); blk2mem maps
// a block to its member.
struct StrMember {
    const void *phi[2];      // the field of either ping-pong half
    uint64_t site_base;
    uint32_t L, first_block;
    int N, pad;
};

template <class P>
__global__ void __launch_bounds__(kStrBlock)
angle_batch_kernel(const StrMember *__restrict__ mem, const uint32_t *__restrict__ blk2mem,
                   int half, double *__restrict__ ang)
{
    const StrMember M = mem[blk2mem[blockIdx.x]];
    P phi;
    phi.p = (decltype(phi.p))M.phi[half];
    const uint32_t base = (blockIdx.x - M.first_block) * (uint32_t)kStrSpan;
    for (int it = 0; it < kStrIters; ++it) {
        const uint32_t m = base + it * kStrBlock + threadIdx.x;
        if (m < M.L) {
            double a1, a2;
            phi.get(m, a1, a2);
            ang[M.site_base + m] = scaled_angle_fast(a1, a2);
        }
    }
}

template <class P>
__global__ void __launch_bounds__(kStrBlock)
strings_flag_batch_kernel(const StrMember *__restrict__ mem, const uint32_t *__restrict__ blk2mem,
                          int half, const double *__restrict__ ang, double accept,
                          uint64_t *__restrict__ words, uint32_t *__restrict__ bcount)
{
    const StrMember M = mem[blk2mem[blockIdx.x]];
    P phi;
    phi.p = (decltype(phi.p))M.phi[half];
    const size_t wbase = (size_t)M.first_block * (kStrSpan / 64);
    flag_block(M.N, M.L, ang + M.site_base, phi, accept, words + wbase, bcount + M.first_block,
               blockIdx.x - M.first_block);
}

// 3D finder in one pass over phi.  A work-group of kSfW waves owns a tile
// of kSfTJ rows (j) x 64 columns (k) and marches along i over a chunk of
// planes.  Per plane it computes the angles of kSfTJ + 1 rows x 65 columns
// (the +1 row and column are the plaquettes' j+1 / k+1 corners) into LDS
// -- kSfTJ + 2 = 32 passes of 64 lanes (the 65th column is one pass of its
// own), four per wave -- and tests the plaquettes of plane i from the LDS
// planes i and i+1.  Each output row of 64 sites is one ballot word,
// stored whole when N is a multiple of 64 (the row is then one aligned
// word) and OR-ed into the zeroed words otherwise.
//
// Screening.  The LDS angles are FP32 (angle_turns_f of the components
// rounded to float).  A plaquette sum is a sum of three circular distances
// min(|b-a|, 1-|b-a|), each 1-Lipschitz in both angles and blind to the
// branch cut, so it is off the FP64 sum by at most six angle errors plus
// the float rounding of the sums: below 2e-6 turns (angles within 1.2e-7
// turns of the float inputs' exact ones, the inputs within half an ulp),
// against the margin kSfMargin = 2^-14 = 6.1e-5.  A
// plaquette above accept + margin is flagged, one below accept - margin
// is not; a site whose plaquettes are all undecided that way (or NaN), or
// that touches an input outside [1e-30, 1e30] in magnitude (where the
// float angle could be off: zeros, underflow, overflow, NaN), is decided
// by the FP64 angles of its six corners (scaled_angle, plaquettes3: the
// expressions of string_finding.cpp).  The flags are therefore those of
// the FP64 test, bit for bit; the FP64 path runs for a handful of sites
// per lattice.
constexpr int kSfW = 8;
constexpr int kSfTJ = 30;
constexpr float kSfMargin = 1.0f / 16384.0f;

struct SfGeom {
    int N, nTj, nTk, CI;
};

// float components of phi at m; false where the float angle is not
// trusted (both components below 1e-30 in magnitude, either above 1e30,
// or NaN)
__device__ __forceinline__ bool phi_float(const PhiF64 &p, uint32_t m, float &x, float &y)
{
    dvec2 v = p.p[m];
    double ax = fabs(v.x), ay = fabs(v.y);
    x = (float)v.x;
    y = (float)v.y;
    return (ax >= 1e-30 || ay >= 1e-30) && ax <= 1e30 && ay <= 1e30;
}
__device__ __forceinline__ bool phi_float(const PhiF32 &p, uint32_t m, float &x, float &y)
{
    fvec4 v = p.p[m];
    float ax = fabsf(v.x), ay = fabsf(v.y);
    x = v.x;
    y = v.y;
    return (ax >= 1e-30f || ay >= 1e-30f) && ax <= 1e30f && ay <= 1e30f;
}

// (atan2(x, y) + pi) / (2 pi) in FP32 for the screen: atan on [0, 1] by an
// odd polynomial t P(t^2) of degree 13 (coefficients fitted to atan by
// reweighted least squares, max error 2.5e-7 rad in exact arithmetic), the
// octant folded by |x| vs |y| and the signs.  Evaluated in float with the
// hardware reciprocal, the result is within 1.2e-7 turns (circular
// distance) of the exact scaled angle of the float inputs -- checked densely
// on the host, with the reciprocal one ulp off either way, by the tests.
// Where the exact angle is 0 or 1 (x = -0, y < 0) this may return the other
// end, which the circular distance does not see.
constexpr float kAtanC[7] = {0.9999961256980896f, -0.3331737518310547f, 0.19807878136634827f, -0.1323358416557312f, 0.07962814718484879f, -0.033608146011829376f, 0.006813105195760727f};

__device__ __forceinline__ float angle_turns_f(float x, float y)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mn * __builtin_amdgcn_rcpf(mx);
    const float s = t * t;
    float p = kAtanC[6];
#pragma unroll
    for (int k = 5; k >= 0; --k) p = __builtin_fmaf(p, s, kAtanC[k]);
    p = p * t;                                          // atan(min/max)
    p = (ax > ay) ? ((float)(M_PI / 2) - p) : p;
    p = (y < 0.0f) ? ((float)M_PI - p) : p;
    p = (x < 0.0f) ? -p : p;                            // atan2(x, y)
    return __builtin_fmaf(p, (float)(1.0 / (2 * M_PI)), 0.5f);
}

__device__ __forceinline__ float edge_theta_f(float a, float b)
{
    float d = fabsf(b - a);
    return fminf(d, 1.0f - d);
}

template <class P, bool ALIGNED>
__global__ void __launch_bounds__(64 * kSfW)
strings_march_kernel(SfGeom g, P phi, double accept, unsigned long long *__restrict__ words)
{
    __shared__ float A[2][kSfTJ + 1][65];
    // the wave index is uniform (a scalar register): row arithmetic is scalar
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int N = g.N;
    // XCD-aware order (work-groups go round-robin to the 8 XCDs): each XCD
    // takes a contiguous run of tiles, k fastest, so the 65th column and the
    // 31st row a tile reads from its neighbours are read on the same L2
    uint32_t b = blockIdx.x;
    if ((gridDim.x & 7u) == 0) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
    const int tk = (int)(b % (uint32_t)g.nTk);
    b /= (uint32_t)g.nTk;
    const int tj = (int)(b % (uint32_t)g.nTj);
    b /= (uint32_t)g.nTj;
    const int i0 = (int)b * g.CI;
    const int i1 = (i0 + g.CI < N) ? i0 + g.CI : N;
    const int j0 = tj * kSfTJ, k0 = tk * 64;
    const float hi = (float)accept + kSfMargin, lo = (float)accept - kSfMargin;

    // periodic wrap of an index below N + 64
    auto wrap = [N](int x) {
        while (x >= N) x -= N;
        return x;
    };
    auto site = [N](int i, int j, int k) {
        return ((uint32_t)i * (uint32_t)N + (uint32_t)j) * (uint32_t)N + (uint32_t)k;
    };
    auto angle_d = [&](int i, int j, int k) {
        double a1, a2;
        phi.get(site(i, j, k), a1, a2);
        return scaled_angle(a1, a2);
    };
    // the four angle passes of this wave per plane: pass q = wave + kSfW t
    // below kSfTJ + 1 is row j0 + q (column k0 + lane), pass kSfTJ + 1 is
    // column k0 + 64 of row j0 + lane (lanes up to kSfTJ); in-plane offsets
    // j N + k and LDS slots, fixed for the whole march
    constexpr int kPass = (kSfTJ + 2) / kSfW;
    static_assert(kPass * kSfW == kSfTJ + 2, "angle passes must fill the waves evenly");
    uint32_t off[kPass];
    int slot[kPass];
    bool live[kPass];
    for (int t = 0; t < kPass; ++t) {
        const int q = wave + kSfW * t;
        const bool row = q <= kSfTJ;
        const int j = wrap(j0 + (row ? q : lane)), k = wrap(k0 + (row ? lane : 64));
        off[t] = (uint32_t)j * (uint32_t)N + (uint32_t)k;
        slot[t] = row ? q * 65 + lane : lane * 65 + 64;
        live[t] = row || lane <= kSfTJ;
    }
    const uint32_t NN = (uint32_t)N * (uint32_t)N;
    // raw components of one plane (in flight while the previous plane is
    // processed) and their FP32 angles into LDS plane s (NaN where the float
    // angle is not trusted)
    struct Raw {
        float x[kPass], y[kPass];
        bool ok[kPass];
    };
    auto fetch = [&](int i, Raw &r) {
        for (int t = 0; t < kPass; ++t)
            r.ok[t] = live[t] && phi_float(phi, (uint32_t)i * NN + off[t], r.x[t], r.y[t]);
    };
    auto store = [&](const Raw &r, int s) {
        float *a = &A[s][0][0];
        for (int t = 0; t < kPass; ++t) {
            if (!live[t]) continue;
            const float th = angle_turns_f(r.x[t], r.y[t]);
            a[slot[t]] = r.ok[t] ? th : __builtin_nanf("");
        }
    };

    int s = 0;
    Raw cur, nxt;
    fetch(i0, cur);
    store(cur, 0);
    fetch(i0 + 1 == N ? 0 : i0 + 1, cur);
    for (int i = i0; i < i1; ++i) {
        const int ip = (i + 1 == N) ? 0 : i + 1;
        if (i + 2 <= i1) fetch(i + 2 >= N ? i + 2 - N : i + 2, nxt);
        store(cur, s ^ 1);
        __syncthreads();
        for (int r = wave; r < kSfTJ; r += kSfW) {
            const int j = j0 + r, k = k0 + lane;
            if (j >= N) break;                          // uniform over the wave
            const float n0 = A[s][r][lane], ni = A[s ^ 1][r][lane], nij = A[s ^ 1][r + 1][lane];
            const float nj = A[s][r + 1][lane], njk = A[s][r + 1][lane + 1];
            const float nk = A[s][r][lane + 1];
            const float sa = edge_theta_f(n0, ni) + edge_theta_f(ni, nij) + edge_theta_f(nij, nj);
            const float sb = edge_theta_f(n0, ni) + edge_theta_f(ni, njk) + edge_theta_f(njk, nk);
            const float sc = edge_theta_f(n0, nj) + edge_theta_f(nj, njk) + edge_theta_f(njk, nk);
            const bool in = k < N;
            bool f = in && (sa > hi || sb > hi || sc > hi);
            const bool undecided = in && !f && !(sa < lo && sb < lo && sc < lo);
            if (__ballot(undecided)) {
                if (undecided) {
                    const int jp = wrap(j + 1), kp = wrap(k + 1);
                    f = plaquettes3(angle_d(i, j, k), angle_d(ip, j, k), angle_d(ip, jp, k),
                                    angle_d(i, jp, k), angle_d(i, jp, kp), angle_d(i, j, kp),
                                    accept);
                }
            }
            const unsigned long long w = __ballot(f);
            const uint64_t m0 = ((uint64_t)i * (uint64_t)N + (uint64_t)j) * (uint64_t)N +
                                (uint64_t)k0;
            if (ALIGNED) {
                if (lane == 0) words[m0 >> 6] = w;
            } else if (lane == 0 && w) {
                const unsigned sh = (unsigned)(m0 & 63);
                atomicOr(&words[m0 >> 6], w << sh);
                if (sh && (w >> (64 - sh))) atomicOr(&words[(m0 >> 6) + 1], w >> (64 - sh));
            }
        }
        __syncthreads();
        s ^= 1;
        cur = nxt;
    }
}

// flagged sites per block of kStrSpan sites (32 words), one word per lane
__global__ void __launch_bounds__(256)
block_counts_kernel(const uint64_t *__restrict__ words, uint64_t nwords, uint32_t *__restrict__ bcount)
{
    const uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint32_t c = (w < nwords) ? (uint32_t)__popcll(words[w]) : 0u;
    for (int o = 16; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 31) == 0 && w < nwords) bcount[w >> 5] = c;
}

// Exclusive scan of the n block counts into off[0..n), off[n] = total, in
// three launches: tile_scan_kernel scans tiles of kScanTile counts (four
// per work-item) to tile-relative offsets and tile sums; top_scan_kernel
// scans the tile sums (one per work-item: lattices hold fewer than 2^31
// sites, so n <= 2^20 and there are at most 256 tiles) and writes the
// total; add_tile_offsets_kernel adds each tile's offset.
constexpr int kScanTile = 4096;

// exclusive prefix of v over the 1024 work-items; *total gets the sum
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *lds, T *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    if (wave == 0) {
        T w = (lane < 16) ? lds[lane] : (T)0;
        T z = w;
        for (int o = 1; o < 16; o <<= 1) {
            T y = __shfl_up(z, o, 64);
            if (lane >= o) z += y;
        }
        if (lane < 16) lds[lane] = z - w;                // exclusive over waves
        if (lane == 15) lds[16] = z;
    }
    __syncthreads();
    *total = lds[16];
    return lds[wave] + x - v;
}

__global__ void __launch_bounds__(1024)
tile_scan_kernel(const uint32_t *__restrict__ cnt, uint64_t *__restrict__ off,
                 uint64_t *__restrict__ tsum, uint32_t n)
{
    __shared__ uint32_t lds[17];
    const uint32_t e0 = blockIdx.x * (uint32_t)kScanTile + 4u * threadIdx.x;
    uint32_t c[4], s = 0;
    for (int u = 0; u < 4; ++u) {
        c[u] = (e0 + u < n) ? cnt[e0 + u] : 0u;
        s += c[u];
    }
    uint32_t tot;
    uint32_t acc = block_exclusive_scan<uint32_t>(s, lds, &tot);
    for (int u = 0; u < 4; ++u) {
        if (e0 + u < n) off[e0 + u] = acc;
        acc += c[u];
    }
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024)
top_scan_kernel(uint64_t *__restrict__ tsum, uint32_t ntiles, uint64_t *__restrict__ total)
{
    __shared__ uint64_t lds[17];
    const uint64_t v = (threadIdx.x < ntiles) ? tsum[threadIdx.x] : 0ull;
    uint64_t tot;
    const uint64_t x = block_exclusive_scan<uint64_t>(v, lds, &tot);
    if (threadIdx.x < ntiles) tsum[threadIdx.x] = x;
    if (threadIdx.x == 0) *total = tot;
}

__global__ void __launch_bounds__(256)
add_tile_offsets_kernel(uint64_t *__restrict__ off, const uint64_t *__restrict__ toff, uint32_t n)
{
    const uint32_t e = blockIdx.x * 256u + threadIdx.x;
    if (e < n) off[e] += toff[e / (uint32_t)kScanTile];
}

template <int ND>
__global__ void __launch_bounds__(kStrBlock)
strings_write_kernel(int N, uint32_t L, const uint64_t *__restrict__ words,
                     const uint64_t *__restrict__ off, int32_t *__restrict__ coords, uint64_t cap)
{
    __shared__ uint32_t wave_cnt[kStrBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t pos = off[blockIdx.x];
    uint32_t base = blockIdx.x * (uint32_t)kStrSpan;
    for (int it = 0; it < kStrIters; ++it) {
        uint32_t wbase = base + it * kStrBlock;
        uint32_t wi = wbase / 64 + wave;
        uint64_t w = ((uint64_t)wi * 64 < L) ? words[wi] : 0ull;
        if (lane == 0) wave_cnt[wave] = (uint32_t)__popcll(w);
        __syncthreads();
        uint64_t before = pos;
        for (int v = 0; v < wave; ++v) before += wave_cnt[v];
        uint64_t below = (lane == 0) ? 0ull : (w & ((~0ull) >> (64 - lane)));
        if ((w >> lane) & 1ull) {
            uint64_t p = before + (uint64_t)__popcll(below);
            uint32_t m = wbase + threadIdx.x;
            if (p < cap) {
                Site<ND> s = decompose<ND>(m, N);
                for (int d = 0; d < ND; ++d) coords[p * ND + d] = s.c[d];
            }
        }
        uint32_t tot = 0;
        for (int v = 0; v < kStrBlock / 64; ++v) tot += wave_cnt[v];
        pos += tot;
        __syncthreads();
    }
}

// 2D core count, s_run.cpp:101-109: list-adjacent pairs differing by (1,0)
// or (0,1)
__global__ void __launch_bounds__(256)
adjacent_pairs_kernel(const int32_t *__restrict__ coords, uint64_t n,
                      unsigned long long *__restrict__ out)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    unsigned c = 0;
    if (i + 1 < n) {
        int dy = coords[2 * (i + 1) + 1] - coords[2 * i + 1];
        int dx = coords[2 * (i + 1)] - coords[2 * i];
        c = (dy == 0 && dx == 1) + (dy == 1 && dx == 0);
    }
    // wave reduce then one atomic per wave (integer: order-independent)
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, (unsigned long long)c);
}

// Flagged sites of one member from its ballot words, and in 2D the
// list-adjacent pairs of s_run.cpp:101-109 without building the list: in
// ascending site order (m = i + N j) the entry after a flagged site m is the
// next flagged site, and the pair counts when that is m + 1 in the same row
// (di = 1, dj = 0) or m + N (di = 0, dj = 1).  Adds to cnt[0] and cnt[1].
// flagged sites in word w of one lattice's words, and (2D) the pairs of
// list-adjacent flagged sites whose first site lies in w
template <int ND>
__device__ __forceinline__ void count_word(const uint64_t *__restrict__ words, uint64_t nwords,
                                           int N, uint64_t w, unsigned long long &n,
                                           unsigned long long &pairs)
{
    uint64_t bits = words[w];
    n += (unsigned long long)__popcll(bits);
    if (ND == 2) {
        while (bits) {
            const uint64_t m = w * 64 + (uint64_t)(__ffsll((unsigned long long)bits) - 1);
            bits &= bits - 1;
            uint64_t next = ~0ull;                   // next flagged site after m
            if (bits) {
                next = w * 64 + (uint64_t)(__ffsll((unsigned long long)bits) - 1);
            } else {
                for (uint64_t v = w + 1; v < nwords && v * 64 <= m + (uint64_t)N; ++v) {
                    const uint64_t x = words[v];
                    if (x) {
                        next = v * 64 + (uint64_t)(__ffsll((unsigned long long)x) - 1);
                        break;
                    }
                }
            }
            if ((next == m + 1 && m % (uint64_t)N != (uint64_t)(N - 1)) ||
                next == m + (uint64_t)N)
                ++pairs;
        }
    }
}

template <int ND>
__global__ void __launch_bounds__(256)
member_count_kernel(const uint64_t *__restrict__ words, uint32_t L, int N,
                    unsigned long long *__restrict__ cnt)
{
    __shared__ unsigned long long part[2][4];
    const uint64_t nwords = ((uint64_t)L + 63) / 64;
    unsigned long long n = 0, pairs = 0;
    // grid-stride over the words, one pair of atomics per work-group
    for (uint64_t w = blockIdx.x * 256ull + threadIdx.x; w < nwords;
         w += (uint64_t)gridDim.x * 256ull)
        count_word<ND>(words, nwords, N, w, n, pairs);
    for (int o = 32; o > 0; o >>= 1) {
        n += __shfl_down(n, o, 64);
        pairs += __shfl_down(pairs, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][wave] = n;
        part[1][wave] = pairs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long tn = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        const unsigned long long tp = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        if (tn) atomicAdd(&cnt[0], tn);
        if (tp) atomicAdd(&cnt[1], tp);
    }
}

// the ensemble's 2D counts: eight blocks per work-group, each block's 32
// words on 32 lanes (one member per block), two atomics per block
__global__ void __launch_bounds__(256)
count_batch_kernel(const StrMember *__restrict__ mem, const uint32_t *__restrict__ blk2mem,
                   uint32_t nblocks, const uint64_t *__restrict__ words,
                   unsigned long long *__restrict__ cnt)
{
    const uint32_t blk = blockIdx.x * 8u + (threadIdx.x >> 5);
    unsigned long long n = 0, pairs = 0;
    uint32_t r = 0;
    if (blk < nblocks) {
        r = blk2mem[blk];
        const StrMember M = mem[r];
        const uint64_t nwords = ((uint64_t)M.L + 63) / 64;
        const uint64_t w = (uint64_t)(blk - M.first_block) * (kStrSpan / 64) + (threadIdx.x & 31);
        if (w < nwords)
            count_word<2>(words + (size_t)M.first_block * (kStrSpan / 64), nwords, M.N, w, n,
                          pairs);
    }
    for (int o = 16; o > 0; o >>= 1) {
        n += __shfl_xor(n, o, 64);
        pairs += __shfl_xor(pairs, o, 64);
    }
    if ((threadIdx.x & 31) == 0 && blk < nblocks) {
        if (n) atomicAdd(&cnt[2 * r], n);
        if (pairs) atomicAdd(&cnt[2 * r + 1], pairs);
    }
}

static double accept_threshold(int thr)
{
    float f = 0.5f * (1.0f - (float)thr / 100.0f);   // string_finding.cpp:8, float
    return (double)f;
}

// one-pass 3D finder: ballot words of every site, then the per-block
// counts the scan and the coordinate writer use
static int launch_strings_march(Context *c, int N, uint32_t L, const void *phi, double accept,
                                uint64_t *words, uint32_t *bcount)
{
    SfGeom g;
    g.N = N;
    g.nTj = (N + kSfTJ - 1) / kSfTJ;
    g.nTk = (N + 63) / 64;
    const uint64_t tiles = (uint64_t)g.nTj * (uint64_t)g.nTk;
    // longest chunk (64 planes down to 4) that still gives two work-groups
    // per CU: the first plane of each chunk is the previous chunk's last
    int CI = 64;
    while (CI > 4 && tiles * (uint64_t)((N + CI - 1) / CI) < 2 * (uint64_t)c->num_cu) CI >>= 1;
    g.CI = CI;
    const uint32_t grid = (uint32_t)(tiles * (uint64_t)((N + CI - 1) / CI));
    const uint64_t nwords = ((uint64_t)L + 63) / 64;
    unsigned long long *w = (unsigned long long *)words;
    if (N % 64 == 0) {
        if (c->prec == AXS_FP64)
            hipLaunchKernelGGL((strings_march_kernel<PhiF64, true>), dim3(grid), dim3(64 * kSfW),
                               0, c->stream, g, PhiF64{(const dvec2 *)phi}, accept, w);
        else
            hipLaunchKernelGGL((strings_march_kernel<PhiF32, true>), dim3(grid), dim3(64 * kSfW),
                               0, c->stream, g, PhiF32{(const fvec4 *)phi}, accept, w);
    } else {
        AXS_HIP_TRY(hipMemsetAsync(words, 0, nwords * sizeof(uint64_t), c->stream));
        if (c->prec == AXS_FP64)
            hipLaunchKernelGGL((strings_march_kernel<PhiF64, false>), dim3(grid), dim3(64 * kSfW),
                               0, c->stream, g, PhiF64{(const dvec2 *)phi}, accept, w);
        else
            hipLaunchKernelGGL((strings_march_kernel<PhiF32, false>), dim3(grid), dim3(64 * kSfW),
                               0, c->stream, g, PhiF32{(const fvec4 *)phi}, accept, w);
    }
    hipLaunchKernelGGL(block_counts_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0,
                       c->stream, words, nwords, bcount);
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

// Layout of the finder's device buffer for lattices of up to L sites:
// ballot words, block counts, block offsets (+ total), a 256-byte slot for
// the 2D pair count, the scan's tile sums, then two counters per member for
// the ensemble-wide count.
struct StrLayout {
    uint32_t nblocks, ntiles;
    size_t words_b, cnt_b, off_b, tail_b, counters_b;
    StrLayout(uint32_t L, size_t nmembers)
    {
        auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
        nblocks = (L + kStrSpan - 1) / kStrSpan;
        ntiles = (nblocks + kScanTile - 1) / kScanTile;
        words_b = ((size_t)nblocks * kStrSpan / 64) * 8;
        cnt_b = align((size_t)nblocks * 4);
        off_b = align((size_t)(nblocks + 1) * 8);
        tail_b = 256 + align((size_t)ntiles * 8);
        counters_b = align(nmembers * 16);
    }
    size_t bytes() const { return words_b + cnt_b + off_b + tail_b + counters_b; }
};

// ballot words and block counts of member m (the search itself)
static int search_member(Context *c, Member &m, uint64_t *words, uint32_t *bcount)
{
    const uint32_t L = (uint32_t)m.L;
    const uint32_t nblocks = (L + kStrSpan - 1) / kStrSpan;
    // the angles need phi only (no FP32 velocity reconstruction)
    const void *phi = m.state[c->cur][PHI];
    const double accept = accept_threshold(c->p.thr);
    if (c->nd == 3) return launch_strings_march(c, m.N, L, phi, accept, words, bcount);
    AXS_TRY(ensure_scratch(c, (size_t)L * sizeof(double)));
    double *ang = (double *)c->scratch;
    const unsigned ga = (L + 255) / 256;
    if (c->prec == AXS_FP64) {
        const PhiF64 p{(const dvec2 *)phi};
        hipLaunchKernelGGL(angle_kernel<PhiF64>, dim3(ga), dim3(256), 0, c->stream, L, p, ang);
        hipLaunchKernelGGL(strings_flag_kernel<PhiF64>, dim3(nblocks), dim3(kStrBlock), 0,
                           c->stream, m.N, L, ang, p, accept, words, bcount);
    } else {
        const PhiF32 p{(const fvec4 *)phi};
        hipLaunchKernelGGL(angle_kernel<PhiF32>, dim3(ga), dim3(256), 0, c->stream, L, p, ang);
        hipLaunchKernelGGL(strings_flag_kernel<PhiF32>, dim3(nblocks), dim3(kStrBlock), 0,
                           c->stream, m.N, L, ang, p, accept, words, bcount);
    }
    AXS_HIP_TRY(hipGetLastError());
    return AXS_OK;
}

int run_axion(Context *c, Member &m, double *out)
{
    if (!m.has_state) {
        set_error("member has no state: call axs_set_member_state first");
        return AXS_ERR_STATE;
    }
    const uint32_t L = (uint32_t)m.L;
    AXS_TRY(ensure_scratch(c, (size_t)L * sizeof(double)));
    double *d = (double *)c->scratch;
    const void *phi = m.state[c->cur][PHI];
    const unsigned g = (L + 255) / 256;
    if (c->prec == AXS_FP64)
        hipLaunchKernelGGL(axion_kernel<PhiF64>, dim3(g), dim3(256), 0, c->stream, L,
                           PhiF64{(const dvec2 *)phi}, d);
    else
        hipLaunchKernelGGL(axion_kernel<PhiF32>, dim3(g), dim3(256), 0, c->stream, L,
                           PhiF32{(const fvec4 *)phi}, d);
    AXS_HIP_TRY(hipGetLastError());
    return copy_to_host(c, out, d, (size_t)L * sizeof(double));
}

int run_strings(Context *c, Member &m, int32_t *coords, int64_t cap, int64_t *count,
                int64_t *num_cores)
{
    if (!m.has_state) {
        set_error("member has no state: call axs_set_member_state first");
        return AXS_ERR_STATE;
    }
    const int member = (int)(&m - c->members.data());
    const uint32_t L = (uint32_t)m.L;
    const StrLayout lay(L, c->members.size());
    const uint32_t nblocks = lay.nblocks, ntiles = lay.ntiles;
    const bool cached = c->str_member == member && c->str_version == m.version && c->strbuf &&
                        c->strbuf_bytes >= lay.bytes();
    if (!cached) {
        c->str_member = -1;
        AXS_TRY(ensure_device(c, &c->strbuf, &c->strbuf_bytes, lay.bytes()));
    }
    char *sp = (char *)c->strbuf;
    uint64_t *words = (uint64_t *)sp;
    uint32_t *bcount = (uint32_t *)(sp + lay.words_b);
    uint64_t *off = (uint64_t *)(sp + lay.words_b + lay.cnt_b);
    unsigned long long *pairs = (unsigned long long *)(sp + lay.words_b + lay.cnt_b + lay.off_b);
    uint64_t *tsum = (uint64_t *)(sp + lay.words_b + lay.cnt_b + lay.off_b + 256);

    auto scan = [&]() {
        hipLaunchKernelGGL(tile_scan_kernel, dim3(ntiles), dim3(1024), 0, c->stream, bcount, off,
                           tsum, nblocks);
        hipLaunchKernelGGL(top_scan_kernel, dim3(1), dim3(1024), 0, c->stream, tsum, ntiles,
                           off + nblocks);
        hipLaunchKernelGGL(add_tile_offsets_kernel, dim3((nblocks + 255) / 256), dim3(256), 0,
                           c->stream, off, tsum, nblocks);
        return hipGetLastError();
    };
    uint64_t total = 0;
    if (cached) {
        total = c->str_total;
    } else {
        AXS_TRY(search_member(c, m, words, bcount));
        AXS_HIP_TRY(scan());
        AXS_HIP_TRY(hipMemcpyAsync(&total, off + nblocks, 8, hipMemcpyDeviceToHost, c->stream));
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
        c->str_member = member;
        c->str_version = m.version;
        c->str_total = total;
        c->str_scanned = true;
    }
    *count = (int64_t)total;

    bool need_list = (coords != nullptr && cap > 0) || (num_cores != nullptr && c->nd == 2);
    if (!need_list) {
        if (num_cores) *num_cores = (int64_t)total;
        return (coords == nullptr || cap >= (int64_t)total) ? AXS_OK : AXS_ERR_CAPACITY;
    }
    if (!c->str_scanned) {         // words cached by a count: offsets for the writer
        AXS_HIP_TRY(scan());
        c->str_scanned = true;
    }

    // coordinate list in its own buffer (scratch keeps words / offsets)
    size_t list_b = (size_t)total * c->nd * sizeof(int32_t) + 256;
    AXS_TRY(ensure_device(c, &c->listbuf, &c->listbuf_bytes, list_b));
    int32_t *dlist = (int32_t *)c->listbuf;

    if (total > 0) {
        auto kw = (c->nd == 2) ? strings_write_kernel<2> : strings_write_kernel<3>;
        hipLaunchKernelGGL(kw, dim3(nblocks), dim3(kStrBlock), 0, c->stream, m.N, L, words, off,
                           dlist, (uint64_t)total);
    }
    AXS_HIP_TRY(hipGetLastError());

    if (num_cores) {
        if (c->nd == 2 && total > 1) {
            AXS_HIP_TRY(hipMemsetAsync(pairs, 0, 8, c->stream));
            unsigned g = (unsigned)((total + 255) / 256);
            hipLaunchKernelGGL(adjacent_pairs_kernel, dim3(g), dim3(256), 0, c->stream, dlist,
                               (uint64_t)total, pairs);
            AXS_HIP_TRY(hipGetLastError());
            unsigned long long np = 0;
            AXS_HIP_TRY(hipMemcpyAsync(&np, pairs, 8, hipMemcpyDeviceToHost, c->stream));
            AXS_HIP_TRY(hipStreamSynchronize(c->stream));
            *num_cores = (int64_t)total - (int64_t)np;
        } else {
            *num_cores = (int64_t)total;
        }
    }

    if (coords && cap > 0) {
        int64_t ncopy = (cap < (int64_t)total) ? cap : (int64_t)total;
        if (ncopy > 0)
            AXS_TRY(copy_to_host(c, coords, dlist, (size_t)ncopy * c->nd * sizeof(int32_t)));
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return (coords == nullptr || cap >= (int64_t)total) ? AXS_OK : AXS_ERR_CAPACITY;
}

// axs_string_counts of a 2D ensemble: three launches for all members (the
// angles, the plaquette test and the counts over the concatenated block
// space), one copy, one synchronisation
static int run_string_counts_2d_batch(Context *c, int64_t *n_plaquettes, int64_t *num_cores)
{
    const size_t nm = c->members.size();
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<StrMember> tab(nm);
    uint32_t nblocks = 0;
    uint64_t sites = 0;
    for (size_t r = 0; r < nm; ++r) {
        const Member &m = c->members[r];
        tab[r] = StrMember{{m.state[0][PHI], m.state[1][PHI]}, sites, (uint32_t)m.L, nblocks,
                           m.N, 0};
        nblocks += ((uint32_t)m.L + kStrSpan - 1) / kStrSpan;
        sites += (uint64_t)m.L;
    }
    const size_t tab_b = align(nm * sizeof(StrMember));
    if (!c->strtab) {          // members and their buffers are fixed: built once
        std::vector<uint32_t> blk(nblocks);
        for (size_t r = 0; r < nm; ++r) {
            const uint32_t nb = (tab[r].L + kStrSpan - 1) / kStrSpan;
            for (uint32_t b = 0; b < nb; ++b) blk[tab[r].first_block + b] = (uint32_t)r;
        }
        AXS_HIP_TRY(hipMalloc(&c->strtab, tab_b + (size_t)nblocks * 4));
        AXS_HIP_TRY(hipMemcpy(c->strtab, tab.data(), nm * sizeof(StrMember),
                              hipMemcpyHostToDevice));
        AXS_HIP_TRY(hipMemcpy((char *)c->strtab + tab_b, blk.data(), (size_t)nblocks * 4,
                              hipMemcpyHostToDevice));
        c->strtab_blocks = nblocks;
    }
    const StrMember *dmem = (const StrMember *)c->strtab;
    const uint32_t *dblk = (const uint32_t *)((char *)c->strtab + tab_b);
    const size_t words_b = (size_t)nblocks * (kStrSpan / 64) * 8, cnt_b = align((size_t)nblocks * 4);
    c->str_member = -1;
    AXS_TRY(ensure_device(c, &c->strbuf, &c->strbuf_bytes, words_b + cnt_b + align(nm * 16)));
    AXS_TRY(ensure_scratch(c, (size_t)sites * sizeof(double)));
    char *sp = (char *)c->strbuf;
    uint64_t *words = (uint64_t *)sp;
    uint32_t *bcount = (uint32_t *)(sp + words_b);
    unsigned long long *counters = (unsigned long long *)(sp + words_b + cnt_b);
    double *ang = (double *)c->scratch;
    AXS_HIP_TRY(hipMemsetAsync(counters, 0, nm * 16, c->stream));
    const double accept = accept_threshold(c->p.thr);
    if (c->prec == AXS_FP64) {
        hipLaunchKernelGGL(angle_batch_kernel<PhiF64>, dim3(nblocks), dim3(kStrBlock), 0,
                           c->stream, dmem, dblk, c->cur, ang);
        hipLaunchKernelGGL(strings_flag_batch_kernel<PhiF64>, dim3(nblocks), dim3(kStrBlock), 0,
                           c->stream, dmem, dblk, c->cur, ang, accept, words, bcount);
    } else {
        hipLaunchKernelGGL(angle_batch_kernel<PhiF32>, dim3(nblocks), dim3(kStrBlock), 0,
                           c->stream, dmem, dblk, c->cur, ang);
        hipLaunchKernelGGL(strings_flag_batch_kernel<PhiF32>, dim3(nblocks), dim3(kStrBlock), 0,
                           c->stream, dmem, dblk, c->cur, ang, accept, words, bcount);
    }
    hipLaunchKernelGGL(count_batch_kernel, dim3((nblocks + 7) / 8), dim3(256), 0, c->stream, dmem,
                       dblk, nblocks, words, counters);
    AXS_HIP_TRY(hipGetLastError());
    std::vector<unsigned long long> h(2 * nm);
    AXS_HIP_TRY(hipMemcpyAsync(h.data(), counters, nm * 16, hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t r = 0; r < nm; ++r) {
        if (n_plaquettes) n_plaquettes[r] = (int64_t)h[2 * r];
        if (num_cores) num_cores[r] = (int64_t)(h[2 * r] - h[2 * r + 1]);
    }
    return AXS_OK;
}

// axs_string_counts: every member searched in turn on the stream, counted
// from its ballot words into its own counters, one copy and one
// synchronisation for the ensemble (the shared buffers are sized for the
// largest member first, so nothing is re-allocated between the launches).
int run_string_counts(Context *c, int64_t *n_plaquettes, int64_t *num_cores)
{
    uint32_t Lmax = 0;
    for (const Member &m : c->members) {
        if (!m.has_state) {
            set_error("member has no state: call axs_set_member_state first");
            return AXS_ERR_STATE;
        }
        Lmax = (uint32_t)m.L > Lmax ? (uint32_t)m.L : Lmax;
    }
    const size_t nm = c->members.size();
    if (c->nd == 2 && nm > 1) return run_string_counts_2d_batch(c, n_plaquettes, num_cores);
    const StrLayout lay(Lmax, nm);
    c->str_member = -1;
    AXS_TRY(ensure_device(c, &c->strbuf, &c->strbuf_bytes, lay.bytes()));
    if (c->nd == 2) AXS_TRY(ensure_scratch(c, (size_t)Lmax * sizeof(double)));
    char *sp = (char *)c->strbuf;
    uint64_t *words = (uint64_t *)sp;
    unsigned long long *counters =
        (unsigned long long *)(sp + lay.words_b + lay.cnt_b + lay.off_b + lay.tail_b);
    AXS_HIP_TRY(hipMemsetAsync(counters, 0, nm * 16, c->stream));
    for (size_t r = 0; r < nm; ++r) {
        Member &m = c->members[r];
        // block counts where the member's own layout keeps them (below the
        // counters of the largest layout), so the last member's search can
        // be reused by a coordinate list
        const StrLayout lr((uint32_t)m.L, nm);
        AXS_TRY(search_member(c, m, words, (uint32_t *)(sp + lr.words_b)));
        const uint32_t L = (uint32_t)m.L;
        const uint64_t nw = ((uint64_t)L + 63) / 64;
        const unsigned g = (unsigned)std::min<uint64_t>((nw + 255) / 256, 4ull * c->num_cu);
        auto k = (c->nd == 2) ? member_count_kernel<2> : member_count_kernel<3>;
        hipLaunchKernelGGL(k, dim3(g), dim3(256), 0, c->stream, words, L, m.N, counters + 2 * r);
        AXS_HIP_TRY(hipGetLastError());
    }
    std::vector<unsigned long long> h(2 * nm);
    AXS_HIP_TRY(hipMemcpyAsync(h.data(), counters, nm * 16, hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t r = 0; r < nm; ++r) {
        if (n_plaquettes) n_plaquettes[r] = (int64_t)h[2 * r];
        if (num_cores) num_cores[r] = (int64_t)(h[2 * r] - h[2 * r + 1]);
    }
    // the words and block counts of the last member stay valid: a
    // coordinate list of it next needs only the scan
    c->str_member = (int)nm - 1;
    c->str_version = c->members[nm - 1].version;
    c->str_total = h[2 * (nm - 1)];
    c->str_scanned = false;
    return AXS_OK;
}

}  // namespace axs
