// axs_strings.hip -- plaquette string finder (reference
// src/common/string_finding.cpp:3-86 on axion = atan2(phi1, phi2),
// src/standard/s_run.cpp:92-94) with ordered compaction.
//
// The reference appends flagged sites to thread-local vectors and merges
// them, so its list order depends on the OpenMP thread count; here the list
// is always in ascending site order (the single-thread order), which also
// fixes the 2D core count of s_run.cpp:101-109.
//
// Pass 0 computes each site's angle (atan2(phi1, phi2) + pi) / (2 pi) in
// FP64 once -- the reference's axion array (s_run.cpp:92-94) scaled as in
// string_finding.cpp:14-16 -- so the FP64 atan2 runs once per site rather
// than once per plaquette corner.  Pass 1 evaluates the plaquette test per
// site from those angles and stores one 64-bit ballot word per wave (a
// bitmask of flagged sites) plus a count per block; pass 2 scans the block
// counts; pass 3 re-reads the ballot words and writes the coordinates at
// their ordered positions.  Passes 0-2 are skipped when the same member's
// unchanged state was searched by the previous call.
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kStrBlock = 256;
constexpr int kStrIters = 8;                    // sites per thread per block
constexpr int kStrSpan = kStrBlock * kStrIters; // sites per block

template <class P>
__global__ void __launch_bounds__(256)
angle_kernel(uint32_t L, P phi, double *__restrict__ ang)
{
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) {
        double a1, a2;
        phi.get(m, a1, a2);
        double a = atan2(a1, a2);
        ang[m] = (a + M_PI) / (2 * M_PI);
    }
}

// std::min(|b-a|, 1-|b-a|), string_finding.cpp:22-24
__device__ __forceinline__ double edge_theta(double a, double b)
{
    double d = fabs(b - a);
    double e = 1 - d;
    return (e < d) ? e : d;
}

template <int ND>
__device__ __forceinline__ bool plaquette_flag(const double *__restrict__ ang, uint32_t m, int N,
                                               double accept)
{
    Site<ND> s = decompose<ND>(m, N);
    if (ND == 2) {
        // corners (i,j), (i+1,j), (i+1,j+1), (i,j+1): string_finding.cpp:17-20
        uint32_t q2 = shifted<ND>(s, 0, 1, N);
        Site<ND> s2 = decompose<ND>(q2, N);
        uint32_t q3 = shifted<ND>(s2, 1, 1, N);
        uint32_t q4 = shifted<ND>(s, 1, 1, N);
        double n1 = ang[m];
        double n2 = ang[q2];
        double n3 = ang[q3];
        double n4 = ang[q4];
        double sum = edge_theta(n1, n2) + edge_theta(n2, n3) + edge_theta(n3, n4);
        return sum > accept;
    } else {
        // string_finding.cpp:49-62; plaquette b uses corner (i, j+1, k+1)
        uint32_t qi = shifted<ND>(s, 0, 1, N);                     // (i+1, j,   k)
        Site<ND> si = decompose<ND>(qi, N);
        uint32_t qij = shifted<ND>(si, 1, 1, N);                   // (i+1, j+1, k)
        uint32_t qj = shifted<ND>(s, 1, 1, N);                     // (i,   j+1, k)
        Site<ND> sj = decompose<ND>(qj, N);
        uint32_t qjk = shifted<ND>(sj, 2, 1, N);                   // (i,   j+1, k+1)
        uint32_t qk = shifted<ND>(s, 2, 1, N);                     // (i,   j,   k+1)
        double n0 = ang[m];
        double ni = ang[qi];
        double nij = ang[qij];
        double nj = ang[qj];
        double njk = ang[qjk];
        double nk = ang[qk];
        double sa = edge_theta(n0, ni) + edge_theta(ni, nij) + edge_theta(nij, nj);
        double sb = edge_theta(n0, ni) + edge_theta(ni, njk) + edge_theta(njk, nk);
        double sc = edge_theta(n0, nj) + edge_theta(nj, njk) + edge_theta(njk, nk);
        return sa > accept || sb > accept || sc > accept;
    }
}

template <int ND>
__global__ void __launch_bounds__(kStrBlock)
strings_flag_kernel(int N, uint32_t L, const double *__restrict__ ang, double accept,
                    uint64_t *__restrict__ words, uint32_t *__restrict__ bcount)
{
    __shared__ uint32_t wave_tot[kStrBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t total = 0;
    uint32_t base = blockIdx.x * (uint32_t)kStrSpan;
    for (int it = 0; it < kStrIters; ++it) {
        uint32_t m = base + it * kStrBlock + threadIdx.x;
        bool f = (m < L) && plaquette_flag<ND>(ang, m, N, accept);
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
        bcount[blockIdx.x] = t;
    }
}

// exclusive scan of n block counts by one workgroup (1024 work-items),
// offsets[n] receives the total
__global__ void __launch_bounds__(1024)
scan_counts_kernel(const uint32_t *__restrict__ cnt, uint64_t *__restrict__ off, uint32_t n)
{
    __shared__ uint64_t part[1024];
    uint32_t per = (n + 1023) / 1024;
    uint32_t b = threadIdx.x * per, e = min(n, b + per);
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (int t = 0; t < 1024; ++t) {
            uint64_t v = part[t];
            part[t] = acc;
            acc += v;
        }
        off[n] = acc;
    }
    __syncthreads();
    uint64_t acc = part[threadIdx.x];
    for (uint32_t i = b; i < e; ++i) {
        off[i] = acc;
        acc += cnt[i];
    }
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

static double accept_threshold(int thr)
{
    float f = 0.5f * (1.0f - (float)thr / 100.0f);   // string_finding.cpp:8, float
    return (double)f;
}

int run_strings(Context *c, Member &m, int32_t *coords, int64_t cap, int64_t *count,
                int64_t *num_cores)
{
    if (!m.has_state) {
        set_error("member has no state: call axs_set_member_state first");
        return AXS_ERR_STATE;
    }
    const int member = (int)(&m - c->members.data());
    uint32_t L = (uint32_t)m.L;
    uint32_t nblocks = (L + kStrSpan - 1) / kStrSpan;
    size_t words_b = ((size_t)nblocks * kStrSpan / 64) * 8;
    size_t cnt_b = ((size_t)nblocks * 4 + 255) & ~(size_t)255;
    size_t off_b = ((size_t)(nblocks + 1) * 8 + 255) & ~(size_t)255;
    size_t tail_b = 256;
    const bool cached = c->str_member == member && c->str_version == m.version && c->strbuf &&
                        c->strbuf_bytes >= words_b + cnt_b + off_b + tail_b;
    if (!cached) {
        c->str_member = -1;
        AXS_TRY(ensure_device(c, &c->strbuf, &c->strbuf_bytes, words_b + cnt_b + off_b + tail_b));
    }
    char *sp = (char *)c->strbuf;
    uint64_t *words = (uint64_t *)sp;
    uint32_t *bcount = (uint32_t *)(sp + words_b);
    uint64_t *off = (uint64_t *)(sp + words_b + cnt_b);
    unsigned long long *pairs = (unsigned long long *)(sp + words_b + cnt_b + off_b);

    uint64_t total = 0;
    if (cached) {
        total = c->str_total;
    } else {
        // the angles need phi only (no FP32 velocity reconstruction)
        const void *phi = m.state[c->cur][PHI];
        AXS_TRY(ensure_scratch(c, (size_t)L * sizeof(double)));
        double *ang = (double *)c->scratch;
        const unsigned ga = (L + 255) / 256;
        if (c->prec == AXS_FP64)
            hipLaunchKernelGGL(angle_kernel<PhiF64>, dim3(ga), dim3(256), 0, c->stream, L,
                               PhiF64{(const dvec2 *)phi}, ang);
        else
            hipLaunchKernelGGL(angle_kernel<PhiF32>, dim3(ga), dim3(256), 0, c->stream, L,
                               PhiF32{(const fvec4 *)phi}, ang);
        double accept = accept_threshold(c->p.thr);
        auto k = (c->nd == 2) ? strings_flag_kernel<2> : strings_flag_kernel<3>;
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(kStrBlock), 0, c->stream, m.N, L, ang, accept,
                           words, bcount);
        hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, c->stream, bcount, off,
                           nblocks);
        AXS_HIP_TRY(hipGetLastError());
        AXS_HIP_TRY(hipMemcpyAsync(&total, off + nblocks, 8, hipMemcpyDeviceToHost, c->stream));
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
        c->str_member = member;
        c->str_version = m.version;
        c->str_total = total;
    }
    *count = (int64_t)total;

    bool need_list = (coords != nullptr && cap > 0) || (num_cores != nullptr && c->nd == 2);
    if (!need_list) {
        if (num_cores) *num_cores = (int64_t)total;
        return (coords == nullptr || cap >= (int64_t)total) ? AXS_OK : AXS_ERR_CAPACITY;
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
            AXS_HIP_TRY(hipMemcpyAsync(coords, dlist, (size_t)ncopy * c->nd * sizeof(int32_t),
                                       hipMemcpyDeviceToHost, c->stream));
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return (coords == nullptr || cap >= (int64_t)total) ? AXS_OK : AXS_ERR_CAPACITY;
}

}  // namespace axs
