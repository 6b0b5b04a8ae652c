// axs_pk.hip -- power spectra (reference output_powerspec,
// src/common/pk_estimation.cpp:37-214, called on phi2 at s_run.cpp:206; any
// stored real component or the complex field phi1 + i phi2 / phidot1 + i
// phidot2 through axs_member_powerspec_field).
//
// The reference multiplies the field by (-1)^(i+j+k) and takes a complex
// forward FFT of the whole N^d array, so that array index i corresponds to
// the signed frequency i - N/2.  The shift only relabels modes; here the
// unshifted field goes through a rocFFT real-to-complex transform (half the
// spectrum, half the memory) and each stored mode (fi, fj, fz >= 0) stands
// for itself and, when 0 < fz < N/2, for its Hermitian partner -k with the
// same |F|^2 and |k| (weight 2).
//
// Binning is done per bin rather than per mode: a bin is a range [lo, hi)
// of the integer |k|^2, so for every (fi, fj) row the modes that fall in
// the bin are one contiguous run of fz.  The rows of each bin's bounding
// box are cut into chunks of kPkRows; one workgroup sums one (bin, chunk)
// item in compensated FP64 (a work-item per row, then a fixed-order tree),
// and a fold kernel merges each bin's items (a work-group per bin: strided
// runs in chunk order, then a fixed pairwise tree) -- balanced over the
// bins (the outer log bins hold most modes), deterministic, no atomics.
#include <rocfft/rocfft.h>

#include <math.h>

#include <algorithm>

#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

struct FftPlan {
    int N = 0, nd = 0, batch = 1;
    bool own = false;           // the transform of our own (power-of-two sides)
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    void *work = nullptr;
    size_t work_bytes = 0;
    double *in = nullptr;       // rocFFT: real input copy (rocFFT may overwrite input)
    double *out = nullptr;      // rocFFT: complex half spectrum, interleaved
    double2 *ya = nullptr, *yb = nullptr;   // own: the passes' intermediate arrays
    double2 *tw = nullptr;      // own: W_N^m, m < N
    ~FftPlan()
    {
        if (plan) rocfft_plan_destroy(plan);
        if (info) rocfft_execution_info_destroy(info);
        if (work) (void)hipFree(work);
        if (in) (void)hipFree(in);
        if (out) (void)hipFree(out);
        if (ya) (void)hipFree(ya);
        if (yb) (void)hipFree(yb);
        if (tw) (void)hipFree(tw);
    }
};

static int rocfft_ready = 0;

// batched real-to-complex plan for `batch` lattices of side N, with its
// input / output / work buffers; kept for the context's lifetime
static int get_plan(Context *c, int N, int batch, FftPlan **out)
{
    for (FftPlan *p : c->fft_plans)
        if (!p->own && p->N == N && p->nd == c->nd && p->batch == batch) {
            *out = p;
            return AXS_OK;
        }
    if (!rocfft_ready) {
        if (rocfft_setup() != rocfft_status_success) {
            set_error("rocfft_setup failed");
            return AXS_ERR_FFT;
        }
        rocfft_ready = 1;
    }
    FftPlan *p = new FftPlan();
    p->N = N;
    p->nd = c->nd;
    p->batch = batch;
    size_t lengths[3] = {(size_t)N, (size_t)N, (size_t)N};  // fastest first
    size_t L = ((c->nd == 3) ? (size_t)N * N * N : (size_t)N * N) * (size_t)batch;
    size_t half = L / N * (N / 2 + 1);
    if (rocfft_plan_create(&p->plan, rocfft_placement_notinplace,
                           rocfft_transform_type_real_forward, rocfft_precision_double,
                           (size_t)c->nd, lengths, (size_t)batch, nullptr) != rocfft_status_success) {
        delete p;
        set_error("rocfft_plan_create failed for N=%d ndims=%d", N, c->nd);
        return AXS_ERR_FFT;
    }
    if (rocfft_plan_get_work_buffer_size(p->plan, &p->work_bytes) != rocfft_status_success ||
        rocfft_execution_info_create(&p->info) != rocfft_status_success) {
        delete p;
        set_error("rocfft execution info setup failed");
        return AXS_ERR_FFT;
    }
    hipError_t e = hipMalloc(&p->in, L * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&p->out, half * 2 * sizeof(double));
    if (e == hipSuccess && p->work_bytes) e = hipMalloc(&p->work, p->work_bytes);
    if (e != hipSuccess) {
        delete p;
        set_error("power-spectrum buffers: %s", hipGetErrorString(e));
        return AXS_ERR_OUT_OF_MEMORY;
    }
    if (p->work_bytes &&
        rocfft_execution_info_set_work_buffer(p->info, p->work, p->work_bytes) !=
            rocfft_status_success) {
        delete p;
        set_error("rocfft_execution_info_set_work_buffer failed");
        return AXS_ERR_FFT;
    }
    c->fft_plans.push_back(p);
    *out = p;
    return AXS_OK;
}

// One real field per transform of a batch: a component of a stored pair.
// kind 0: FP64 pairs (phi or phidot), 1: the FP32 (phi1, phi2, u1, u2)
// quadruple (phi only), 2: FP32 pairs (the phidot read-back); comp 0 or 1.
struct PkSrc {
    const void *p;
    int kind, comp;
};

// the value of site m of a source
__device__ __forceinline__ double pk_value(const PkSrc &s, size_t m)
{
    if (s.kind == 0) {
        const dvec2 w = ((const dvec2 *)s.p)[m];
        return s.comp ? w.y : w.x;
    }
    if (s.kind == 1) {
        const fvec2 w = ((const fvec2 *)s.p)[2 * m];                    // (phi1, phi2)
        return (double)(s.comp ? w.y : w.x);
    }
    const fvec2 w = ((const fvec2 *)s.p)[m];
    return (double)(s.comp ? w.y : w.x);
}

// the selected component of every lattice of a batch (one per blockIdx.y)
// as contiguous double arrays, lattice b at y + b L
__global__ void __launch_bounds__(256)
pack_kernel(uint32_t L, const PkSrc *__restrict__ src, double *__restrict__ y)
{
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < L) y[(size_t)blockIdx.y * L + m] = pk_value(src[blockIdx.y], m);
}

struct KahanP {
    double s, c;
};

__device__ __forceinline__ void kaddp(KahanP &k, double x)
{
    double t = k.s + x;
    if (fabs(k.s) >= fabs(x))
        k.c += (k.s - t) + x;
    else
        k.c += (x - t) + k.s;
    k.s = t;
}

__device__ __forceinline__ KahanP kmergep(KahanP a, KahanP b)
{
    double t = a.s + b.s;
    double bp = t - a.s;
    double e = (a.s - (t - bp)) + (b.s - bp);
    return KahanP{t, a.c + b.c + e};
}

// smallest integer r >= 0 with r*r >= v (v >= 0)
__device__ __forceinline__ int64_t isqrt_ceil(int64_t v)
{
    if (v <= 0) return 0;
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r < v) ++r;
    while (r > 0 && (r - 1) * (r - 1) >= v) --r;
    return r;
}

constexpr int kPkRows = 256;    // bounding-box rows per work item

// Item t of bin b (item_begin[b] <= t < item_begin[b+1]) sums the modes with
// lo2[b] <= |k|^2 < hi2[b] on rows [chunk*kPkRows, chunk*kPkRows + kPkRows)
// of the bin's bounding box |fi|,|fj| <= R (3D) or |fj| <= R (2D), rows in
// row-major order.  F: half spectrum, fast axis length H = N/2+1.
template <int ND>
__global__ void __launch_bounds__(kPkRows)
bin_kernel(int N, const double2 *__restrict__ F, int nbins, const uint32_t *__restrict__ item_begin,
           const int64_t *__restrict__ lo2, const int64_t *__restrict__ hi2,
           KahanP *__restrict__ part_pk, KahanP *__restrict__ part_ks,
           long long *__restrict__ part_cnt)
{
    __shared__ KahanP sh_pk[kPkRows / 64], sh_ks[kPkRows / 64];
    __shared__ long long sh_n[kPkRows / 64];
    const uint32_t item = blockIdx.x;
    const int H0 = N / 2 + 1;
    F += (size_t)blockIdx.y * (size_t)((ND == 3) ? (int64_t)N * N * H0 : (int64_t)N * H0);
    const size_t pslot = (size_t)blockIdx.y * gridDim.x + item;
    int lo_b = 0, hi_b = nbins - 1;              // last bin starting at or before item
    while (lo_b < hi_b) {
        int mid = (lo_b + hi_b + 1) >> 1;
        if (item_begin[mid] <= item)
            lo_b = mid;
        else
            hi_b = mid - 1;
    }
    const int b = lo_b;
    const int64_t chunk = item - item_begin[b];
    const int64_t lo = lo2[b], hi = hi2[b];
    const int H = N / 2 + 1;
    KahanP pk{0.0, 0.0}, ks{0.0, 0.0};
    long long cnt = 0;
    {
        int64_t R = isqrt_ceil(hi);           // |f| < R on every axis
        if (R > N / 2) R = N / 2;
        int64_t side = 2 * R + 1;
        int64_t nrows = (ND == 3) ? side * side : side;
        int64_t r = chunk * kPkRows + threadIdx.x;
        if (hi > lo && r < nrows) {
            int64_t fi = 0, fj;
            if (ND == 3) {
                fi = r / side - R;
                fj = r % side - R;
            } else {
                fj = r - R;
            }
            int64_t r2 = fi * fi + fj * fj;
            if (fi > -N / 2 && fj > -N / 2 && r2 < hi) {   // |k| >= N/2 otherwise
                int64_t z0 = isqrt_ceil(lo - r2);
                int64_t z1 = isqrt_ceil(hi - r2);        // exclusive
                if (z1 > H) z1 = H;
                int64_t ii = fi < 0 ? fi + N : fi, jj = fj < 0 ? fj + N : fj;
                int64_t rowbase = (ND == 3) ? (ii * N + jj) * H : jj * H;
                for (int64_t z = z0; z < z1; ++z) {
                    double2 f = F[rowbase + z];
                    double w = (z == 0 || z == N / 2) ? 1.0 : 2.0;
                    double p2 = f.x * f.x + f.y * f.y;
                    double km = sqrt((double)(r2 + z * z));
                    kaddp(pk, w * p2);
                    kaddp(ks, w * km);
                    cnt += (long long)w;
                }
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        KahanP a{__shfl_down(pk.s, o, 64), __shfl_down(pk.c, o, 64)};
        KahanP k2{__shfl_down(ks.s, o, 64), __shfl_down(ks.c, o, 64)};
        long long n2 = __shfl_down(cnt, o, 64);
        pk = kmergep(pk, a);
        ks = kmergep(ks, k2);
        cnt += n2;
    }
    if (lane == 0) {
        sh_pk[wave] = pk;
        sh_ks[wave] = ks;
        sh_n[wave] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        KahanP a = sh_pk[0], k = sh_ks[0];
        long long n = sh_n[0];
        for (int w = 1; w < kPkRows / 64; ++w) {
            a = kmergep(a, sh_pk[w]);
            k = kmergep(k, sh_ks[w]);
            n += sh_n[w];
        }
        part_pk[pslot] = a;
        part_ks[pslot] = k;
        part_cnt[pslot] = n;
    }
}

// bin blockIdx.x of lattice blockIdx.y: work-item t merges the bin's items
// t, t + 256, ... in order, then a fixed pairwise tree over the work-items
// (deterministic for a given item count)
__global__ void __launch_bounds__(256)
bin_fold_kernel(int nbins, uint32_t nitems, const uint32_t *__restrict__ item_begin,
                const KahanP *__restrict__ part_pk, const KahanP *__restrict__ part_ks,
                const long long *__restrict__ part_cnt, double *__restrict__ out_pk,
                double *__restrict__ out_ks, long long *__restrict__ out_cnt)
{
    __shared__ KahanP sa[256], sk[256];
    __shared__ long long sn[256];
    const int b = blockIdx.x;
    const size_t pofs = (size_t)blockIdx.y * nitems, oofs = (size_t)blockIdx.y * nbins;
    KahanP a{0.0, 0.0}, k{0.0, 0.0};
    long long n = 0;
    for (uint32_t t = item_begin[b] + threadIdx.x; t < item_begin[b + 1]; t += 256) {
        a = kmergep(a, part_pk[pofs + t]);
        k = kmergep(k, part_ks[pofs + t]);
        n += part_cnt[pofs + t];
    }
    sa[threadIdx.x] = a;
    sk[threadIdx.x] = k;
    sn[threadIdx.x] = n;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sa[threadIdx.x] = kmergep(sa[threadIdx.x], sa[threadIdx.x + h]);
            sk[threadIdx.x] = kmergep(sk[threadIdx.x], sk[threadIdx.x + h]);
            sn[threadIdx.x] += sn[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out_pk[oofs + b] = sa[0].s + sa[0].c;
        out_ks[oofs + b] = sk[0].s + sk[0].c;
        out_cnt[oofs + b] = sn[0];
    }
}

// ---------------------------------------------------------------------------
// Row-order binning, for bin tables that fit the work-group's LDS.  The half
// spectrum is read once, in storage order: wave w of the grid takes rows w,
// w + W, ... (W waves in all), its 64 lanes 64 consecutive modes z of a row
// at a time.  Along a row |k|^2 = r2 + z^2 grows with z, so the lanes' bins
// never decrease across the wave and each bin present is one run of lanes: a
// segmented inclusive scan (DPP row shifts and row broadcasts, registers
// only) leaves each run's sum in its last lane, which adds it (compensated)
// to the wave's accumulator for that bin in LDS -- the runs' bins differ, so
// no two lanes touch one accumulator.  At the end the work-group merges its
// waves' accumulators in wave order into one partial per bin, and
// fold_rows_kernel merges the work-groups' partials per bin in work-group
// order: deterministic for a given lattice.
// A mode's bin: s = floor(|k|) from a float square root corrected by
// integer compares, the first bin that can hold |k|^2 >= s^2 from a table of
// the shells s, then up the bins while |k|^2 >= hi (at most a few steps,
// only where bins are narrower than a shell).
// GEOM sums the data-independent |k| (weighted as the modes) and counts the
// modes from the runs' lengths, without reading the spectrum.
constexpr int kScanWaves = 4;
constexpr size_t kScanLdsMax = 64 * 1024;

// v of lane (lane - d) in the same 16-lane row (row_shr:d), of lane 15 of
// the previous row (row_bcast:15, rows 1 and 3) or of lane 31 (row_bcast:31,
// rows 2 and 3); lanes without a source read 0
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_take(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xf, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

static size_t scan_lds_bytes(int N, int nbins, bool geom)
{
    size_t b = (size_t)kScanWaves * nbins * sizeof(KahanP);
    if (geom) b += (size_t)kScanWaves * nbins * sizeof(long long);
    b += (size_t)nbins * sizeof(uint32_t);
    b += ((size_t)(N / 2) * sizeof(uint16_t) + 3) & ~(size_t)3;
    return b;
}

template <int ND, bool GEOM>
__global__ void __launch_bounds__(64 * kScanWaves)
bin_scan_kernel(int N, const double2 *__restrict__ F, int nbins, const uint32_t *__restrict__ hi2g,
                const uint16_t *__restrict__ shellg, uint32_t nrows, KahanP *__restrict__ part,
                long long *__restrict__ part_cnt)
{
    extern __shared__ __align__(16) unsigned char smem[];
    KahanP *acc = reinterpret_cast<KahanP *>(smem);                    // [wave][bin]
    long long *acnt = reinterpret_cast<long long *>(acc + kScanWaves * nbins);
    uint32_t *hi2 = reinterpret_cast<uint32_t *>(acnt + (GEOM ? kScanWaves * nbins : 0));
    uint16_t *shell = reinterpret_cast<uint16_t *>(hi2 + nbins);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int half = N / 2;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        hi2[i] = hi2g[i];
        for (int w = 0; w < kScanWaves; ++w) {
            acc[w * nbins + i] = KahanP{0.0, 0.0};
            if (GEOM) acnt[w * nbins + i] = 0;
        }
    }
    for (int i = threadIdx.x; i < half; i += blockDim.x) shell[i] = shellg[i];
    __syncthreads();
    const uint32_t H = (uint32_t)half + 1;
    if (!GEOM) F += (size_t)blockIdx.y * (size_t)nrows * H;
    KahanP *wacc = acc + wave * nbins;
    long long *wcnt = acnt + wave * nbins;
    const uint32_t top = (uint32_t)half * (uint32_t)half;         // |k| < N/2
    const uint32_t W = gridDim.x * kScanWaves;
    const uint64_t upto = (2ull << lane) - 1;                     // lanes 0 .. lane
    for (uint32_t r = blockIdx.x * kScanWaves + (uint32_t)wave; r < nrows; r += W) {
        int fi = 0, fj;
        if (ND == 3) {
            const uint32_t ii = r / (uint32_t)N, jj = r - ii * (uint32_t)N;
            fi = (int)ii < half ? (int)ii : (int)ii - N;
            fj = (int)jj < half ? (int)jj : (int)jj - N;
        } else {
            fj = (int)r < half ? (int)r : (int)r - N;
        }
        const uint32_t r2 = (uint32_t)(fi * fi) + (uint32_t)(fj * fj);
        if (r2 >= top) continue;                                 // the whole row has |k| >= N/2
        const uint32_t zmax = (uint32_t)isqrt_ceil((int64_t)(top - r2));   // z < zmax <=> |k|^2 < top
        const uint32_t zs = (r2 == 0) ? 1u : 0u;                 // not the k = 0 mode
        const double2 *row = F + (size_t)r * H;
        for (uint32_t z0 = 0; z0 < zmax; z0 += 64) {
            const uint32_t z = z0 + (uint32_t)lane;
            const bool on = z >= zs && z < zmax;
            const uint32_t k2 = r2 + z * z;
            int b;
            double v = 0.0;
            if (on) {
                uint32_t sh = (uint32_t)__builtin_amdgcn_sqrtf((float)k2);
                if (sh * sh > k2) --sh;
                if ((sh + 1) * (sh + 1) <= k2) ++sh;
                b = shell[sh];
                while (hi2[b] <= k2) ++b;
                if (GEOM) {
                    const double km = sqrt((double)k2);
                    v = (z == 0) ? km : 2.0 * km;
                } else {
                    const double2 f = row[z];
                    const double p2 = f.x * f.x + f.y * f.y;
                    v = (z == 0) ? p2 : 2.0 * p2;
                }
            } else {
                b = (z < zs) ? -1 : nbins;
            }
            const int bprev = __builtin_amdgcn_update_dpp(-2, b, 0x138, 0xf, 0xf, false);
            const uint64_t heads = __ballot(bprev != b);        // first lane of each run
            const int start = 63 - __clzll(heads & upto);
            const int len = lane - start;                        // earlier lanes of the run
            double u;
            u = dpp_take<0x111, 0xf>(v);
            if (len >= 1) v += u;
            u = dpp_take<0x112, 0xf>(v);
            if (len >= 2) v += u;
            u = dpp_take<0x114, 0xf>(v);
            if (len >= 4) v += u;
            u = dpp_take<0x118, 0xf>(v);
            if (len >= 8) v += u;
            u = dpp_take<0x142, 0xa>(v);
            if (len > (lane & 15)) v += u;                       // run reaches the row before
            u = dpp_take<0x143, 0xc>(v);
            if (len > (lane & 31)) v += u;                       // run reaches lanes 0..31
            const bool last = lane == 63 || (((heads >> lane) >> 1) & 1ull);
            if (last && b >= 0 && b < nbins) {
                KahanP a = wacc[b];
                kaddp(a, v);
                wacc[b] = a;
                if (GEOM) wcnt[b] += 2 * (long long)(len + 1) - ((z - (uint32_t)len == 0) ? 1 : 0);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        KahanP a = acc[i];
        for (int w = 1; w < kScanWaves; ++w) a = kmergep(a, acc[w * nbins + i]);
        const size_t slot = ((size_t)blockIdx.y * nbins + i) * gridDim.x + blockIdx.x;
        part[slot] = a;
        if (GEOM) {
            long long n = 0;
            for (int w = 0; w < kScanWaves; ++w) n += acnt[w * nbins + i];
            part_cnt[slot] = n;
        }
    }
}

// bin blockIdx.x of lattice blockIdx.y: the nblk work-groups' partials of
// bin_scan_kernel merged in work-group order (strided runs, then a fixed
// pairwise tree); counts summed when part_cnt is given
__global__ void __launch_bounds__(256)
fold_rows_kernel(int nbins, uint32_t nblk, const KahanP *__restrict__ part,
                 const long long *__restrict__ part_cnt, double *__restrict__ out,
                 long long *__restrict__ out_cnt)
{
    __shared__ KahanP sa[256];
    __shared__ long long sn[256];
    const size_t base = ((size_t)blockIdx.y * nbins + blockIdx.x) * nblk;
    KahanP a{0.0, 0.0};
    long long n = 0;
    for (uint32_t t = threadIdx.x; t < nblk; t += 256) {
        a = kmergep(a, part[base + t]);
        if (part_cnt) n += part_cnt[base + t];
    }
    sa[threadIdx.x] = a;
    sn[threadIdx.x] = n;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sa[threadIdx.x] = kmergep(sa[threadIdx.x], sa[threadIdx.x + h]);
            sn[threadIdx.x] += sn[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t o = (size_t)blockIdx.y * nbins + blockIdx.x;
        out[o] = sa[0].s + sa[0].c;
        if (out_cnt) out_cnt[o] = sn[0];
    }
}

// ---------------------------------------------------------------------------
// Power-of-two sides: the transform in passes of our own that each read and
// write HBM once, the first reading the field straight from the state and
// the last binning the modes as it produces them (no pack, no spectrum
// written, no binning pass).  3D, lattice f[i][j][z], M = N/2:
//   Z  rows along z (real): g_m = f_2m + i f_2m+1, an M-point complex DFT
//      G, then X_k = E_k + W_N^k O_k with E_k = (G_k + conj G_{M-k}) / 2,
//      O_k = -i (G_k - conj G_{M-k}) / 2 (G_M = G_0), for k < M (the plane
//      k_z = N/2 has |k| >= N/2 and is never binned); written as Y[i][k][j].
//   Y  rows along j: N-point DFTs, written as Y'[k_z][k_j][i].
//   X  rows along i: N-point DFTs, then every mode's w |F|^2 (w = 1 at
//      k_z = 0, else 2 for the Hermitian partner) binned as bin_scan_kernel
//      does; along a row |k| runs up over the first half and down over the
//      second, so the 64 lanes of a chunk (N >= 128) still hold each bin
//      present as one run.
// 2D (f[j][i]) is Z along i, written as Y[k][j], then the binning pass
// along j.  A work-group transforms kFftPts complex points in LDS, rows of
// one length L: radix-4 Stockham stages (one radix-2 stage first when
// log2 L is odd), each thread holding 16 points in registers between a
// stage's reads and writes (two barriers per stage, one buffer).  Rows are
// padded by one point in LDS so that the column reads of the transposed
// writes fall in distinct banks.  Twiddles: one table of W_N^m, m < N.
constexpr int kFftPts = 4096;
constexpr int kFftThr = 256;

__device__ __forceinline__ double2 cmulw(double2 a, double2 w)
{
    return double2{fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x)};
}

__device__ __forceinline__ void dft2(double2 (&v)[2])
{
    const double2 a = v[0], b = v[1];
    v[0] = double2{a.x + b.x, a.y + b.y};
    v[1] = double2{a.x - b.x, a.y - b.y};
}

// forward 4-point DFT (W_4 = -i)
__device__ __forceinline__ void dft4(double2 (&v)[4])
{
    const double2 t0{v[0].x + v[2].x, v[0].y + v[2].y}, t1{v[0].x - v[2].x, v[0].y - v[2].y};
    const double2 t2{v[1].x + v[3].x, v[1].y + v[3].y};
    const double2 t3{v[1].y - v[3].y, v[3].x - v[1].x};            // -i (v1 - v3)
    v[0] = double2{t0.x + t2.x, t0.y + t2.y};
    v[2] = double2{t0.x - t2.x, t0.y - t2.y};
    v[1] = double2{t1.x + t3.x, t1.y + t3.y};
    v[3] = double2{t1.x - t3.x, t1.y - t3.y};
}

// 16-point DFT as 4 x 4 (Cooley-Tukey): DFT4 over v[b + 4a] for each b,
// twiddles W_16^{bp}, DFT4 over b; output p + 4q
__device__ __forceinline__ void dft16(double2 (&v)[16])
{
    constexpr double c1 = 0x1.d906bcf328d46p-1, s1 = 0x1.87de2a6aea963p-2;   // cos, sin pi/8
    constexpr double h = 0x1.6a09e667f3bcdp-1;                               // sqrt(1/2)
    double2 y[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double2 t[4] = {v[b], v[b + 4], v[b + 8], v[b + 12]};
        dft4(t);
#pragma unroll
        for (int p = 0; p < 4; ++p) y[b][p] = t[p];
    }
    // W_16^1 = (c1, -s1), ^2 = (h, -h), ^3 = (s1, -c1), ^4 = -i, ^6 = (-h, -h), ^9 = (-c1, s1)
    y[1][1] = cmulw(y[1][1], double2{c1, -s1});
    y[1][2] = cmulw(y[1][2], double2{h, -h});
    y[1][3] = cmulw(y[1][3], double2{s1, -c1});
    y[2][1] = cmulw(y[2][1], double2{h, -h});
    y[2][2] = double2{y[2][2].y, -y[2][2].x};
    y[2][3] = cmulw(y[2][3], double2{-h, -h});
    y[3][1] = cmulw(y[3][1], double2{s1, -c1});
    y[3][2] = cmulw(y[3][2], double2{-h, -h});
    y[3][3] = cmulw(y[3][3], double2{-c1, s1});
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        double2 t[4] = {y[0][p], y[1][p], y[2][p], y[3][p]};
        dft4(t);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[p + 4 * q] = t[q];
    }
}

// LDS position of point i of a row: one pad every 16 points, so that the
// first stage's strided writes fall in distinct banks
__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }

// row stride in LDS of rows of L points (padded; the odd extra point keeps
// the rows' starts in different banks)
template <int L>
constexpr int fft_ls() { return L + L / 16 + 1; }

// One radix-RAD Stockham stage over all rows in buf (rows of L points at
// stride LS, kFftPts points): Ns = product of the earlier radices; butterfly
// j of a row reads points j + k L/RAD, twiddles them by W_L^{(j mod Ns) k
// L/(Ns RAD)} = tw[(...) ts], and writes point k to (j - j mod Ns) RAD +
// j mod Ns + k Ns.
template <int L, int RAD, int PTS>
__device__ __forceinline__ void fft_stage(double2 *buf, const double2 *__restrict__ tw, int ts,
                                          int Ns)
{
    constexpr int Q = L / RAD, LS = fft_ls<L>(), THR = PTS / 16;
    constexpr int NB = PTS / RAD / THR;
    double2 v[NB][RAD];
    int ob[NB], oj[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = threadIdx.x + q * THR;
        const int row = b / Q, j = b - row * Q;
        const int jm = j & (Ns - 1);
        const double2 *rb = buf + row * LS;
#pragma unroll
        for (int k = 0; k < RAD; ++k) v[q][k] = rb[fpad(j + k * Q)];
        if (Ns > 1) {
            const int step = jm * (L / (Ns * RAD)) * ts;
#pragma unroll
            for (int k = 1; k < RAD; ++k) v[q][k] = cmulw(v[q][k], tw[k * step]);
        }
        if constexpr (RAD == 16)
            dft16(v[q]);
        else if constexpr (RAD == 4)
            dft4(v[q]);
        else
            dft2(v[q]);
        ob[q] = row * LS;
        oj[q] = (j - jm) * RAD + jm;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int k = 0; k < RAD; ++k) buf[ob[q] + fpad(oj[q] + k * Ns)] = v[q][k];
    __syncthreads();
}

// forward DFT of every row in buf (PTS points, PTS / 16 threads);
// tw[m ts] = W_L^m.  Radix-16 stages while 16 divides what is left, then
// radix 4 and 2.
template <int L, int PTS = kFftPts>
__device__ __forceinline__ void fft_rows(double2 *buf, const double2 *__restrict__ tw, int ts)
{
    constexpr int n = __builtin_ctz(L);
    constexpr int n16 = n / 4, rest = n % 4;
    int Ns = 1;
#pragma unroll
    for (int t = 0; t < n16; ++t, Ns *= 16) fft_stage<L, 16, PTS>(buf, tw, ts, Ns);
    if constexpr (rest >= 2) {
        fft_stage<L, 4, PTS>(buf, tw, ts, Ns);
        Ns *= 4;
    }
    if constexpr ((rest & 1) != 0) fft_stage<L, 2, PTS>(buf, tw, ts, Ns);
}

// Z pass: rows of N reals of lattice blockIdx.y (rpl rows: N^2 in 3D, N in
// 2D), R = kFftPts / M rows per work-group; writes the k < M half of each
// row as Y[lattice][i][k][j] (3D) or Y[lattice][k][j] (2D)
template <int ND, int N, class V>
__device__ __forceinline__ void fft_z_load(double2 *buf, const V *__restrict__ pv, int comp,
                                           size_t site0)
{
    constexpr int M = N / 2, LS = fft_ls<M>();
    // the source pointer came from memory (a generic pointer): name it a
    // global one, so that the loads are global loads that need not wait for
    // each other or for the LDS stores
    using GV = const __attribute__((address_space(1))) V;
    GV *p = (GV *)pv;
    // one site per lane (consecutive lanes, consecutive sites: whole lines),
    // even sites into the real parts and odd ones into the imaginary parts
#pragma unroll 8
    for (int q = 0; q < 2 * kFftPts / kFftThr; ++q) {
        const int e = threadIdx.x + q * kFftThr;
        const int r = e / N, z = e - r * N;
        const V w = p[site0 + (size_t)e];
        reinterpret_cast<double *>(buf + r * LS + fpad(z >> 1))[z & 1] =
            (double)(comp ? w.y : w.x);
    }
}

template <int ND, int N>
__global__ void __launch_bounds__(kFftThr)
fft_z_kernel(const PkSrc *__restrict__ src, double2 *__restrict__ y,
             const double2 *__restrict__ tw)
{
    constexpr int M = N / 2, R = kFftPts / M, LS = fft_ls<M>();
    constexpr uint32_t rpl = (ND == 3) ? (uint32_t)N * N : (uint32_t)N;
    __shared__ double2 buf[R * LS];
    const uint32_t b = blockIdx.y, row0 = blockIdx.x * R;
    const PkSrc sc = src[b];
    const size_t site0 = (size_t)row0 * N;
    if (sc.kind == 0)
        fft_z_load<ND, N>(buf, (const dvec2 *)sc.p, sc.comp, site0);
    else if (sc.kind == 1)      // (phi1, phi2, u1, u2): the first pair of each quadruple
        fft_z_load<ND, N>(buf, (const fvec4 *)sc.p, sc.comp, site0);
    else
        fft_z_load<ND, N>(buf, (const fvec2 *)sc.p, sc.comp, site0);
    __syncthreads();
    fft_rows<M>(buf, tw, 2);                         // W_M^m = W_N^2m
#pragma unroll 4
    for (int q = 0; q < kFftPts / kFftThr; ++q) {
        const int e = threadIdx.x + q * kFftThr;
        const int r = e % R, k = e / R;
        const double2 a = buf[r * LS + fpad(k)], c = buf[r * LS + fpad((M - k) & (M - 1))];
        const double2 E{0.5 * (a.x + c.x), 0.5 * (a.y - c.y)};
        const double2 O{0.5 * (a.y + c.y), -0.5 * (a.x - c.x)};
        const double2 w = cmulw(O, tw[k]);
        const double2 X{E.x + w.x, E.y + w.y};
        const uint32_t rl = row0 + (uint32_t)r;
        size_t idx;
        if (ND == 3) {
            const uint32_t i = rl / (uint32_t)N, j = rl - i * (uint32_t)N;
            idx = (((size_t)b * N + i) * M + (uint32_t)k) * N + j;
        } else {
            idx = ((size_t)b * M + (uint32_t)k) * N + rl;
        }
        y[idx] = X;
    }
    (void)rpl;
}

// Y pass (3D): rows Y[b][i][k_z][*] along j, R = kFftPtsY / N rows of
// consecutive i per group (eight at N = 1024: whole 128-byte lines in the
// transposed write); writes Y'[b][k_z][k_j][i].  A group's rows fill most
// of a CU's LDS, so one work-group runs per CU: each takes a contiguous run
// of groups and loads the next group's points into registers while the
// current one is transformed and stored, and reads the twiddles from a copy
// in LDS (a global twiddle load would wait for those earlier loads).
constexpr int kFftPtsY = 8192;

template <int N>
__global__ void __launch_bounds__(kFftPtsY / 16)
fft_y_kernel(const double2 *__restrict__ ya, double2 *__restrict__ yb,
             const double2 *__restrict__ tw, uint32_t ngroups)
{
    constexpr int M = N / 2, R = kFftPtsY / N, LS = fft_ls<N>(), GI = N / R;
    constexpr int THR = kFftPtsY / 16, PER = kFftPtsY / THR;
    static_assert(N <= 1024, "rows and twiddles must fit the LDS");
    __shared__ double2 buf[R * LS];
    __shared__ double2 twl[N];
    for (int m = threadIdx.x; m < N; m += THR) twl[m] = tw[m];
    const uint32_t g0 = (uint32_t)(((uint64_t)ngroups * blockIdx.x) / gridDim.x);
    const uint32_t g1 = (uint32_t)(((uint64_t)ngroups * (blockIdx.x + 1)) / gridDim.x);
    // point q of a thread is e = tid + q THR of the group: row e / N, column
    // e mod N.  N and THR are powers of two, so its offsets split into a part
    // of the thread and a constant of q: few registers stay live across the
    // transform besides the prefetched points.
    const uint32_t tid = threadIdx.x;
    const uint32_t rl = (N <= THR) ? tid / N : 0, jl = (N <= THR) ? tid % N : tid;
    const uint32_t ld_l = rl * (uint32_t)(M * N) + jl, st_l = rl * LS + fpad(jl);
    auto ld_q = [](int q) { return (uint32_t)((q * THR / N) * M * N + (q * THR) % N); };
    auto st_q = [](int q) { return (q * THR / N) * LS + fpad((q * THR) % N); };
    auto src = [&](uint32_t gq) {
        const uint32_t ig = gq % GI, t = gq / GI, kz = t % M, b = t / M;
        return ya + (((size_t)b * N + ig * R) * M + kz) * N + ld_l;
    };
    // the prefetched points as plain doubles (a double2 copy goes through
    // memory and would keep the array in scratch)
    double nxr[PER], nxi[PER];
    auto fetch = [&](uint32_t gq) {
        const double2 *s = src(gq);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const double2 v = s[ld_q(q)];
            nxr[q] = v.x;
            nxi[q] = v.y;
        }
    };
    auto put = [&]() {
#pragma unroll
        for (int q = 0; q < PER; ++q) buf[st_l + st_q(q)] = double2{nxr[q], nxi[q]};
    };
    if (g0 < g1) {
        fetch(g0);
        put();
    }
    // transposed write: point q is row tid mod R, column k_j = tid / R + q THR / R
    const uint32_t wr = tid % R, wk = tid / R;
    // Per group: the next group's loads are issued before the transform; the
    // result is read out of LDS, the next group's points written in, and
    // only then the result stored, so that no store is pending when the
    // loads are waited for.
    for (uint32_t gq = g0; gq < g1; ++gq) {
        __syncthreads();
        fetch(gq + 1 < g1 ? gq + 1 : gq);      // no branch around the loads
        fft_rows<N, kFftPtsY>(buf, twl, 1);
        double outr[PER], outi[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const double2 v = buf[wr * LS + fpad(wk + (uint32_t)q * (THR / R))];
            outr[q] = v.x;
            outi[q] = v.y;
        }
        __syncthreads();
        put();
        const uint32_t ig = gq % GI, t = gq / GI, kz = t % M, b = t / M;
        double2 *dst = yb + ((size_t)b * M + kz) * N * N + ig * R + (size_t)wk * N + wr;
#pragma unroll
        for (int q = 0; q < PER; ++q)
            dst[(size_t)q * (THR / R) * N] = double2{outr[q], outi[q]};
    }
}

// Last pass with the binning: rows along the last axis (3D: Y'[b][k_z][k_j]
// [*] along i; 2D: Y[b][k][*] along j), R = kFftPts / N rows per group; each
// work-group takes a contiguous run of groups (blockIdx.y = lattice); each
// wave bins rows wave, wave + W, ... of a group into its accumulators as
// bin_scan_kernel does, and each work-group writes one partial per bin at
// the end.  Groups of 2048 points (one row where N > 2048): smaller
// work-groups, more of them per CU, to overlap their loads, transforms and
// binning.
template <int N>
constexpr int fft_bin_pts() { return N > 2048 ? N : 2048; }

template <int ND, int N>
__global__ void __launch_bounds__(fft_bin_pts<N>() / 16)
fft_bin_kernel(const double2 *__restrict__ y, const double2 *__restrict__ tw, int nbins,
               const uint32_t *__restrict__ hi2g, const uint16_t *__restrict__ shellg,
               uint32_t rpl, KahanP *__restrict__ part)
{
    constexpr int PB = fft_bin_pts<N>(), THR = PB / 16;
    constexpr int R = PB / N, LS = fft_ls<N>(), half = N / 2;
    constexpr int W = THR / 64;
    __shared__ double2 buf[R * LS];
    // each wave's bin sums are plain doubles (a wave adds at most two runs
    // per bin and row, over 64 rows at 1024^3: a relative error of at most
    // ~128 ulp, far below the 1e-9 the spectra are checked to), merged
    // compensated across waves and work-groups; half the LDS of compensated
    // pairs, so that 512 bins leave room for three work-groups per CU
    extern __shared__ __align__(16) unsigned char smem[];
    double *acc = reinterpret_cast<double *>(smem);                    // [wave][bin]
    uint32_t *hi2 = reinterpret_cast<uint32_t *>(acc + W * nbins);
    uint16_t *shell = reinterpret_cast<uint16_t *>(hi2 + nbins);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < nbins; i += THR) {
        hi2[i] = hi2g[i];
        for (int w = 0; w < W; ++w) acc[w * nbins + i] = 0.0;
    }
    for (int i = threadIdx.x; i < half; i += THR) shell[i] = shellg[i];
    __syncthreads();
    double *wacc = acc + wave * nbins;
    const double2 *yl = y + (size_t)blockIdx.y * rpl * N;
    const uint32_t top = (uint32_t)half * (uint32_t)half;
    const uint64_t upto = (2ull << lane) - 1;
    const uint32_t ngroups = rpl / R;
    const uint32_t g0 = (uint32_t)(((uint64_t)ngroups * blockIdx.x) / gridDim.x);
    const uint32_t g1 = (uint32_t)(((uint64_t)ngroups * (blockIdx.x + 1)) / gridDim.x);
    for (uint32_t gq = g0; gq < g1; ++gq) {
        const uint32_t row0 = gq * R;
        // a group whose rows all lie outside |k| < N/2 is not transformed
        uint32_t rmin = 0xffffffffu;
        for (int r = 0; r < R; ++r) {
            const uint32_t g = row0 + (uint32_t)r;
            uint32_t r2;
            if (ND == 3) {
                const uint32_t kz = g / (uint32_t)N, kj = g - kz * (uint32_t)N;
                const int fj = (int)kj < half ? (int)kj : (int)kj - N;
                r2 = kz * kz + (uint32_t)(fj * fj);
            } else {
                r2 = g * g;
            }
            rmin = r2 < rmin ? r2 : rmin;
        }
        if (rmin >= top) continue;
#pragma unroll 4
        for (int q = 0; q < PB / THR; ++q) {
            const int e = threadIdx.x + q * THR;
            const int r = e / N, x = e - r * N;
            buf[r * LS + fpad(x)] = yl[(size_t)(row0 + r) * N + x];
        }
        __syncthreads();
        fft_rows<N, PB>(buf, tw, 1);
        for (int r = wave; r < R; r += W) {
            const uint32_t g = row0 + (uint32_t)r;
            uint32_t r2;
            bool zero_plane;
            if (ND == 3) {
                const uint32_t kz = g / (uint32_t)N, kj = g - kz * (uint32_t)N;
                const int fj = (int)kj < half ? (int)kj : (int)kj - N;
                r2 = kz * kz + (uint32_t)(fj * fj);
                zero_plane = kz == 0;
            } else {
                r2 = g * g;
                zero_plane = g == 0;
            }
            if (r2 >= top) continue;                 // the whole row has |k| >= N/2
            for (int x0 = 0; x0 < N; x0 += 64) {
                // chunks lie in one half of the row (N >= 128): the smallest
                // |f| of the chunk decides whether any lane is inside
                const int fmin = x0 < half ? x0 : N - (x0 + 63);
                if (r2 + (uint32_t)(fmin * fmin) >= top) continue;
                const int x = x0 + lane;
                const int fx = x < half ? x : x - N;
                const uint32_t k2 = r2 + (uint32_t)(fx * fx);
                const bool on = k2 < top && k2 > 0;
                int b;
                double v = 0.0;
                if (on) {
                    uint32_t sh = (uint32_t)__builtin_amdgcn_sqrtf((float)k2);
                    if (sh * sh > k2) --sh;
                    if ((sh + 1) * (sh + 1) <= k2) ++sh;
                    b = shell[sh];
                    while (hi2[b] <= k2) ++b;
                    const double2 f = buf[r * LS + fpad(x)];
                    const double p2 = f.x * f.x + f.y * f.y;
                    v = zero_plane ? p2 : 2.0 * p2;
                } else {
                    b = (k2 == 0) ? -1 : nbins;
                }
                const int bprev = __builtin_amdgcn_update_dpp(-2, b, 0x138, 0xf, 0xf, false);
                const uint64_t heads = __ballot(bprev != b);
                const int start = 63 - __clzll(heads & upto);
                const int len = lane - start;
                double u;
                u = dpp_take<0x111, 0xf>(v);
                if (len >= 1) v += u;
                u = dpp_take<0x112, 0xf>(v);
                if (len >= 2) v += u;
                u = dpp_take<0x114, 0xf>(v);
                if (len >= 4) v += u;
                u = dpp_take<0x118, 0xf>(v);
                if (len >= 8) v += u;
                u = dpp_take<0x142, 0xa>(v);
                if (len > (lane & 15)) v += u;
                u = dpp_take<0x143, 0xc>(v);
                if (len > (lane & 31)) v += u;
                const bool last = lane == 63 || (((heads >> lane) >> 1) & 1ull);
                if (last && b >= 0 && b < nbins) wacc[b] += v;
            }
        }
        __syncthreads();                             // before the next group's loads
    }
    for (int i = threadIdx.x; i < nbins; i += THR) {
        KahanP a{acc[i], 0.0};
        for (int w = 1; w < W; ++w) a = kmergep(a, KahanP{acc[w * nbins + i], 0.0});
        part[((size_t)blockIdx.y * nbins + i) * gridDim.x + blockIdx.x] = a;
    }
}

// bins of a lattice of side N as ranges [lo, hi) of the integer |k|^2, and
// the work items (bounding-box rows in chunks of kPkRows) of each bin
struct BinLayout {
    int N = 0, nd = 0, binning = 0, nbins = 0;
    std::vector<int64_t> lo, hi;
    std::vector<uint32_t> items;      // first item of each bin, items[nbins] = total
    // row-order binning: hi as 32-bit, and for each shell s < N/2 the first
    // bin whose range can hold |k|^2 >= s^2
    std::vector<uint32_t> hi32;
    std::vector<uint16_t> shell;
    bool scan = false;                // row-order binning fits the LDS
};

static int bin_layout(int N, int nd, int binning, int nbins, BinLayout &B)
{
    if (binning == AXS_BIN_LINEAR)
        nbins = N / 2;
    else if (binning != AXS_BIN_LOG || nbins < 1 || nbins > 4096) {
        set_error("powerspec: bad binning %d / nbins %d", binning, nbins);
        return AXS_ERR_INVALID_ARG;
    }
    B.N = N;
    B.nd = nd;
    B.binning = binning;
    B.nbins = nbins;
    B.lo.assign(nbins, 0);
    B.hi.assign(nbins, 0);
    std::vector<int64_t> &lo = B.lo, &hi = B.hi;
    int64_t top = (int64_t)(N / 2) * (N / 2);      // |k| < N/2  (pk_estimation.cpp:93)
    if (binning == AXS_BIN_LINEAR) {
        // floor(|k|) - 1 == b  <=>  (b+1)^2 <= |k|^2 < (b+2)^2    (:95)
        for (int b = 0; b < nbins; ++b) {
            lo[b] = (int64_t)(b + 1) * (b + 1);
            hi[b] = (int64_t)(b + 2) * (b + 2);
            if (hi[b] > top) hi[b] = top;
            if (lo[b] > hi[b]) lo[b] = hi[b];
        }
    } else {
        // squared edges e2[b] = (N/2)^(2b/nbins); integer range [ceil(e2[b]), ceil(e2[b+1]))
        std::vector<double> e2(nbins + 1);
        for (int b = 0; b <= nbins; ++b) e2[b] = pow((double)(N / 2), 2.0 * b / nbins);
        e2[nbins] = (double)top;
        for (int b = 0; b < nbins; ++b) {
            lo[b] = (int64_t)ceil(e2[b]);
            hi[b] = (int64_t)ceil(e2[b + 1]);
            if (lo[b] < 1) lo[b] = 1;
            if (hi[b] > top) hi[b] = top;
            if (lo[b] > hi[b]) lo[b] = hi[b];
        }
    }
    B.items.assign(nbins + 1, 0);
    for (int b = 0; b < nbins; ++b) {
        int64_t rows = 0;
        if (hi[b] > lo[b]) {
            int64_t R = (int64_t)ceil(sqrt((double)hi[b]));
            while (R * R < hi[b]) ++R;
            while (R > 0 && (R - 1) * (R - 1) >= hi[b]) --R;
            if (R > N / 2) R = N / 2;
            int64_t side = 2 * R + 1;
            rows = (nd == 3) ? side * side : side;
        }
        B.items[b + 1] = B.items[b] + (uint32_t)((rows + kPkRows - 1) / kPkRows);
    }
    B.scan = N <= 65536 && scan_lds_bytes(N, nbins, true) <= kScanLdsMax;
    if (B.scan) {
        B.hi32.assign(hi.begin(), hi.end());
        B.shell.assign(N / 2, 0);
        int b = 0;
        for (int64_t sh = 0; sh < N / 2; ++sh) {
            while (b < nbins - 1 && hi[b] <= sh * sh) ++b;
            B.shell[sh] = (uint16_t)b;
        }
    }
    return AXS_OK;
}

// work-groups of the row-order binning per lattice: about 16 rows per wave,
// at most 2048 (a function of the side only, so a lattice's sums do not
// depend on the batch it is transformed in)
static uint32_t scan_blocks(int nd, int N)
{
    const uint64_t nrows = (nd == 3) ? (uint64_t)N * N : (uint64_t)N;
    const uint64_t want = (nrows + 16 * kScanWaves - 1) / (16 * kScanWaves);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, 2048));
}

// Device tables and partial sums of the row-order binning of n spectra
// (nblk work-groups per spectrum), in the context's scratch.
struct ScanTabs {
    uint32_t *hi = nullptr;
    uint16_t *shell = nullptr;
    PkSrc *src = nullptr;
    KahanP *part = nullptr;
    long long *pcnt = nullptr;
    double *out = nullptr;
    long long *cnt = nullptr;
};

static int scan_tabs(Context *c, const BinLayout &B, int n, uint32_t nblk, ScanTabs &t)
{
    const int N = B.N;
    const size_t nb = (size_t)B.nbins;
    const uint32_t gblk = scan_blocks(c->nd, N);                  // geometry pass
    const size_t nparts = std::max<size_t>((size_t)n * nblk, gblk);
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_hi = 0, o_sh = align(nb * 4);
    const size_t o_src = o_sh + align((size_t)(N / 2) * 2);
    const size_t o_part = o_src + align((size_t)n * sizeof(PkSrc));
    const size_t o_pcnt = o_part + align(nb * nparts * sizeof(KahanP));
    const size_t o_out = o_pcnt + align(nb * gblk * sizeof(long long));
    const size_t o_cnt = o_out + align((size_t)n * nb * 8);
    const size_t need = o_cnt + align(nb * 8);
    AXS_TRY(ensure_scratch(c, need));
    char *sp = (char *)c->scratch;
    t.hi = (uint32_t *)(sp + o_hi);
    t.shell = (uint16_t *)(sp + o_sh);
    t.src = (PkSrc *)(sp + o_src);
    t.part = (KahanP *)(sp + o_part);
    t.pcnt = (long long *)(sp + o_pcnt);
    t.out = (double *)(sp + o_out);
    t.cnt = (long long *)(sp + o_cnt);
    AXS_HIP_TRY(hipMemcpyAsync(t.hi, B.hi32.data(), nb * 4, hipMemcpyHostToDevice, c->stream));
    AXS_HIP_TRY(hipMemcpyAsync(t.shell, B.shell.data(), (size_t)(N / 2) * 2,
                               hipMemcpyHostToDevice, c->stream));
    return AXS_OK;
}

// the bin table's |k| sums and mode counts: from the context, or by the
// geometry pass of bin_scan_kernel (which reads no spectrum) the first time
static int bin_geometry(Context *c, const BinLayout &B, const ScanTabs &t,
                        Context::PkGeom **out)
{
    const int N = B.N, nbins = B.nbins;
    const size_t nb = (size_t)nbins;
    for (auto &e : c->pk_geom)
        if (e.N == N && e.nd == c->nd && e.binning == B.binning && e.nbins == nbins) {
            *out = &e;
            return AXS_OK;
        }
    Context::PkGeom e;
    e.N = N;
    e.nd = c->nd;
    e.binning = B.binning;
    e.nbins = nbins;
    e.ks.assign(nb, 0.0);
    e.cnt.assign(nb, 0);
    const uint32_t nrows = (c->nd == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    const uint32_t gblk = scan_blocks(c->nd, N);
    auto kg = (c->nd == 3) ? bin_scan_kernel<3, true> : bin_scan_kernel<2, true>;
    hipLaunchKernelGGL(kg, dim3(gblk, 1), dim3(64 * kScanWaves), scan_lds_bytes(N, nbins, true),
                       c->stream, N, (const double2 *)nullptr, nbins, t.hi, t.shell, nrows,
                       t.part, t.pcnt);
    hipLaunchKernelGGL(fold_rows_kernel, dim3((unsigned)nbins, 1), dim3(256), 0, c->stream, nbins,
                       gblk, t.part, t.pcnt, t.out, t.cnt);
    AXS_HIP_TRY(hipGetLastError());
    AXS_HIP_TRY(hipMemcpyAsync(e.ks.data(), t.out, nb * 8, hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipMemcpyAsync(e.cnt.data(), t.cnt, nb * 8, hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    c->pk_geom.push_back(std::move(e));
    *out = &c->pk_geom.back();
    return AXS_OK;
}

// fold the n x nbins x nblk power partials, copy the sums, and the
// geometry's |k| sums and counts beside them
static int bin_finish(Context *c, int n, const BinLayout &B, const ScanTabs &t, uint32_t nblk,
                      const Context::PkGeom *g, std::vector<double> &hpk,
                      std::vector<double> &hks, std::vector<long long> &hcnt)
{
    const int nbins = B.nbins;
    const size_t nb = (size_t)nbins;
    hipLaunchKernelGGL(fold_rows_kernel, dim3((unsigned)nbins, (unsigned)n), dim3(256), 0,
                       c->stream, nbins, nblk, t.part, (const long long *)nullptr, t.out,
                       (long long *)nullptr);
    AXS_HIP_TRY(hipGetLastError());
    hpk.assign((size_t)n * nb, 0.0);
    AXS_HIP_TRY(hipMemcpyAsync(hpk.data(), t.out, hpk.size() * 8, hipMemcpyDeviceToHost,
                               c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    hks.resize((size_t)n * nb);
    hcnt.resize((size_t)n * nb);
    for (int q = 0; q < n; ++q) {
        std::copy(g->ks.begin(), g->ks.end(), hks.begin() + (size_t)q * nb);
        std::copy(g->cnt.begin(), g->cnt.end(), hcnt.begin() + (size_t)q * nb);
    }
    return AXS_OK;
}

// Row-order binning of the n half spectra in plan->out (rocFFT's output).
static int bin_rows(Context *c, FftPlan *plan, int n, const BinLayout &B,
                    std::vector<double> &hpk, std::vector<double> &hks,
                    std::vector<long long> &hcnt)
{
    const int N = B.N, nbins = B.nbins;
    const uint32_t nrows = (c->nd == 3) ? (uint32_t)N * (uint32_t)N : (uint32_t)N;
    const uint32_t nblk = scan_blocks(c->nd, N);
    ScanTabs t;
    AXS_TRY(scan_tabs(c, B, n, nblk, t));
    Context::PkGeom *g = nullptr;
    AXS_TRY(bin_geometry(c, B, t, &g));
    auto kd = (c->nd == 3) ? bin_scan_kernel<3, false> : bin_scan_kernel<2, false>;
    hipLaunchKernelGGL(kd, dim3(nblk, (unsigned)n), dim3(64 * kScanWaves),
                       scan_lds_bytes(N, nbins, false), c->stream, N,
                       (const double2 *)plan->out, nbins, t.hi, t.shell, nrows, t.part,
                       (long long *)nullptr);
    return bin_finish(c, n, B, t, nblk, g, hpk, hks, hcnt);
}

// sides the transform of our own takes: powers of two, 128 .. 1024 (3D) or
// 4096 (2D), with a bin table that fits the binning pass's LDS
static bool own_fft_ok(const Context *c, int N, const BinLayout &B)
{
    if (c->pk_rocfft || (N & (N - 1)) != 0 || N < 128 || N > (c->nd == 3 ? 1024 : 4096))
        return false;
    const int waves = std::max(2048, N) / 16 / 64;            // fft_bin_kernel's
    const size_t lds = (size_t)waves * B.nbins * sizeof(double) + (size_t)B.nbins * 4 +
                       (size_t)(N / 2) * 2;
    return B.scan && lds <= kScanLdsMax;
}

static int get_own(Context *c, int N, int batch, FftPlan **out)
{
    for (FftPlan *p : c->fft_plans)
        if (p->own && p->N == N && p->nd == c->nd && p->batch == batch) {
            *out = p;
            return AXS_OK;
        }
    FftPlan *p = new FftPlan();
    p->own = true;
    p->N = N;
    p->nd = c->nd;
    p->batch = batch;
    const size_t M = (size_t)N / 2;
    const size_t na = (size_t)batch * ((c->nd == 3) ? (size_t)N * M * N : M * N);
    hipError_t e = hipMalloc(&p->ya, na * sizeof(double2));
    if (e == hipSuccess && c->nd == 3) e = hipMalloc(&p->yb, na * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc(&p->tw, (size_t)N * sizeof(double2));
    if (e != hipSuccess) {
        delete p;
        set_error("power-spectrum buffers: %s", hipGetErrorString(e));
        return AXS_ERR_OUT_OF_MEMORY;
    }
    std::vector<double2> tw(N);
    for (int m = 0; m < N; ++m) {
        const long double a = 2.0L * 3.141592653589793238462643383279502884L * m / N;
        tw[m] = double2{(double)cosl(a), (double)-sinl(a)};
    }
    e = hipMemcpy(p->tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete p;
        set_error("power-spectrum twiddles: %s", hipGetErrorString(e));
        return AXS_ERR_HIP;
    }
    c->fft_plans.push_back(p);
    *out = p;
    return AXS_OK;
}

template <int N>
static void launch_own(Context *c, const FftPlan *p, int n, const BinLayout &B,
                       const ScanTabs &t, uint32_t nblk)
{
    constexpr int M = N / 2;
    if constexpr (N > 1024) {       // 3D sides up to 1024 (own_fft_ok)
    } else if (c->nd == 3) {
        const uint32_t rpl = (uint32_t)N * N;
        constexpr int RZ = kFftPts / M, RY = kFftPtsY / N;
        hipLaunchKernelGGL((fft_z_kernel<3, N>), dim3(rpl / RZ, (unsigned)n), dim3(kFftThr), 0,
                           c->stream, t.src, p->ya, p->tw);
        const uint32_t ngy = (uint32_t)n * M * (N / RY);
        hipLaunchKernelGGL((fft_y_kernel<N>),
                           dim3(std::min<uint32_t>(ngy, (uint32_t)std::max(1, c->num_cu))),
                           dim3(kFftPtsY / 16), 0, c->stream, p->ya, p->yb, p->tw, ngy);
        constexpr int TB = fft_bin_pts<N>() / 16;
        const size_t lds = (size_t)(TB / 64) * B.nbins * sizeof(double) +
                           (size_t)B.nbins * 4 + (size_t)M * 2;
        hipLaunchKernelGGL((fft_bin_kernel<3, N>), dim3(nblk, (unsigned)n), dim3(TB), lds,
                           c->stream, p->yb, p->tw, B.nbins, t.hi, t.shell, (uint32_t)(M * N),
                           t.part);
        return;
    }
    if (c->nd != 3) {
        const uint32_t rpl = (uint32_t)N;
        constexpr int RZ = kFftPts / M;
        hipLaunchKernelGGL((fft_z_kernel<2, N>), dim3(rpl / RZ, (unsigned)n), dim3(kFftThr), 0,
                           c->stream, t.src, p->ya, p->tw);
        constexpr int TB = fft_bin_pts<N>() / 16;
        const size_t lds = (size_t)(TB / 64) * B.nbins * sizeof(double) +
                           (size_t)B.nbins * 4 + (size_t)M * 2;
        hipLaunchKernelGGL((fft_bin_kernel<2, N>), dim3(nblk, (unsigned)n), dim3(TB), lds,
                           c->stream, p->ya, p->tw, B.nbins, t.hi, t.shell, (uint32_t)M, t.part);
    }
}

// The n spectra of side N by the transform of our own (own_fft_ok).
static int pk_raw_own(Context *c, const PkSrc *srcs, int n, int N, const BinLayout &B,
                      std::vector<double> &hpk, std::vector<double> &hks,
                      std::vector<long long> &hcnt)
{
    FftPlan *p = nullptr;
    AXS_TRY(get_own(c, N, n, &p));
    const uint32_t rows = (c->nd == 3) ? (uint32_t)(N / 2) * (uint32_t)N : (uint32_t)(N / 2);
    const uint32_t groups = rows / (uint32_t)(std::max(2048, N) / N);   // fft_bin_kernel's
    const uint32_t nblk = std::max<uint32_t>(1, std::min<uint32_t>(groups, 4096));
    ScanTabs t;
    AXS_TRY(scan_tabs(c, B, n, nblk, t));
    Context::PkGeom *g = nullptr;
    AXS_TRY(bin_geometry(c, B, t, &g));
    AXS_HIP_TRY(hipMemcpyAsync(t.src, srcs, (size_t)n * sizeof(PkSrc), hipMemcpyHostToDevice,
                               c->stream));
    switch (N) {
    case 128: launch_own<128>(c, p, n, B, t, nblk); break;
    case 256: launch_own<256>(c, p, n, B, t, nblk); break;
    case 512: launch_own<512>(c, p, n, B, t, nblk); break;
    case 1024: launch_own<1024>(c, p, n, B, t, nblk); break;
    case 2048: launch_own<2048>(c, p, n, B, t, nblk); break;
    default: launch_own<4096>(c, p, n, B, t, nblk); break;
    }
    AXS_HIP_TRY(hipGetLastError());
    return bin_finish(c, n, B, t, nblk, g, hpk, hks, hcnt);
}

// Raw bin sums (|F|^2, |k| and mode counts per bin, before the
// normalisation) of n real fields of side N in one batched transform:
// source q's bins go to hpk / hks / hcnt + q nbins.
static int pk_raw(Context *c, const PkSrc *srcs, int n, int N, const BinLayout &B,
                  std::vector<double> &hpk, std::vector<double> &hks,
                  std::vector<long long> &hcnt)
{
    const int nbins = B.nbins;
    if (own_fft_ok(c, N, B)) return pk_raw_own(c, srcs, n, N, B, hpk, hks, hcnt);
    FftPlan *plan = nullptr;
    AXS_TRY(get_plan(c, N, n, &plan));
    const uint32_t L = (uint32_t)((c->nd == 3) ? (int64_t)N * N * N : (int64_t)N * N);

    const size_t nb = (size_t)nbins;
    const uint32_t nitems = B.items[nb];
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_src = 0, o_lo = align((size_t)n * sizeof(PkSrc)), o_hi = o_lo + align(nb * 8);
    const size_t o_ib = o_hi + align(nb * 8);
    const size_t o_ppk = o_ib + align((nb + 1) * 4);
    const size_t o_pks = o_ppk + align((size_t)n * nitems * 16);
    const size_t o_pcnt = o_pks + align((size_t)n * nitems * 16);
    const size_t o_pk = o_pcnt + align((size_t)n * nitems * 8);
    const size_t o_ks = o_pk + align((size_t)n * nb * 8);
    const size_t o_cnt = o_ks + align((size_t)n * nb * 8);
    const size_t need = o_cnt + align((size_t)n * nb * 8);
    AXS_TRY(ensure_scratch(c, need));
    char *sp = (char *)c->scratch;
    PkSrc *dsrc = (PkSrc *)(sp + o_src);
    int64_t *dlo = (int64_t *)(sp + o_lo);
    int64_t *dhi = (int64_t *)(sp + o_hi);
    uint32_t *dib = (uint32_t *)(sp + o_ib);
    KahanP *ppk = (KahanP *)(sp + o_ppk);
    KahanP *pks = (KahanP *)(sp + o_pks);
    long long *pcnt = (long long *)(sp + o_pcnt);
    double *dpk = (double *)(sp + o_pk);
    double *dks = (double *)(sp + o_ks);
    long long *dcnt = (long long *)(sp + o_cnt);

    AXS_HIP_TRY(hipMemcpyAsync(dsrc, srcs, (size_t)n * sizeof(PkSrc), hipMemcpyHostToDevice,
                               c->stream));
    hipLaunchKernelGGL(pack_kernel, dim3((L + 255) / 256, (unsigned)n), dim3(256), 0, c->stream,
                       L, dsrc, plan->in);
    AXS_HIP_TRY(hipGetLastError());
    if (rocfft_execution_info_set_stream(plan->info, c->stream) != rocfft_status_success) {
        set_error("rocfft_execution_info_set_stream failed");
        return AXS_ERR_FFT;
    }
    void *ib[1] = {plan->in};
    void *ob[1] = {plan->out};
    if (rocfft_execute(plan->plan, ib, ob, plan->info) != rocfft_status_success) {
        set_error("rocfft_execute failed");
        return AXS_ERR_FFT;
    }
    if (B.scan) return bin_rows(c, plan, n, B, hpk, hks, hcnt);

    AXS_HIP_TRY(hipMemcpyAsync(dlo, B.lo.data(), nb * 8, hipMemcpyHostToDevice, c->stream));
    AXS_HIP_TRY(hipMemcpyAsync(dhi, B.hi.data(), nb * 8, hipMemcpyHostToDevice, c->stream));
    AXS_HIP_TRY(hipMemcpyAsync(dib, B.items.data(), (nb + 1) * 4, hipMemcpyHostToDevice,
                               c->stream));
    if (nitems > 0) {
        auto k = (c->nd == 3) ? bin_kernel<3> : bin_kernel<2>;
        hipLaunchKernelGGL(k, dim3(nitems, (unsigned)n), dim3(kPkRows), 0, c->stream, N,
                           (const double2 *)plan->out, nbins, dib, dlo, dhi, ppk, pks, pcnt);
    }
    hipLaunchKernelGGL(bin_fold_kernel, dim3((unsigned)nbins, (unsigned)n), dim3(256), 0,
                       c->stream, nbins, nitems, dib, ppk, pks, pcnt, dpk, dks, dcnt);
    AXS_HIP_TRY(hipGetLastError());
    hpk.assign((size_t)n * nb, 0.0);
    hks.assign((size_t)n * nb, 0.0);
    hcnt.assign((size_t)n * nb, 0);
    AXS_HIP_TRY(hipMemcpyAsync(hpk.data(), dpk, hpk.size() * 8, hipMemcpyDeviceToHost,
                               c->stream));
    AXS_HIP_TRY(hipMemcpyAsync(hks.data(), dks, hks.size() * 8, hipMemcpyDeviceToHost,
                               c->stream));
    AXS_HIP_TRY(hipMemcpyAsync(hcnt.data(), dcnt, hcnt.size() * 8, hipMemcpyDeviceToHost,
                               c->stream));
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    return AXS_OK;
}

// the reference's normalisation of one spectrum's raw bin sums
static void pk_finalize(const Context *c, int N, int binning, int nbins, const double *rpk,
                        const double *rks, const long long *rcnt, double *ks, double *pk,
                        int32_t *count)
{
    float Lbox = N * c->p.space_step;                           // pk_estimation.cpp:41
    for (int e = 0; e < nbins; ++e) {
        count[e] = (int32_t)rcnt[e];
        if (rcnt[e] == 0) {
            ks[e] = 0.0;
            pk[e] = 0.0;
            continue;
        }
        double kbar = rks[e] / (double)rcnt[e];                 // :118 / :196
        double p = rpk[e];
        if (binning == AXS_BIN_LINEAR) {
            if (c->nd == 2)
                p /= 2.0f * M_PI * kbar;                        // :119
            else
                p /= 4.0f * M_PI * (kbar * kbar);               // :197
        } else {
            p /= (double)rcnt[e];
        }
        ks[e] = kbar * (2.0f * M_PI / Lbox);                    // :120 / :198
        pk[e] = p;
    }
}

// source of a stored component: field 0/1 = phi1/phi2 (the state), 2/3 =
// phidot1/phidot2 (FP32: the read-back rebuilt from u)
static int pk_source(Context *c, Member &m, int comp, PkSrc *out)
{
    if (!m.has_state) {
        set_error("member has no state: call axs_set_member_state first");
        return AXS_ERR_STATE;
    }
    if (comp < 2) {
        *out = PkSrc{m.state[c->cur][PHI], c->prec == AXS_FP64 ? 0 : 1, comp};
        return AXS_OK;
    }
    FieldView fv;
    AXS_TRY(current_fields(c, m, &fv));
    *out = PkSrc{fv.vel, c->prec == AXS_FP64 ? 0 : 2, comp - 2};
    return AXS_OK;
}

// Spectra of the n members idx[0..n), all of side N, in one batched
// transform (phi2, s_run.cpp:206): member idx[q]'s bins go to
// ks / pk / count + q nbins.
static int pk_batch(Context *c, const int *idx, int n, int binning, int nbins_in, double *ks,
                    double *pk, int32_t *count)
{
    const int N = c->members[idx[0]].N;
    BinLayout B;
    AXS_TRY(bin_layout(N, c->nd, binning, nbins_in, B));
    std::vector<PkSrc> srcs(n);
    for (int q = 0; q < n; ++q) AXS_TRY(pk_source(c, c->members[idx[q]], 1, &srcs[q]));
    std::vector<double> hpk, hks;
    std::vector<long long> hcnt;
    AXS_TRY(pk_raw(c, srcs.data(), n, N, B, hpk, hks, hcnt));
    const size_t nb = (size_t)B.nbins;
    for (int q = 0; q < n; ++q)
        pk_finalize(c, N, binning, B.nbins, hpk.data() + q * nb, hks.data() + q * nb,
                    hcnt.data() + q * nb, ks + q * nb, pk + q * nb, count + q * nb);
    return AXS_OK;
}

// The spectrum of one field of a member (output_powerspec(re, im),
// pk_estimation.cpp:37-214): a real component, or the complex phi1 + i
// phi2 / phidot1 + i phidot2.  Every mode the bins take has |k| < N/2, so
// its partner -k is taken too; for a complex field f = a + i b the pair
// holds |F(k)|^2 + |F(-k)|^2 = 2 (|A(k)|^2 + |B(k)|^2), so a bin's sum is
// the sum of the two real components' sums: two real transforms in one
// batch, their raw sums added before the normalisation.
int run_powerspec_field(Context *c, Member &m, int field, int binning, int nbins, double *ks,
                        double *pk, int32_t *count)
{
    int comps[2], nc;
    switch (field) {
    case AXS_FIELD_PHI2: comps[0] = 1; nc = 1; break;
    case AXS_FIELD_PHI1: comps[0] = 0; nc = 1; break;
    case AXS_FIELD_PHI: comps[0] = 0; comps[1] = 1; nc = 2; break;
    case AXS_FIELD_PHIDOT1: comps[0] = 2; nc = 1; break;
    case AXS_FIELD_PHIDOT2: comps[0] = 3; nc = 1; break;
    case AXS_FIELD_PHIDOT: comps[0] = 2; comps[1] = 3; nc = 2; break;
    default:
        set_error("powerspec: unknown field %d", field);
        return AXS_ERR_INVALID_ARG;
    }
    BinLayout B;
    AXS_TRY(bin_layout(m.N, c->nd, binning, nbins, B));
    PkSrc srcs[2];
    for (int q = 0; q < nc; ++q) AXS_TRY(pk_source(c, m, comps[q], &srcs[q]));
    std::vector<double> hpk, hks;
    std::vector<long long> hcnt;
    AXS_TRY(pk_raw(c, srcs, nc, m.N, B, hpk, hks, hcnt));
    const size_t nb = (size_t)B.nbins;
    if (nc == 2)
        for (size_t e = 0; e < nb; ++e) hpk[e] += hpk[nb + e];
    pk_finalize(c, m.N, binning, B.nbins, hpk.data(), hks.data(), hcnt.data(), ks, pk, count);
    return AXS_OK;
}

int run_powerspec(Context *c, Member &m, int binning, int nbins, double *ks, double *pk,
                  int32_t *count)
{
    const int member = (int)(&m - c->members.data());
    return pk_batch(c, &member, 1, binning, nbins, ks, pk, count);
}

// Every member's spectrum: members of one side in batched transforms of at
// most kPkBatchBytes of input each, in member order within a side; member
// r's bins at offset sum_{q<r} nbins_q.
constexpr size_t kPkBatchBytes = (size_t)4 << 30;

int run_powerspec_all(Context *c, int binning, int nbins, double *ks, double *pk, int32_t *count)
{
    const int nm = (int)c->members.size();
    std::vector<size_t> first(nm + 1, 0);
    for (int r = 0; r < nm; ++r) {
        const int nb = (binning == AXS_BIN_LINEAR) ? c->members[r].N / 2 : nbins;
        first[r + 1] = first[r] + (size_t)(nb > 0 ? nb : 0);
    }
    std::vector<char> done(nm, 0);
    for (int r = 0; r < nm; ++r) {
        if (done[r]) continue;
        const int N = c->members[r].N;
        const size_t Lb = (size_t)c->members[r].L * 8;
        const int cap = (int)std::max<size_t>(1, kPkBatchBytes / Lb);
        std::vector<int> group;
        for (int q = r; q < nm; ++q)
            if (!done[q] && c->members[q].N == N) group.push_back(q);
        for (size_t g0 = 0; g0 < group.size(); g0 += (size_t)cap) {
            const int n = (int)std::min<size_t>((size_t)cap, group.size() - g0);
            const int nb = (binning == AXS_BIN_LINEAR) ? N / 2 : nbins;
            std::vector<double> bks((size_t)n * nb), bpk((size_t)n * nb);
            std::vector<int32_t> bcnt((size_t)n * nb);
            AXS_TRY(pk_batch(c, group.data() + g0, n, binning, nbins, bks.data(), bpk.data(),
                             bcnt.data()));
            for (int q = 0; q < n; ++q) {
                const size_t dst = first[group[g0 + q]], srco = (size_t)q * nb;
                std::copy(bks.begin() + srco, bks.begin() + srco + nb, ks + dst);
                std::copy(bpk.begin() + srco, bpk.begin() + srco + nb, pk + dst);
                std::copy(bcnt.begin() + srco, bcnt.begin() + srco + nb, count + dst);
            }
        }
        for (int q : group) done[q] = 1;
    }
    return AXS_OK;
}

void destroy_plans(Context *c)
{
    for (FftPlan *p : c->fft_plans) delete p;
    c->fft_plans.clear();
}

}  // namespace axs
