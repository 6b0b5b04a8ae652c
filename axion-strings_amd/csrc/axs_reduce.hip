// axs_reduce.hip -- whole-lattice reductions in compensated FP64:
//   background sums of s_run.cpp:135-150 (six quantities);
//   energy components (kinetic, gradient, potential).
//
// Each work-item keeps Neumaier (sum, compensation) pairs over a fixed
// subset of sites (background: a strided set; energy: columns along the
// slowest axis, marched so that the slow-axis neighbour stays in
// registers); waves combine pairs with cross-lane shuffles, the block
// writes one pair per quantity, and a single-block pass folds the block
// pairs in block order.  The partition depends only on the lattice size,
// so results are identical from run to run.
#include <algorithm>

#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 2048;

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

int run_background(Context *c, Member &m, double out[6])
{
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

int run_energy(Context *c, Member &m, double out[3])
{
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
    AXS_TRY(ensure_scratch(c, part_b + 64));
    Kahan *part = (Kahan *)c->scratch;
    double *dout = (double *)((char *)c->scratch + part_b);
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
    double dx = (double)c->p.space_step;
    double vol = (c->nd == 3) ? dx * dx * dx : dx * dx;
    out[0] = raw[0] * vol;
    out[1] = raw[1] / (dx * dx) * vol;
    out[2] = raw[2] * vol;
    return AXS_OK;
}

}  // namespace axs
