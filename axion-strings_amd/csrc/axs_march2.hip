// axs_march2.hip -- FP32 3D step with temporal blocking: two time steps per
// pass over HBM (standard 7-point stencil).
//
// A work-group owns an output tile of (4R-4) rows (j) x 60 columns (k) and
// marches along i over a chunk of CI planes.  Its four waves cover 4R rows
// of 64 consecutive k: the output tile plus a two-site halo on every side,
// read from HBM once per chunk.  For each i the pipeline is
//
//   load the state (phi, u) of plane i+2 on the 4R x 64 region,
//      drift x1 = phi + dt u                               (registers, LDS X1)
//   first step at plane i+1 on the region minus a one-site rim:
//      ker1 = F(x1, tau1), u1 = u + dt ker1, x2 = x1 + dt u1   (registers, LDS X2)
//   second step at plane i on the output tile:
//      ker2 = F(x2, tau2), u2 = u1 + dt ker2, write (x2, u2)
//
// with the drifted planes i-1, i, i+1 of each level in registers (the i
// neighbours) and the middle plane in LDS (the j and k neighbours).  Every
// intermediate value is computed by the same expressions as in the one-step
// kernels (Laplacian order centre, +-i, +-j, +-k; no contraction), so one
// pass here is bit-identical to two one-step launches.  Tiles at the
// lattice edge overhang it: loads wrap periodically, stores are masked.
//
// LDS: X1 double-buffered, X2 triple-buffered, so one barrier per plane
// separates every write from the reads of the same buffer.
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kTBW = 64;            // columns loaded per row (one wave)
constexpr int kTBOut = kTBW - 4;    // columns written per row

struct TBGeom {
    int N, CI, nTj, nTk, remap;
};

template <int R>
struct TBLds {
    static constexpr int P = 4 * R * kTBW;   // one component of one plane
    float x1[2][2][P];
    float x2[3][2][P];
};

template <int R, bool PF>
__global__ void __launch_bounds__(256)
march2_f32_kernel(TBGeom g, const fvec4 *__restrict__ in, fvec4 *__restrict__ out,
                  StepScalarsF s1, StepScalarsF s2)
{
    constexpr int ER = 4 * R;           // rows loaded
    constexpr int TJ = ER - 4;          // rows written
    __shared__ TBLds<R> sh;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int N = g.N;
    const uint32_t NN = (uint32_t)N * (uint32_t)N;

    uint32_t b = blockIdx.x;
    if (g.remap) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
    const int tk = (int)(b % (uint32_t)g.nTk);
    b /= (uint32_t)g.nTk;
    const int tj = (int)(b % (uint32_t)g.nTj);
    const int ch = (int)(b / (uint32_t)g.nTj);
    const int j0 = tj * TJ, k0 = tk * kTBOut, i0 = ch * g.CI;

    // column of this lane, wrapped; -2 <= k0 - 2 + lane < 2N since N >= 64
    int kk = k0 - 2 + lane;
    kk = (kk < 0) ? kk + N : ((kk >= N) ? kk - N : kk);
    const bool col1 = lane >= 1 && lane < kTBW - 1;
    const bool col2 = lane >= 2 && lane < kTBW - 2 && (k0 + lane - 2) < N;

    uint32_t rowoff[R];
    bool row1[R], row2[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = wave * R + r;                 // row within the loaded region
        int j = j0 - 2 + e;                         // -2 <= j < 2N since TJ < N
        j = (j < 0) ? j + N : ((j >= N) ? j - N : j);
        rowoff[r] = (uint32_t)j * (uint32_t)N + (uint32_t)kk;
        row1[r] = e >= 1 && e < ER - 1;
        row2[r] = e >= 2 && e < ER - 2 && (j0 + e - 2) < N;
    }
    auto plane = [&](int i) -> uint32_t {
        int w = (i < 0) ? i + N : ((i >= N) ? i - N : i);
        return (uint32_t)w * NN;
    };
    auto load = [&](int i, fvec4 (&raw)[R]) {
        const uint32_t po = plane(i);
#pragma unroll
        for (int r = 0; r < R; ++r) raw[r] = in[po + rowoff[r]];
    };

    // x1 of planes (i, i+1, i+2) = (x1a, x1b, x1c); u of planes i+1, i+2;
    // x2 of planes (i-1, i, i+1) = (x2a, x2b, x2c); u1 of planes i, i+1
    float x1a[R][2], x1b[R][2], x1c[R][2];
    float u0b[R][2], u0c[R][2];
    float x2a[R][2], x2b[R][2], x2c[R][2];
    float u1b[R][2], u1c[R][2];
    fvec4 raw[R], pf[R];

    const float c0a = s1.c[0], c0b = s2.c[0];

    // prologue: planes i0-2 (registers) and i0-1 (registers, X1)
    load(i0 - 2, raw);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        x1a[r][0] = raw[r].x + s1.dt * raw[r].z;
        x1a[r][1] = raw[r].y + s1.dt * raw[r].w;
    }
    load(i0 - 1, raw);
    {
        float *W1 = sh.x1[(i0 - 1) & 1][0], *W2 = sh.x1[(i0 - 1) & 1][1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            x1b[r][0] = raw[r].x + s1.dt * raw[r].z;
            x1b[r][1] = raw[r].y + s1.dt * raw[r].w;
            u0b[r][0] = raw[r].z;
            u0b[r][1] = raw[r].w;
            const int o = (wave * R + r) * kTBW + lane;
            W1[o] = x1b[r][0];
            W2[o] = x1b[r][1];
        }
    }
    if (PF) load(i0, pf);
    __syncthreads();

    int w2 = 0;   // X2 slot of plane i+1; plane i sits in slot w2 - 1 (mod 3)
    for (int it = -2; it < g.CI; ++it) {
        const int i = i0 + it;
        // plane i+2: x1 and u
        if (PF) {
#pragma unroll
            for (int r = 0; r < R; ++r) raw[r] = pf[r];
            if (it + 1 < g.CI) load(i + 3, pf);
        } else {
            load(i + 2, raw);
        }
        {
            float *W1 = sh.x1[(i + 2) & 1][0], *W2 = sh.x1[(i + 2) & 1][1];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                x1c[r][0] = raw[r].x + s1.dt * raw[r].z;
                x1c[r][1] = raw[r].y + s1.dt * raw[r].w;
                u0c[r][0] = raw[r].z;
                u0c[r][1] = raw[r].w;
                const int o = (wave * R + r) * kTBW + lane;
                W1[o] = x1c[r][0];
                W2[o] = x1c[r][1];
            }
        }
        // first step at plane i+1 (its x1 has been in X1 since the last barrier)
        {
            const float *L1 = sh.x1[(i + 1) & 1][0], *L2 = sh.x1[(i + 1) & 1][1];
            float *W1 = sh.x2[w2][0], *W2 = sh.x2[w2][1];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!(row1[r] && col1)) continue;
                const int o = (wave * R + r) * kTBW + lane;
                float a = c0a * x1b[r][0], c = c0a * x1b[r][1];
                a = a + x1c[r][0];
                c = c + x1c[r][1];        // i+1 neighbour
                a = a + x1a[r][0];
                c = c + x1a[r][1];        // i-1 neighbour
                a = a + L1[o + kTBW];
                c = c + L2[o + kTBW];     // j+1
                a = a + L1[o - kTBW];
                c = c + L2[o - kTBW];     // j-1
                a = a + L1[o + 1];
                c = c + L2[o + 1];        // k+1
                a = a + L1[o - 1];
                c = c + L2[o - 1];        // k-1
                const float x = x1b[r][0], y = x1b[r][1];
                const float S = x * x + y * y - s1.tau2 + s1.T03;
                const float k1 = s1.alpha * a + s1.cl * x * S;
                const float k2 = s1.alpha * c + s1.cl * y * S;
                const float ua = u0b[r][0] + s1.dt * k1;
                const float ub = u0b[r][1] + s1.dt * k2;
                u1c[r][0] = ua;
                u1c[r][1] = ub;
                x2c[r][0] = x + s2.dt * ua;
                x2c[r][1] = y + s2.dt * ub;
                W1[o] = x2c[r][0];
                W2[o] = x2c[r][1];
            }
        }
        __syncthreads();
        // second step at plane i (its x2 was written one plane ago)
        if (it >= 0) {
            const int rs = (w2 == 0) ? 2 : w2 - 1;
            const float *L1 = sh.x2[rs][0], *L2 = sh.x2[rs][1];
            const uint32_t po = plane(i);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!(row2[r] && col2)) continue;
                const int o = (wave * R + r) * kTBW + lane;
                float a = c0b * x2b[r][0], c = c0b * x2b[r][1];
                a = a + x2c[r][0];
                c = c + x2c[r][1];
                a = a + x2a[r][0];
                c = c + x2a[r][1];
                a = a + L1[o + kTBW];
                c = c + L2[o + kTBW];
                a = a + L1[o - kTBW];
                c = c + L2[o - kTBW];
                a = a + L1[o + 1];
                c = c + L2[o + 1];
                a = a + L1[o - 1];
                c = c + L2[o - 1];
                const float x = x2b[r][0], y = x2b[r][1];
                const float S = x * x + y * y - s2.tau2 + s2.T03;
                const float k1 = s2.alpha * a + s2.cl * x * S;
                const float k2 = s2.alpha * c + s2.cl * y * S;
                fvec4 v;
                v.x = x;
                v.y = y;
                v.z = u1b[r][0] + s2.dt * k1;
                v.w = u1b[r][1] + s2.dt * k2;
                store_nt(out + po + rowoff[r], v);
            }
        }
        w2 = (w2 == 2) ? 0 : w2 + 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                x1a[r][q] = x1b[r][q];
                x1b[r][q] = x1c[r][q];
                u0b[r][q] = u0c[r][q];
                x2a[r][q] = x2b[r][q];
                x2b[r][q] = x2c[r][q];
                u1b[r][q] = u1c[r][q];
            }
        }
    }
}

bool march2_supported(const Context *c, int N)
{
    return c->nd == 3 && c->prec == AXS_FP32 && c->p.stencil_setting == 0 && N >= 64;
}

template <int R>
static int launch_tb(Context *c, int N, const fvec4 *in, fvec4 *out, const StepScalarsF &a,
                     const StepScalarsF &b, int ci, bool pf)
{
    constexpr int TJ = 4 * R - 4;
    TBGeom g;
    g.N = N;
    g.nTj = (N + TJ - 1) / TJ;
    g.nTk = (N + kTBOut - 1) / kTBOut;
    if (ci <= 0 || N % ci != 0) {
        // short chunks keep neighbouring tiles in step (measured at 256^3
        // and 1024^3: 16 and 32 planes beat 64 and 128)
        ci = (N >= 512) ? 32 : 16;
        while (N % ci != 0) --ci;
    }
    g.CI = ci;
    const uint32_t nb = (uint32_t)g.nTj * (uint32_t)g.nTk * (uint32_t)(N / ci);
    g.remap = (nb % 8 == 0) ? 1 : 0;
    if (pf)
        hipLaunchKernelGGL((march2_f32_kernel<R, true>), dim3(nb), dim3(256), 0, c->stream, g, in,
                           out, a, b);
    else
        hipLaunchKernelGGL((march2_f32_kernel<R, false>), dim3(nb), dim3(256), 0, c->stream, g,
                           in, out, a, b);
    return AXS_OK;
}

int launch_march2(Context *c, Member &m, int src, int dst, const StepScalars &a,
                  const StepScalars &b)
{
    const StepScalarsF fa = to_float(a), fb = to_float(b);
    const fvec4 *in = (const fvec4 *)m.state[src][PHI];
    fvec4 *out = (fvec4 *)m.state[dst][PHI];
    const bool pf = (c->tune.flags & kTunePrefetch) != 0;
    int rows = c->tune.rtb;
    if (rows == 0) rows = (m.N >= 512) ? 6 : 4;   // measured best at 1024^3 / 256^3
    switch (rows) {
    case 4: return launch_tb<4>(c, m.N, in, out, fa, fb, c->tune.ci32, pf);
    case 8: return launch_tb<8>(c, m.N, in, out, fa, fb, c->tune.ci32, pf);
    default: return launch_tb<6>(c, m.N, in, out, fa, fb, c->tune.ci32, pf);
    }
}

}  // namespace axs

namespace axs {

bool two_step_active(const Context *c)
{
    if (c->tune.depth != 2 || c->tune.kernel == AXS_KERNEL_SIMPLE) return false;
    for (const Member &m : c->members)
        if (!march2_supported(c, m.N)) return false;
    return true;
}

int launch_two_steps(Context *c, const StepScalars &a, const StepScalars &b)
{
    const int src = c->cur, dst = c->cur ^ 1;
    for (Member &m : c->members) AXS_TRY(launch_march2(c, m, src, dst, a, b));
    AXS_HIP_TRY(hipGetLastError());
    c->cur = dst;
    return AXS_OK;
}

}  // namespace axs
