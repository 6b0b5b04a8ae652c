// axs_march2d.hip -- FP32 2D step with temporal blocking: D time steps per
// pass over HBM (5-point stencil), for single lattices and ensembles.
//
// Layout m = i + N*j (i contiguous).  One wave owns a strip of 64
// consecutive columns i and marches along j over a chunk of CJ rows; level 0
// is the state loaded from HBM, level l (1..D) the lattice after l steps,
// computed on the strip minus an l-column rim, so the wave writes 64 - 2D
// columns.  Per row j:
//
//   load (phi, u) of row j+D, drift x1 = phi + dt u           (level 0 -> x1)
//   for l = 1..D, at row q = j+D-l:
//      ker = F(x_l, tau_l);  u_l = u_{l-1} + dt ker;  x_{l+1} = x_l + dt u_l
//   write (x_D, u_D) of row j
//
// Each level keeps x_l of rows q-1 and q and u_{l-1} of row q in registers;
// row q+1 is the value the level below has just produced (the j
// neighbours), and the i neighbours are the neighbouring lanes, read with
// cross-lane shuffles -- no LDS, no barriers, every wave independent.  The
// Laplacian sums in the one-step kernels' order (centre, +i, -i, +j, -j)
// with the same expressions, so a pass is bit-identical to D one-step
// launches.  Work items (member, strip, chunk) are dealt to waves from a
// persistent list; strips at the lattice edge overhang it (columns wrap
// periodically, stores are masked).
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kW2 = 64;           // columns per strip (one wave)
constexpr int kWaves2D = 4;       // waves per work-group

bool march2d_supported(const Context *c, int N)
{
    return c->nd == 2 && c->prec == AXS_FP32 && c->p.stencil_setting == 0 &&
           c->p.eom == AXS_EOM_FIELD_RESCALED && c->p.laplacian == AXS_LAPLACIAN_STENCIL &&
           N >= kW2;
}

bool march2d_depth_built(int D) { return D == 2 || D == 3 || D == 4 || D == 6 || D == 8; }

template <int D, bool PF>
__device__ __forceinline__ void march2d_strip(const int N, const int CJ, const int ti,
                                              const int ch, const fvec4 *__restrict__ in,
                                              fvec4 *__restrict__ out, const TBScalars<D> &sc)
{
    constexpr int TI = kW2 - 2 * D;      // columns written
    const int lane = threadIdx.x & 63;
    const int i0 = ti * TI, j0 = ch * CJ;
    int ii = i0 - D + lane;              // -D <= ii < 2N for N >= 64
    ii = (ii < 0) ? ii + N : ((ii >= N) ? ii - N : ii);
    const bool colout = lane >= D && lane < kW2 - D && (i0 + lane - D) < N;

    auto row = [&](int j) -> uint32_t {
        int w = (j < 0) ? j + N : ((j >= N) ? j - N : j);
        return (uint32_t)w * (uint32_t)N + (uint32_t)ii;
    };

    // per level l = 1..D (index l-1): x_l of rows q-1 (xm), q (x0), u_{l-1} of row q
    float xm[D][2], x0[D][2], uq[D][2];
#pragma unroll
    for (int l = 0; l < D; ++l)
#pragma unroll
        for (int q = 0; q < 2; ++q) xm[l][q] = x0[l][q] = uq[l][q] = 0.f;
    fvec4 raw, pf;
    if (PF) pf = in[row(j0 - D)];

    for (int it = -2 * D; it < CJ; ++it) {
        const int j = j0 + it;
        if (PF) {
            raw = pf;
            if (it + 1 < CJ) pf = in[row(j + D + 1)];
        } else {
            raw = in[row(j + D)];
        }
        const StepScalarsF &s0 = sc.s[0];
        float fx[2], fu[2];                 // x_l and u_{l-1} of row q+1, handed up
        fx[0] = raw.x + s0.dt * raw.z;
        fx[1] = raw.y + s0.dt * raw.w;
        fu[0] = raw.z;
        fu[1] = raw.w;
#pragma unroll
        for (int l = 1; l <= D; ++l) {
            const StepScalarsF &s = sc.s[l - 1];
            const bool active = it >= 2 * l - 2 * D;   // x_l of rows q-1, q, q+1 exist
            float nx[2] = {0.f, 0.f}, nu[2] = {0.f, 0.f};
            if (active) {
                const float x = x0[l - 1][0], y = x0[l - 1][1];
                // i neighbours from the adjacent lanes (all lanes take part)
                const float xr = __shfl_down(x, 1), yr = __shfl_down(y, 1);
                const float xl = __shfl_up(x, 1), yl = __shfl_up(y, 1);
                if (lane >= l && lane < kW2 - l) {
                    float a = s.c[0] * x, c = s.c[0] * y;
                    a = a + xr;
                    c = c + yr;                 // i+1
                    a = a + xl;
                    c = c + yl;                 // i-1
                    a = a + fx[0];
                    c = c + fx[1];              // j+1
                    a = a + xm[l - 1][0];
                    c = c + xm[l - 1][1];       // j-1
                    const float S = x * x + y * y - s.tau2 + s.T03;
                    const float k1 = s.alpha * a + s.cl * x * S;
                    const float k2 = s.alpha * c + s.cl * y * S;
                    const float ua = uq[l - 1][0] + s.dt * k1;
                    const float ub = uq[l - 1][1] + s.dt * k2;
                    if (l == D) {
                        if (it >= 0 && colout) {
                            fvec4 v;
                            v.x = x;
                            v.y = y;
                            v.z = ua;
                            v.w = ub;
                            store_nt(out + row(j), v);
                        }
                    } else {
                        const StepScalarsF &sn = sc.s[l < D ? l : D - 1];
                        nx[0] = x + sn.dt * ua;
                        nx[1] = y + sn.dt * ub;
                        nu[0] = ua;
                        nu[1] = ub;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                xm[l - 1][q] = x0[l - 1][q];
                x0[l - 1][q] = fx[q];
                uq[l - 1][q] = fu[q];
                fx[q] = nx[q];
                fu[q] = nu[q];
            }
        }
    }
}

// one work item per wave: (member, strip, chunk) from the member table
template <int D, bool PF>
__global__ void __launch_bounds__(64 * kWaves2D)
march2d_kernel(const BatchMember *__restrict__ mem, int n_members, uint32_t n_items,
               TBScalars<D> sc)
{
    const uint32_t wave = blockIdx.x * kWaves2D + (threadIdx.x >> 6);
    const uint32_t stride = gridDim.x * kWaves2D;
    for (uint32_t item = wave; item < n_items; item += stride) {
        int lo = 0, hi = n_members - 1;        // last member starting at or before item
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (mem[mid].item_begin <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchMember &m = mem[lo];
        const uint32_t t = item - m.item_begin;
        const int ti = (int)(t % (uint32_t)m.nTj);
        const int ch = (int)(t / (uint32_t)m.nTj);
        march2d_strip<D, PF>(m.N, m.CI, ti, ch, (const fvec4 *)m.in[PHI], (fvec4 *)m.out[PHI],
                             sc);
    }
}

// rows per chunk: the configured count if it divides N, else 64 (measured
// best for the 512-lattice ensemble of sides 128..1024: fewer halo rows per
// chunk than 16 or 32)
static int pick_cj(const Context *c, int N)
{
    int cj = c->tune.ci32;
    if (cj > 0 && N % cj == 0) return cj;
    cj = 64;
    while (N % cj != 0) --cj;
    return cj;
}

int prepare_march2d_items(Context *c)
{
    const int D = multi_step_depth(c);
    if (D < 2) return AXS_OK;
    const int key = (D * 100 + 2) * 10000 + c->tune.ci32;
    if (key == c->mbatch_key && c->mbatch_tab[0]) return AXS_OK;
    const int TI = kW2 - 2 * D;
    std::vector<BatchMember> tab[2];
    uint32_t items = 0;
    for (const Member &m : c->members) {
        BatchMember b{};
        b.N = m.N;
        b.L = (uint32_t)m.L;
        b.item_begin = items;
        b.nTj = (m.N + TI - 1) / TI;       // strips
        b.nTk = 1;
        b.CI = pick_cj(c, m.N);            // rows per chunk
        items += (uint32_t)b.nTj * (uint32_t)(m.N / b.CI);
        for (int h = 0; h < 2; ++h) {
            b.in[PHI] = m.state[h][PHI];
            b.out[PHI] = m.state[h ^ 1][PHI];
            tab[h].push_back(b);
        }
    }
    for (int h = 0; h < 2; ++h) {
        if (c->mbatch_tab[h]) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(c->mbatch_tab[h]);
            c->mbatch_tab[h] = nullptr;
        }
        AXS_HIP_TRY(hipMalloc(&c->mbatch_tab[h], c->members.size() * sizeof(BatchMember)));
        AXS_HIP_TRY(hipMemcpyAsync(c->mbatch_tab[h], tab[h].data(),
                                   tab[h].size() * sizeof(BatchMember), hipMemcpyHostToDevice,
                                   c->stream));
    }
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    c->mbatch_items = items;
    c->mbatch_key = key;
    return AXS_OK;
}

template <int D>
static int launch_d(Context *c, int src, const StepScalars *s, bool pf)
{
    TBScalars<D> sc;
    for (int l = 0; l < D; ++l) sc.s[l] = to_float(s[l]);
    auto kern = pf ? march2d_kernel<D, true> : march2d_kernel<D, false>;
    int per_cu = 0;
    AXS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * kWaves2D, 0));
    const uint32_t cap = (uint32_t)c->num_cu * (uint32_t)(per_cu > 0 ? per_cu : 1);
    const uint32_t need = (c->mbatch_items + kWaves2D - 1) / kWaves2D;
    const uint32_t grid = need < cap ? need : cap;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kWaves2D), 0, c->stream, c->mbatch_tab[src],
                       (int)c->members.size(), c->mbatch_items, sc);
    return AXS_OK;
}

int launch_march2d(Context *c, int src, const StepScalars *s, int depth)
{
    if (!c->mbatch_tab[src]) {
        set_error("2D multi-step work list missing");
        return AXS_ERR_STATE;
    }
    const bool pf = (c->tune.flags & kTunePrefetch) != 0;
    switch (depth) {
    case 2: return launch_d<2>(c, src, s, pf);
    case 3: return launch_d<3>(c, src, s, pf);
    case 4: return launch_d<4>(c, src, s, pf);
    case 6: return launch_d<6>(c, src, s, pf);
    case 8: return launch_d<8>(c, src, s, pf);
    default:
        set_error("2D depth %d is not built", depth);
        return AXS_ERR_UNSUPPORTED;
    }
}

}  // namespace axs
