// axs_march2d.hip -- 2D step with temporal blocking: D time steps per pass
// over HBM (5-point stencil), for single lattices and ensembles, in the FP32
// form and the bit-exact FP64 form (field-rescaled or Hamiltonian).
//
// Layout m = i + N*j (i contiguous).  One wave owns a strip of 64
// consecutive columns i and marches along j over a chunk of CJ rows; level 0
// is the state loaded from HBM, level l (1..D) the lattice after l steps,
// computed on the strip minus an l-column rim, so the wave writes 64 - 2D
// columns.  Per row j:
//
//   load the state of row j+D and drift it                (level 0 -> x_1)
//   for l = 1..D, at row q = j+D-l: one step of the form (axs_forms.hpp)
//   write the state of row j
//
// Each level keeps x_l of rows q-1 and q and the site's own data (FP32: u;
// FP64: phidot, ker) of row q in registers; row q+1 is the value the level
// below has just produced (the j neighbours), and the i neighbours are the
// neighbouring lanes, read with cross-lane shuffles -- no LDS, no barriers,
// every wave independent.  The Laplacian sums in the one-step kernels'
// order (centre, +i, -i, +j, -j) with the same expressions, so a pass is
// bit-identical to D one-step launches.  Work items (member, strip, chunk)
// are dealt to waves from a persistent list; strips at the lattice edge
// overhang it (columns wrap periodically, stores are masked).
#include <type_traits>

#include "axs_forms.hpp"
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kW2 = 64;           // columns per strip (one wave)
constexpr int kWaves2D = 4;       // waves per work-group

bool march2d_supported(const Context *c, int N)
{
    return c->nd == 2 && c->p.stencil_setting == 0 && c->p.laplacian == AXS_LAPLACIAN_STENCIL &&
           N >= kW2;
}

bool march2d_depth_built(int D) { return D == 2 || D == 3 || D == 4 || D == 6 || D == 8; }

template <class Tr, int D, bool PF>
__device__ __forceinline__ void march2d_strip(const int N, const int CJ, const int ti,
                                              const int ch, const typename Tr::In &in,
                                              const typename Tr::Out &out,
                                              const LevelScalars<typename Tr::Sc, D> &sc)
{
    using T = typename Tr::T;
    using Own = typename Tr::Own;
    using Raw = typename Tr::Raw;
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

    // per level l = 1..D (index l-1): x_l of rows q-1 (xm), q (x0), own data of row q
    T xm[D][2], x0[D][2];
    Own oq[D];
#pragma unroll
    for (int l = 0; l < D; ++l) {
#pragma unroll
        for (int q = 0; q < 2; ++q) xm[l][q] = x0[l][q] = (T)0;
        oq[l] = Own{};
    }
    Raw raw, pf;
    if (PF) Tr::fetch(in, row(j0 - D), pf);

    for (int it = -2 * D; it < CJ; ++it) {
        const int j = j0 + it;
        if (PF) {
            raw = pf;
            if (it + 1 < CJ) Tr::fetch(in, row(j + D + 1), pf);
        } else {
            Tr::fetch(in, row(j + D), raw);
        }
        T fx[2];                           // x_l of row q+1, handed up
        Own fo;                            // own data of row q+1
        Tr::drift0(raw, sc.s[0], fx[0], fx[1], fo);
#pragma unroll
        for (int l = 1; l <= D; ++l) {
            const typename Tr::Sc &s = sc.s[l - 1];
            const typename Tr::Sc &sn = sc.s[l < D ? l : D - 1];
            const bool active = it >= 2 * l - 2 * D;   // x_l of rows q-1, q, q+1 exist
            T nx[2] = {(T)0, (T)0};
            Own no{};
            if (active) {
                const T x = x0[l - 1][0], y = x0[l - 1][1];
                // i neighbours from the adjacent lanes (all lanes take part)
                const T xr = __shfl_down(x, 1), yr = __shfl_down(y, 1);
                const T xl = __shfl_up(x, 1), yl = __shfl_up(y, 1);
                if (lane >= l && lane < kW2 - l) {
                    T a = s.c[0] * x, c = s.c[0] * y;
                    a = a + xr;
                    c = c + yr;                 // i+1
                    a = a + xl;
                    c = c + yl;                 // i-1
                    a = a + fx[0];
                    c = c + fx[1];              // j+1
                    a = a + xm[l - 1][0];
                    c = c + xm[l - 1][1];       // j-1
                    Own on;
                    T ex, ey;
                    Tr::level(x, y, a, c, oq[l - 1], s, sn, on, ex, ey);
                    if (l == D) {
                        if (it >= 0 && colout) Tr::store(out, row(j), x, y, on);
                    } else {
                        nx[0] = ex;
                        nx[1] = ey;
                        no = on;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                xm[l - 1][q] = x0[l - 1][q];
                x0[l - 1][q] = fx[q];
                fx[q] = nx[q];
            }
            oq[l - 1] = fo;
            fo = no;
        }
    }
}

template <class Tr>
__device__ __forceinline__ typename Tr::In strip_in(void *const *p)
{
    if constexpr (std::is_same<Tr, F32Form>::value)
        return {(const fvec4 *)p[PHI]};
    else
        return {(const dvec2 *)p[PHI], (const dvec2 *)p[VEL], (const dvec2 *)p[KER]};
}
template <class Tr>
__device__ __forceinline__ typename Tr::Out strip_out(void *const *p)
{
    if constexpr (std::is_same<Tr, F32Form>::value)
        return {(fvec4 *)p[PHI]};
    else
        return {(dvec2 *)p[PHI], (dvec2 *)p[VEL], (dvec2 *)p[KER]};
}

// one work item per wave: (member, strip, chunk) from the member table
template <class Tr, int D, bool PF>
__global__ void __launch_bounds__(64 * kWaves2D)
march2d_kernel(const BatchMember *__restrict__ mem, int n_members, uint32_t n_items,
               LevelScalars<typename Tr::Sc, D> sc)
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
        march2d_strip<Tr, D, PF>(m.N, m.CI, ti, ch, strip_in<Tr>(m.in), strip_out<Tr>(m.out), sc);
    }
}

// rows per chunk: the configured count if it divides N, else 128 (measured
// best for the 512-lattice ensemble of sides 128..1024 in both forms: fewer
// halo rows per chunk than 32 or 64)
static int pick_cj(const Context *c, int N)
{
    int cj = c->tune.ci32;
    if (cj > 0 && N % cj == 0) return cj;
    cj = 128;
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
            for (int a = 0; a < 3; ++a) {
                b.in[a] = m.state[h][a];
                b.out[a] = m.state[h ^ 1][a];
            }
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

template <class Sc>
static Sc strip_scalars(const StepScalars &s)
{
    if constexpr (std::is_same<Sc, StepScalarsF>::value)
        return to_float(s);
    else
        return s;
}

template <class Tr, int D>
static int launch_d(Context *c, int src, const StepScalars *s, bool pf)
{
    LevelScalars<typename Tr::Sc, D> sc;
    for (int l = 0; l < D; ++l) sc.s[l] = strip_scalars<typename Tr::Sc>(s[l]);
    auto kern = pf ? march2d_kernel<Tr, D, true> : march2d_kernel<Tr, D, false>;
    int per_cu = 0;
    AXS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * kWaves2D, 0));
    const uint32_t cap = (uint32_t)c->num_cu * (uint32_t)(per_cu > 0 ? per_cu : 1);
    const uint32_t need = (c->mbatch_items + kWaves2D - 1) / kWaves2D;
    const uint32_t grid = need < cap ? need : cap;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kWaves2D), 0, c->stream, c->mbatch_tab[src],
                       (int)c->members.size(), c->mbatch_items, sc);
    return AXS_OK;
}

template <class Tr>
static int launch_form2d(Context *c, int src, const StepScalars *s, int depth, bool pf)
{
    switch (depth) {
    case 2: return launch_d<Tr, 2>(c, src, s, pf);
    case 3: return launch_d<Tr, 3>(c, src, s, pf);
    case 4: return launch_d<Tr, 4>(c, src, s, pf);
    case 6: return launch_d<Tr, 6>(c, src, s, pf);
    case 8: return launch_d<Tr, 8>(c, src, s, pf);
    default:
        set_error("2D depth %d is not built", depth);
        return AXS_ERR_UNSUPPORTED;
    }
}

int launch_march2d(Context *c, int src, const StepScalars *s, int depth)
{
    if (!c->mbatch_tab[src]) {
        set_error("2D multi-step work list missing");
        return AXS_ERR_STATE;
    }
    const bool pf = (c->tune.flags & kTunePrefetch) != 0;
    if (c->prec == AXS_FP32) return launch_form2d<F32Form>(c, src, s, depth, pf);
    if (c->p.eom == AXS_EOM_HAMILTONIAN) return launch_form2d<F64FormT<true>>(c, src, s, depth, pf);
    return launch_form2d<F64FormT<false>>(c, src, s, depth, pf);
}

}  // namespace axs
