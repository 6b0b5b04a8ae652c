// axs_march2d.hpp -- the 2D multi-step (temporally blocked) step kernels,
// templated on the storage form, the steps per pass D and the stencil
// radius RAD (stencil_setting + 1).
//
// Layout m = i + N*j (i contiguous).  One wave owns a strip of 64
// consecutive columns i and marches along j over a chunk of CJ rows; level 0
// is the state loaded from HBM, level l (1..D) the lattice after l steps,
// computed on the strip minus an l*RAD-column rim, so the wave writes
// 64 - 2 D RAD columns.  Per row j:
//
//   load the state of row j + D RAD and drift it           (level 0 -> x_1)
//   for l = 1..D, at row q = j + (D - l) RAD: one step of the form
//      (axs_forms.hpp)
//   write the state of row j
//
// Each level keeps x_l of rows q-RAD .. q+RAD-1 and the site's own data
// (FP32: u; FP64: phidot, ker) of rows q .. q+RAD-1 in registers; row q+RAD
// is the value the level below has just produced (the j neighbours), and
// the i neighbours are the neighbouring lanes, read with cross-lane
// shuffles -- no LDS, no barriers, every wave independent.  Every lane
// computes every level (a lane mask would save no issue cycles; values
// outside a level's region never reach a stored site).  The Laplacian
// sums in the one-step kernels' order (centre, then for each distance m:
// FP64 +i, -i, +j, -j, the FP32 form +j, -j, +i, -i, each times the
// stencil coefficient) with the same expressions, so a pass is
// bit-identical to D one-step launches.  Work
// items (member, strip, chunk) are dealt to waves from a persistent list;
// strips at the lattice edge overhang it (columns wrap periodically, stores
// are masked).
#pragma once

#include <type_traits>

#include "axs_forms.hpp"
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kW2 = 64;           // columns per strip (one wave)
constexpr int kWaves2D = 4;       // waves per work-group

// rows per unrolled block of the strip loop: three (measured on the
// 512-lattice ensemble: FP32 4%, FP64 5% faster than one; two is between)
// unless the rings and FIFOs hold many registers per lane
template <class Tr, int D, int RAD>
constexpr int strip_unroll()
{
    constexpr int words = D * RAD * (4 * (int)sizeof(typename Tr::T) +
                                     (int)sizeof(typename Tr::Own)) / 4;
    return words <= 96 ? 3 : 1;
}

template <class Tr, int D, int RAD, bool PF>
__device__ __forceinline__ void march2d_strip(const int N, const int CJ, const int ti,
                                              const int ch, const typename Tr::In &in,
                                              const typename Tr::Out &out,
                                              const LevelScalars<typename Tr::Sc, D> &sc)
{
    using T = typename Tr::T;
    using Own = typename Tr::Own;
    using Raw = typename Tr::Raw;
    constexpr int H = D * RAD;           // halo width
    constexpr int TI = kW2 - 2 * H;      // columns written
    static_assert(TI > 0, "strip shape");
    const int lane = threadIdx.x & 63;
    const int i0 = ti * TI, j0 = ch * CJ;
    int ii = i0 - H + lane;              // -H <= ii < 2N for N >= 64
    ii = (ii < 0) ? ii + N : ((ii >= N) ? ii - N : ii);
    const bool colout = lane >= H && lane < kW2 - H && (i0 + lane - H) < N;

    auto row = [&](int j) -> uint32_t {
        int w = (j < 0) ? j + N : ((j >= N) ? j - N : j);
        return (uint32_t)w * (uint32_t)N + (uint32_t)ii;
    };

    // per level l = 1..D (index l-1): x_l of rows q-RAD .. q+RAD-1 (ring)
    // and the own data of rows q .. q+RAD-1 (fifo)
    T ring[D][2 * RAD][2];
    Own fifo[D][RAD];
#pragma unroll
    for (int l = 0; l < D; ++l) {
#pragma unroll
        for (int p = 0; p < 2 * RAD; ++p) ring[l][p][0] = ring[l][p][1] = (T)0;
#pragma unroll
        for (int p = 0; p < RAD; ++p) fifo[l][p] = Own{};
    }
    Raw raw, pf;
    const typename Tr::Vc vc = Tr::vconst(sc.s[0]);
    if (PF) Tr::fetch(in, row(j0 - H), pf);

    auto row_step = [&](const int it) __attribute__((always_inline)) {
        const int j = j0 + it;
        if (PF) {
            raw = pf;
            if (it + 1 < CJ) Tr::fetch(in, row(j + H + 1), pf);
        } else {
            Tr::fetch(in, row(j + H), raw);
        }
        T fx[2];                           // x_l of row q+RAD, handed up
        Own fo;                            // own data of row q+RAD
        Tr::drift0(raw, sc.s[0], fx[0], fx[1], fo);
#pragma unroll
        for (int l = 1; l <= D; ++l) {
            const typename Tr::Sc &s = sc.s[l - 1];
            const typename Tr::Sc &sn = sc.s[l < D ? l : D - 1];
            // every lane computes from the first row on: results of lanes
            // outside level l's region, or of rows before x_l of rows
            // q-RAD .. q+RAD exist (it < 2 RAD (l - D)), only reach sites
            // outside level l+1's region, which are never stored
            T nx[2];
            Own no;
            {
                const T x = ring[l - 1][RAD][0], y = ring[l - 1][RAD][1];
                // i neighbours from the lanes m away (all lanes take part)
                T xr[RAD], yr[RAD], xl[RAD], yl[RAD];
#pragma unroll
                for (int m = 1; m <= RAD; ++m) {
                    // m lanes away: m chained one-lane DPP wave shifts
                    xr[m - 1] = lane_next(m == 1 ? x : xr[m - 2]);
                    yr[m - 1] = lane_next(m == 1 ? y : yr[m - 2]);
                    xl[m - 1] = lane_prev(m == 1 ? x : xl[m - 2]);
                    yl[m - 1] = lane_prev(m == 1 ? y : yl[m - 2]);
                }
                {
                    T a = Tr::centre(s, x), c = Tr::centre(s, y);
#pragma unroll
                    for (int m = 1; m <= RAD; ++m) {
                        // the 5-point stencil's coefficient is 1: its
                        // products are exact and left out
                        const T cm = (T)s.c[m];
                        auto add = [&](T acc, T v) -> T {
                            return (RAD == 1) ? acc + v : Tr::acc(acc, cm, v);
                        };
                        const int up = (RAD + m < 2 * RAD) ? RAD + m : 2 * RAD - 1;
                        const T ux = (m == RAD) ? fx[0] : ring[l - 1][up][0];
                        const T uy = (m == RAD) ? fx[1] : ring[l - 1][up][1];
                        if constexpr (std::is_same<T, float>::value) {
                            // the FP32 form's order: j, then i -- the sum
                            // starts from a register and both lane shifts
                            // fold into their adds as DPP operands
                            a = add(a, ux);
                            c = add(c, uy);                      // j+m
                            a = add(a, ring[l - 1][RAD - m][0]);
                            c = add(c, ring[l - 1][RAD - m][1]); // j-m
                            a = add(a, xr[m - 1]);
                            c = add(c, yr[m - 1]);               // i+m
                            a = add(a, xl[m - 1]);
                            c = add(c, yl[m - 1]);               // i-m
                        } else {
                            // FP64: the reference's COO order, i then j
                            a = add(a, xr[m - 1]);
                            c = add(c, yr[m - 1]);               // i+m
                            a = add(a, xl[m - 1]);
                            c = add(c, yl[m - 1]);               // i-m
                            a = add(a, ux);
                            c = add(c, uy);                      // j+m
                            a = add(a, ring[l - 1][RAD - m][0]);
                            c = add(c, ring[l - 1][RAD - m][1]); // j-m
                        }
                    }
                    Own on;
                    T ex, ey;
                    Tr::level(x, y, a, c, fifo[l - 1][0], s, sn, vc, on, ex, ey);
                    if (l == D) {
                        if (it >= 0 && colout) Tr::store(out, row(j), x, y, on);
                    } else {
                        nx[0] = ex;
                        nx[1] = ey;
                        no = on;
                    }
                }
            }
            // shift level l by one row: the value handed up enters at the top
#pragma unroll
            for (int q = 0; q < 2; ++q) {
#pragma unroll
                for (int p = 0; p + 1 < 2 * RAD; ++p) ring[l - 1][p][q] = ring[l - 1][p + 1][q];
                ring[l - 1][2 * RAD - 1][q] = fx[q];
                fx[q] = nx[q];
            }
#pragma unroll
            for (int p = 0; p + 1 < RAD; ++p) fifo[l - 1][p] = fifo[l - 1][p + 1];
            fifo[l - 1][RAD - 1] = fo;
            fo = no;
        }
    };
    // blocks of P rows, each unrolled, so that the register rings rotate
    // partly by renaming; then the remaining rows
    constexpr int P = strip_unroll<Tr, D, RAD>();
    int it = -2 * H;
    for (; it + P <= CJ; it += P) {
#pragma unroll
        for (int k = 0; k < P; ++k) row_step(it + k);
    }
    for (; it < CJ; ++it) row_step(it);
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
template <class Tr, int D, int RAD, bool PF>
__global__ void __launch_bounds__(64 * kWaves2D)
march2d_kernel(const BatchMember *__restrict__ mem, int n_members, uint32_t n_items,
               LevelScalars<typename Tr::Sc, D> sc)
{
    // the wave index is uniform: held in a scalar register, so the item's
    // member lookup, sides and row offsets are scalar work
    const uint32_t wave = blockIdx.x * kWaves2D + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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
        march2d_strip<Tr, D, RAD, PF>(m.N, m.CI, ti, ch, strip_in<Tr>(m.in),
                                      strip_out<Tr>(m.out), sc);
    }
}

template <class Tr, int D, int RAD>
int launch_march2d_pass(Context *c, int src, const StepScalars *s, bool pf)
{
    LevelScalars<typename Tr::Sc, D> sc;
    for (int l = 0; l < D; ++l) {
        if constexpr (std::is_same<typename Tr::Sc, StepScalarsF>::value)
            sc.s[l] = to_float(s[l]);
        else
            sc.s[l] = s[l];
    }
    auto kern = pf ? march2d_kernel<Tr, D, RAD, true> : march2d_kernel<Tr, D, RAD, false>;
    int per_cu = 0;
    AXS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * kWaves2D, 0));
    const uint32_t cap = (uint32_t)c->num_cu * (uint32_t)(per_cu > 0 ? per_cu : 1);
    const Context::MultiTables &mt = c->mtab[D];
    const uint32_t need = (mt.items + kWaves2D - 1) / kWaves2D;
    const uint32_t grid = need < cap ? need : cap;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kWaves2D), 0, c->stream, mt.tab[src],
                       (int)c->members.size(), mt.items, sc);
    return AXS_OK;
}

}  // namespace axs
