// axs_march2.hpp -- the 3D multi-step (temporally blocked) step kernels,
// templated on the storage form, the steps per pass D, the work-group shape
// (W waves of R rows) and the stencil radius RAD (stencil_setting + 1).
//
// A work-group of W waves covers ER = W*R rows (j) of 64 consecutive k and
// marches along i over a chunk of CI planes.  Level 0 is the state loaded
// from HBM; level l (1..D) is the lattice after l steps, computed on the
// loaded region minus an l*RAD-site rim, so the work-group writes an
// (ER - 2 D RAD) x (64 - 2 D RAD) output tile and reads a D*RAD-site halo on
// every side (re-read by the neighbouring tiles).  Per plane i:
//
//   load the state of plane i + D RAD and drift it           (level 0 -> x_1)
//   for l = 1..D: at plane q = i + (D - l) RAD, one step of the form
//      (axs_forms.hpp) from x_l and the site's own data
//   write the state of plane i
//
// Each level keeps x_l of planes q-RAD .. q+RAD-1 in registers (plane
// q+RAD is the value the level below has just produced) -- the i neighbours
// -- and the own data of planes q .. q+RAD-1 (produced RAD planes before
// they are used); x_l of plane q sits in one LDS plane per level for the j
// and k neighbours, written after a barrier that ends the plane's reads and
// read after a second barrier.  The Laplacian sums in the one-step kernels'
// order -- the form's centre term, then for each distance m: +-i, +-j,
// +-k, each times the stencil coefficient -- with the forms' expressions,
// so a pass is bit-identical to D one-step launches.  Tiles at the lattice
// edge overhang it: loads wrap periodically and stores are masked.
//
// Issue cost: a row outside a level's region is skipped with a scalar
// branch (the wave index is uniform) except in the FP32 7-point form, where
// every row computes (kSkipRows); every lane of a computed row computes
// (lanes outside the region cost no extra issue cycles and their values
// never reach a stored site), the LDS planes carry RAD padding rows so that
// those lanes' reads stay inside them, and the march loop runs in unrolled
// blocks of planes.
#pragma once

#include <type_traits>

#include "axs_forms.hpp"
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kTBW = 64;   // columns loaded per row (one wave)

// The columns of a tile: each row loads the 64 columns [k0 - HK, k0 - HK +
// 64) around the TK columns [k0, k0 + TK) it writes.  With a halo of at most
// four sites the rows keep four halo columns on each side and write 56, and
// the tiles start at k0 = 56 tk + 4: every loaded row segment then begins
// on a 128-byte line (56 sites of 16 bytes are 7 lines) and takes 8 lines,
// not 9.  Deeper halos keep HK = H, TK = 64 - 2H, k0 = TK tk.  The tiles
// cover the columns [origin, origin + N), wrapping at N.
__host__ __device__ constexpr int tile_halo_k(int H) { return H <= 4 ? 4 : H; }
__host__ __device__ constexpr int tile_cols(int H) { return kTBW - 2 * tile_halo_k(H); }
__host__ __device__ constexpr int tile_origin(int H) { return H <= 4 ? 4 : 0; }

// FP32 7-point stencil: the k neighbours of a site are x of the lanes next
// to it, which a DPP whole-wave shift hands to the add as its operand
// (folded into the v_add: no LDS read, no extra instruction).  The end
// lanes read zero, and lie outside every level's region.
template <class T, int RAD>
constexpr bool kLaneK = std::is_same<T, float>::value && RAD == 1;
// Rows outside a level's region: skipped with a scalar branch, except in
// that form, where every row computes -- the branch saves a sixth of the
// arithmetic at depth 4, but merging the values of its two paths costs more
// register moves than that, and the stores after all rows (below) keep the
// DPP operands folded.  The wider stencils keep the branch: computing every
// row there needs more registers than a 16-wave work-group has.
template <class T, int RAD>
constexpr bool kSkipRows = !kLaneK<T, RAD>;
// planes per unrolled block of the march loop: the register rings then
// rotate partly by renaming (measured: three planes 8% faster than one at
// 1024^3, FP32 depth 4; two and six no better).  Shapes whose rings and
// FIFOs hold many registers per lane (weighted double for 16-wave groups,
// which have half the registers) stay at one plane: unrolled, they spill.
template <class Tr, int D, int W, int R, int RAD>
constexpr int plane_unroll()
{
    constexpr int words = D * R * RAD * (4 * (int)sizeof(typename Tr::T) +
                                         (int)sizeof(typename Tr::Own)) / 4;
    return words * (W >= 16 ? 2 : 1) <= 96 ? 3 : 1;
}

struct TBGeom {
    int N, CI, nTj, nTk;
    int band;     // tile order: bands of `band` tile rows, j fastest in a band
    int xcd;      // 1: each XCD takes a contiguous range of the order (xcd_order)
};

// The tile order of a launch, chosen by timing at 1024^3 (whose launches
// hold 13 work-groups per CU in the FP32 form, 116 in the FP64 form): FP32
// walks whole columns of tiles, j fastest, in dispatch order (work-groups
// dealt to the XCDs in turn); FP64 gives each XCD a contiguous range
// (xcd_order) of bands of 8 tile rows, j fastest in a band.  Timings of one
// order differ by up to 10% between processes (the state's placement in
// memory), so the choice rests on comparisons within one process.
template <class Tr>
inline void tile_order(TBGeom &g)
{
    const bool f32 = std::is_same<typename Tr::T, float>::value;
    g.band = f32 ? g.nTj : 8;
    g.xcd = f32 ? 0 : 1;
}

// tile (tj, tk) of logical index t of a layer of nTj x nTk tiles (band >= 1;
// band 1 is k fastest)
__device__ __forceinline__ void tile_of(uint32_t t, const TBGeom &g, int &tj, int &tk)
{
    const uint32_t nTk = (uint32_t)g.nTk, nTj = (uint32_t)g.nTj;
    // band of `band` tile rows (the last one shorter), walked j fastest:
    // consecutive work-groups form a band x (count / band) block of tiles
    const uint32_t bw = (uint32_t)g.band * nTk;
    const uint32_t j0 = (t / bw) * (uint32_t)g.band, loc = t % bw;
    const uint32_t h = (nTj - j0 < (uint32_t)g.band) ? nTj - j0 : (uint32_t)g.band;
    tj = (int)(j0 + loc % h);
    tk = (int)(loc / h);
}

// x_l of the plane level l works on: the ER loaded rows plus RAD padding
// rows above and below, so that the neighbour reads of the rim rows and
// lanes stay inside the array (their values are never used)
template <class T, int D, int W, int R, int RAD>
using TBLds = T[D][2][(W * R + 2 * RAD) * kTBW];

// One work-group: tile (tj, tk) of the lattice of side N, planes
// [ch*CI, ch*CI + CI).
template <class Tr, int D, int W, int R, int RAD, bool PF>
__device__ __forceinline__ void marchd_tile(const int N, const int CI, const int tj, const int tk,
                                            const int ch, const typename Tr::In &in,
                                            const typename Tr::Out &out,
                                            const LevelScalars<typename Tr::Sc, D> &sc,
                                            TBLds<typename Tr::T, D, W, R, RAD> &X)
{
    using T = typename Tr::T;
    using Own = typename Tr::Own;
    using Raw = typename Tr::Raw;
    constexpr int ER = W * R;            // rows loaded
    constexpr int H = D * RAD;           // halo width
    constexpr int TJ = ER - 2 * H;       // rows written
    constexpr int HK = tile_halo_k(H);   // halo columns on each side
    constexpr int TK = tile_cols(H);     // columns written
    constexpr int K0 = tile_origin(H);
    static_assert(TJ > 0 && TK > 0, "tile shape");

    // the wave index is uniform: held in a scalar register, so the row
    // tests below are scalar branches
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t NN = (uint32_t)N * (uint32_t)N;
    const int j0 = tj * TJ, k0 = tk * TK + K0, i0 = ch * CI;

    // column of this lane, wrapped: -HK <= k0 - HK + lane < 2N for N >= 64;
    // it is written if it lies in the lanes' output range and in the first
    // pass of the columns from the tiles' origin.  In the last tile of a row
    // of tiles, the lanes past the halo of its last written column load the
    // column of the last lane still needed: the same lines, so no fetch
    // (their values, like every lane's outside the levels' regions, reach
    // no stored site: a wrong value spreads one stencil radius per level)
    const int wcols = (N + K0 - k0 < TK) ? N + K0 - k0 : TK;       // columns written
    const int lane_ld = (lane < wcols + 2 * HK) ? lane : wcols + 2 * HK - 1;
    int kk = k0 - HK + lane_ld;
    kk = (kk < 0) ? kk + N : ((kk >= N) ? kk - N : kk);
    const bool colout = lane >= HK && lane < kTBW - HK && (k0 + lane - HK) < N + K0;

    uint32_t rowoff[R];
    int erow[R];
    bool rowout[R];
    // rows likewise: past the halo of the last written row, a copy of the
    // last row still needed (the same lines, fetched once)
    const int wrows = (N - j0 < TJ) ? N - j0 : TJ;                   // rows written
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = wave * R + r;                 // row within the loaded region
        const int e_ld = (e < wrows + 2 * H) ? e : wrows + 2 * H - 1;
        int j = j0 - H + e_ld;                      // -H <= j < 2N (ER <= 64 <= N)
        j = (j < 0) ? j + N : ((j >= N) ? j - N : j);
        rowoff[r] = (uint32_t)j * (uint32_t)N + (uint32_t)kk;
        erow[r] = e;
        rowout[r] = e >= H && e < ER - H && (j0 + e - H) < N;
    }
    auto plane = [&](int i) -> uint32_t {
        int w = (i < 0) ? i + N : ((i >= N) ? i - N : i);
        return (uint32_t)w * NN;
    };
    auto load = [&](int i, Raw (&raw)[R]) {
        const uint32_t po = plane(i);
#pragma unroll
        for (int r = 0; r < R; ++r) Tr::fetch(in, po + rowoff[r], raw[r]);
    };
    // row e lies in level l's region.  Every lane of such a row computes:
    // masking the rim lanes would save no issue cycles, and the values of
    // lanes (and rows, planes) outside level l's region only ever reach
    // sites outside level l+1's region, which are never stored
    auto row_in = [&](int e, int l) -> bool { return e >= l * RAD && e < ER - l * RAD; };

    // per level l = 1..D (index l-1): x_l of planes q-RAD .. q+RAD-1 (ring)
    // and the own data of planes q .. q+RAD-1 (fifo)
    T ring[D][2 * RAD][R][2];
    Own fifo[D][RAD][R];
#pragma unroll
    for (int l = 0; l < D; ++l)
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int p = 0; p < 2 * RAD; ++p) ring[l][p][r][0] = ring[l][p][r][1] = (T)0;
#pragma unroll
            for (int p = 0; p < RAD; ++p) fifo[l][p][r] = Own{};
        }
    Raw raw[R], pf[R];
    const typename Tr::Vc vc = Tr::vconst(sc.s[0]);

    if (PF) load(i0 - H, pf);
    auto plane_step = [&](const int it) __attribute__((always_inline)) {
        const int i = i0 + it;
        // level 0: plane i+H
        if (PF) {
#pragma unroll
            for (int r = 0; r < R; ++r) raw[r] = pf[r];
            if (it + 1 < CI) load(i + H + 1, pf);
        } else {
            load(i + H, raw);
        }
        // values handed up the levels: x_l and the own data of plane q+RAD
        T fx[R][2];
        Own fo[R];
#pragma unroll
        for (int r = 0; r < R; ++r) Tr::drift0(raw[r], sc.s[0], fx[r][0], fx[r][1], fo[r]);
#pragma unroll
        for (int l = 1; l <= D; ++l) {
            const typename Tr::Sc &s = sc.s[l - 1];
            const typename Tr::Sc &sn = sc.s[l < D ? l : D - 1];
            const T *L1 = X[l - 1][0], *L2 = X[l - 1][1];
            // level l computes from the first plane on: until x_l of planes
            // q-RAD .. q+RAD exist (it >= 2 RAD (l - D)) its results are
            // unused -- they leave the ring of level l+1 before it fills
            T nx[R][2];
            Own no[R];
            {
                const uint32_t po = plane(i + (D - l) * RAD);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (kSkipRows<T, RAD>) {
                        // rows outside the region: any value (never used)
                        nx[r][0] = __builtin_nondeterministic_value((T)0);
                        nx[r][1] = __builtin_nondeterministic_value((T)0);
                        no[r] = Tr::unset();
                        if (!row_in(erow[r], l)) continue;
                    }
                    const int o = (erow[r] + RAD) * kTBW + lane;
                    const T x = ring[l - 1][RAD][r][0], y = ring[l - 1][RAD][r][1];
                    T a = Tr::centre(s, x), c = Tr::centre(s, y);
#pragma unroll
                    for (int m = 1; m <= RAD; ++m) {
                        // the 7-point stencil's coefficient is 1: its
                        // products are exact and left out
                        const T cm = (T)s.c[m];
                        auto add = [&](T acc, T v) -> T {
                            return (RAD == 1) ? acc + v : Tr::acc(acc, cm, v);
                        };
                        const int up = (RAD + m < 2 * RAD) ? RAD + m : 2 * RAD - 1;
                        const T ux = (m == RAD) ? fx[r][0] : ring[l - 1][up][r][0];
                        const T uy = (m == RAD) ? fx[r][1] : ring[l - 1][up][r][1];
                        a = add(a, ux);
                        c = add(c, uy);                                     // i+m
                        a = add(a, ring[l - 1][RAD - m][r][0]);
                        c = add(c, ring[l - 1][RAD - m][r][1]);             // i-m
                        a = add(a, L1[o + m * kTBW]);
                        c = add(c, L2[o + m * kTBW]);                       // j+m
                        a = add(a, L1[o - m * kTBW]);
                        c = add(c, L2[o - m * kTBW]);                       // j-m
                        if constexpr (kLaneK<T, RAD>) {
                            // k+-1 are the neighbour lanes' x: DPP operands
                            a = add(a, lane_next(x));
                            c = add(c, lane_next(y));                       // k+1
                            a = add(a, lane_prev(x));
                            c = add(c, lane_prev(y));                       // k-1
                        } else {
                            a = add(a, L1[o + m]);
                            c = add(c, L2[o + m]);                          // k+m
                            a = add(a, L1[o - m]);
                            c = add(c, L2[o - m]);                          // k-m
                        }
                    }
                    Own on;
                    T ex, ey;
                    Tr::level(x, y, a, c, fifo[l - 1][0][r], s, sn, vc, on, ex, ey);
                    if (l == D && !kLaneK<T, RAD>) {
                        if (it >= 0 && rowout[r] && colout)
                            Tr::store(out, po + rowoff[r], x, y, on);
                        continue;
                    }
                    // kLaneK: level D's stored site too, stored after every
                    // row's arithmetic, so that the stores' exec-masked
                    // branches split no row's instructions
                    nx[r][0] = (l == D) ? x : ex;
                    nx[r][1] = (l == D) ? y : ey;
                    no[r] = on;
                }
                if constexpr (kLaneK<T, RAD>) {
                    if (l == D) {
                        // the stored values are computed here, for every
                        // lane: sunk into the stores' masked branches, the
                        // adds would lose their DPP operands to separate moves
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            __asm__ volatile("" : "+v"(no[r].u1), "+v"(no[r].u2));
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            if (it >= 0 && rowout[r] && colout)
                                Tr::store(out, po + rowoff[r], nx[r][0], nx[r][1], no[r]);
                    }
                }
            }
            // shift level l by one plane: the value handed up enters at the top
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma unroll
                for (int p = 0; p + 1 < 2 * RAD; ++p) {
                    ring[l - 1][p][r][0] = ring[l - 1][p + 1][r][0];
                    ring[l - 1][p][r][1] = ring[l - 1][p + 1][r][1];
                }
                ring[l - 1][2 * RAD - 1][r][0] = fx[r][0];
                ring[l - 1][2 * RAD - 1][r][1] = fx[r][1];
#pragma unroll
                for (int p = 0; p + 1 < RAD; ++p) fifo[l - 1][p][r] = fifo[l - 1][p + 1][r];
                fifo[l - 1][RAD - 1][r] = fo[r];
            }
            if (l < D) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    fx[r][0] = nx[r][0];
                    fx[r][1] = nx[r][1];
                    fo[r] = no[r];
                }
            }
        }
        __syncthreads();   // every read of this plane's LDS values is done
        // each level's next plane (q+1) is now the middle of its ring
#pragma unroll
        for (int l = 0; l < D; ++l) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int o = (erow[r] + RAD) * kTBW + lane;
                X[l][0][o] = ring[l][RAD][r][0];
                X[l][1][o] = ring[l][RAD][r][1];
            }
        }
        __syncthreads();   // the next plane's values are visible
    };
    // blocks of P planes, each unrolled, so that the register rings rotate
    // by renaming instead of moves; then the remaining planes
    constexpr int P = plane_unroll<Tr, D, W, R, RAD>();
    int it = -2 * H;
    for (; it + P <= CI; it += P) {
#pragma unroll
        for (int k = 0; k < P; ++k) plane_step(it + k);
    }
    for (; it < CI; ++it) plane_step(it);
}

template <class Tr, int D, int W, int R, int RAD, bool PF>
__global__ void __launch_bounds__(64 * W)
marchd_kernel(TBGeom g, typename Tr::In in, typename Tr::Out out,
              LevelScalars<typename Tr::Sc, D> sc)
{
    __shared__ TBLds<typename Tr::T, D, W, R, RAD> X;
    const uint32_t b = g.xcd ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t layer = (uint32_t)g.nTj * (uint32_t)g.nTk;
    int tj, tk;
    tile_of(b % layer, g, tj, tk);
    const int ch = (int)(b / layer);
    marchd_tile<Tr, D, W, R, RAD, PF>(g.N, g.CI, tj, tk, ch, in, out, sc, X);
}

// the form's array views from a member-table entry (device) or a member (host)
template <class Tr>
__host__ __device__ __forceinline__ typename Tr::In form_in(void *const *p)
{
    if constexpr (std::is_same<Tr, F32Form>::value)
        return {(const fvec4 *)p[PHI]};
    else
        return {(const dvec2 *)p[PHI], (const dvec2 *)p[VEL], (const dvec2 *)p[KER]};
}
template <class Tr>
__host__ __device__ __forceinline__ typename Tr::Out form_out(void *const *p)
{
    if constexpr (std::is_same<Tr, F32Form>::value)
        return {(fvec4 *)p[PHI]};
    else
        return {(dvec2 *)p[PHI], (dvec2 *)p[VEL], (dvec2 *)p[KER]};
}

// Ensemble form: one persistent launch for every member.  Work items are
// the (member, tile, chunk) triples, members in order, tiles k fastest;
// work-group b takes items b, b + gridDim.x, ...  The first barrier of a
// tile orders its LDS writes after the previous tile's reads.
template <class Tr, int D, int W, int R, int RAD, bool PF>
__global__ void __launch_bounds__(64 * W)
marchd_batch_kernel(const BatchMember *__restrict__ mem, int n_members, uint32_t n_items,
                    LevelScalars<typename Tr::Sc, D> sc)
{
    __shared__ TBLds<typename Tr::T, D, W, R, RAD> X;
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        int lo = 0, hi = n_members - 1;        // last member starting at or before item
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (mem[mid].item_begin <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchMember &m = mem[lo];
        uint32_t t = item - m.item_begin;
        const int tk = (int)(t % (uint32_t)m.nTk);
        t /= (uint32_t)m.nTk;
        const int tj = (int)(t % (uint32_t)m.nTj);
        const int ch = (int)(t / (uint32_t)m.nTj);
        marchd_tile<Tr, D, W, R, RAD, PF>(m.N, m.CI, tj, tk, ch, form_in<Tr>(m.in),
                                          form_out<Tr>(m.out), sc, X);
    }
}

// planes per chunk of a member of side N, tiles TJ x TK: the configured
// count if it divides N, else the ensemble-wide choice of multi_chunk_cap
inline int multi_chunk(const Context *c, int N, int TJ, int TK, int H)
{
    const int ci = c->tune.ci32;
    if (ci > 0 && N % ci == 0) return ci;
    return divisor_below(N, multi_chunk_cap(c, TJ, TK, H));
}

template <class Sc>
inline Sc level_scalars(const StepScalars &s)
{
    if constexpr (std::is_same<Sc, StepScalarsF>::value)
        return to_float(s);
    else
        return s;
}

template <class Tr, int D, int RAD>
struct SingleLaunch {
    Context *c;
    int N;
    typename Tr::In in;
    typename Tr::Out out;
    LevelScalars<typename Tr::Sc, D> sc;
    bool pf;
    template <int W, int R>
    int run()
    {
        constexpr int TJ = W * R - 2 * D * RAD;
        constexpr int TK = tile_cols(D * RAD);
        static_assert(TJ > 0 && W * R <= 64, "tile shape");
        TBGeom g;
        g.N = N;
        g.nTj = (N + TJ - 1) / TJ;
        g.nTk = (N + TK - 1) / TK;
        g.CI = multi_chunk(c, N, TJ, TK, D * RAD);
        const uint32_t nb = (uint32_t)g.nTj * (uint32_t)g.nTk * (uint32_t)(N / g.CI);
        tile_order<Tr>(g);
        if (pf)
            hipLaunchKernelGGL((marchd_kernel<Tr, D, W, R, RAD, true>), dim3(nb), dim3(64 * W), 0,
                               c->stream, g, in, out, sc);
        else
            hipLaunchKernelGGL((marchd_kernel<Tr, D, W, R, RAD, false>), dim3(nb), dim3(64 * W),
                               0, c->stream, g, in, out, sc);
        return AXS_OK;
    }
};

template <class Tr, int D, int RAD>
struct BatchLaunch {
    Context *c;
    int src;
    LevelScalars<typename Tr::Sc, D> sc;
    bool pf;
    template <int W, int R>
    int run()
    {
        auto kern = pf ? marchd_batch_kernel<Tr, D, W, R, RAD, true>
                       : marchd_batch_kernel<Tr, D, W, R, RAD, false>;
        int per_cu = 0;
        AXS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, 0));
        const uint32_t cap = (uint32_t)c->num_cu * (uint32_t)(per_cu > 0 ? per_cu : 1);
        const Context::MultiTables &mt = c->mtab[D];
        const uint32_t grid = mt.items < cap ? mt.items : cap;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * W), 0, c->stream, mt.tab[src],
                           (int)c->members.size(), mt.items, sc);
        return AXS_OK;
    }
};

// one multi-step pass of every member with work-group shape (W, R):
// Dispatch(W, R, f) calls f.run<W, R>() for the built shapes
template <class Tr, int D, int RAD, class Dispatch>
int launch_members(Context *c, int src, int dst, const StepScalars *s, int W, int Rbatch,
                   int (*rows_of)(const Context *, int, int, int), Dispatch dispatch)
{
    LevelScalars<typename Tr::Sc, D> sc;
    for (int l = 0; l < D; ++l) sc.s[l] = level_scalars<typename Tr::Sc>(s[l]);
    const bool pf = (c->tune.flags & kTunePrefetch) != 0;
    if (c->members.size() > 1 && c->mtab[D].tab[src]) {
        BatchLaunch<Tr, D, RAD> f{c, src, sc, pf};
        return dispatch(W, Rbatch, f);
    }
    for (Member &m : c->members) {
        SingleLaunch<Tr, D, RAD> f{c, m.N, form_in<Tr>(m.state[src]), form_out<Tr>(m.state[dst]),
                                   sc, pf};
        AXS_TRY(dispatch(W, rows_of(c, m.N, D, W), f));
    }
    return AXS_OK;
}

}  // namespace axs
