// axs_march2.hip -- 3D step with temporal blocking: D = 2, 3 or 4 time steps
// per pass over HBM (standard 7-point stencil), in the FP32 form and in the
// bit-exact FP64 form (field-rescaled or Hamiltonian).
//
// A work-group of W waves covers ER = W*R rows (j) of 64 consecutive k and
// marches along i over a chunk of CI planes.  Level 0 is the state loaded
// from HBM; level l (1..D) is the lattice after l steps, computed on the
// loaded region minus an l-site rim, so the work-group writes an
// (ER-2D) x (64-2D) output tile and reads a D-site halo on every side
// (re-read by the neighbouring tiles).  Per plane i the pipeline is
//
//   load the state of plane i+D and drift it              (level 0 -> x_1)
//   for l = 1..D: at plane q = i+D-l, one step of the form (axs_forms.hpp)
//      from x_l and the site's own data, giving x_{l+1} and new own data
//   write the state of plane i
//
// Each level keeps x_l of planes q-1 and q and its own data of plane q in
// registers (the i neighbours come from registers; plane q+1 is the value
// the level below has just produced) and x_l of plane q in one LDS plane
// (the j and k neighbours), written after a barrier that ends the plane's
// reads and read after a second barrier.  Every value is computed by the
// expressions of the one-step kernels (Laplacian order centre, +-i, +-j,
// +-k; no contraction), so a pass is bit-identical to D one-step launches.
// Tiles at the lattice edge overhang it: loads wrap periodically and stores
// are masked.
#include <type_traits>

#include "axs_forms.hpp"
#include "axs_internal.hpp"
#include "axs_lattice.hpp"

namespace axs {

constexpr int kTBW = 64;   // columns loaded per row (one wave)

struct TBGeom {
    int N, CI, nTj, nTk, remap;
};

template <class T, int D, int W, int R>
using TBLds = T[D][2][W * R * kTBW];   // x_l of the plane level l works on

// One work-group: tile (tj, tk) of the lattice of side N, planes
// [ch*CI, ch*CI + CI).
template <class Tr, int D, int W, int R, bool PF>
__device__ __forceinline__ void marchd_tile(const int N, const int CI, const int tj, const int tk,
                                            const int ch, const typename Tr::In &in,
                                            const typename Tr::Out &out,
                                            const LevelScalars<typename Tr::Sc, D> &sc,
                                            TBLds<typename Tr::T, D, W, R> &X)
{
    using T = typename Tr::T;
    using Own = typename Tr::Own;
    using Raw = typename Tr::Raw;
    constexpr int ER = W * R;            // rows loaded
    constexpr int TJ = ER - 2 * D;       // rows written
    constexpr int TK = kTBW - 2 * D;     // columns written

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t NN = (uint32_t)N * (uint32_t)N;
    const int j0 = tj * TJ, k0 = tk * TK, i0 = ch * CI;

    // column of this lane, wrapped: -D <= k0 - D + lane < 2N for N >= 64
    int kk = k0 - D + lane;
    kk = (kk < 0) ? kk + N : ((kk >= N) ? kk - N : kk);
    const bool colout = lane >= D && lane < kTBW - D && (k0 + lane - D) < N;

    uint32_t rowoff[R];
    int erow[R];
    bool rowout[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = wave * R + r;                 // row within the loaded region
        int j = j0 - D + e;                         // -D <= j < 2N (ER <= 64 <= N)
        j = (j < 0) ? j + N : ((j >= N) ? j - N : j);
        rowoff[r] = (uint32_t)j * (uint32_t)N + (uint32_t)kk;
        erow[r] = e;
        rowout[r] = e >= D && e < ER - D && (j0 + e - D) < N;
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
    // site (row e, this lane) lies in level l's region
    auto inside = [&](int e, int l) -> bool {
        return e >= l && e < ER - l && lane >= l && lane < kTBW - l;
    };

    // per level l = 1..D (index l-1): x_l at planes q-1 (xm) and q (x0),
    // the own data at plane q (oq)
    T xm[D][R][2], x0[D][R][2];
    Own oq[D][R];
#pragma unroll
    for (int l = 0; l < D; ++l)
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int q = 0; q < 2; ++q) xm[l][r][q] = x0[l][r][q] = (T)0;
            oq[l][r] = Own{};
        }
    Raw raw[R], pf[R];

    if (PF) load(i0 - D, pf);
    for (int it = -2 * D; it < CI; ++it) {
        const int i = i0 + it;
        // level 0: plane i+D
        if (PF) {
#pragma unroll
            for (int r = 0; r < R; ++r) raw[r] = pf[r];
            if (it + 1 < CI) load(i + D + 1, pf);
        } else {
            load(i + D, raw);
        }
        // values handed up the levels: x_l and the own data at plane q+1
        T fx[R][2];
        Own fo[R];
        T wx[D][R][2];                     // x_l of plane q+1, for the LDS writes
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Tr::drift0(raw[r], sc.s[0], fx[r][0], fx[r][1], fo[r]);
            wx[0][r][0] = fx[r][0];
            wx[0][r][1] = fx[r][1];
        }
#pragma unroll
        for (int l = 1; l <= D; ++l) {
            const typename Tr::Sc &s = sc.s[l - 1];
            const typename Tr::Sc &sn = sc.s[l < D ? l : D - 1];
            const T *L1 = X[l - 1][0], *L2 = X[l - 1][1];
            const bool active = it >= 2 * l - 2 * D;   // x_l of planes q-1, q, q+1 exist
            T nx[R][2];
            Own no[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                nx[r][0] = nx[r][1] = (T)0;
                no[r] = Own{};
            }
            if (active) {
                const uint32_t po = plane(i + D - l);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (!inside(erow[r], l)) continue;
                    const int o = erow[r] * kTBW + lane;
                    const T x = x0[l - 1][r][0], y = x0[l - 1][r][1];
                    T a = s.c[0] * x, c = s.c[0] * y;
                    a = a + fx[r][0];
                    c = c + fx[r][1];        // i+1
                    a = a + xm[l - 1][r][0];
                    c = c + xm[l - 1][r][1]; // i-1
                    a = a + L1[o + kTBW];
                    c = c + L2[o + kTBW];    // j+1
                    a = a + L1[o - kTBW];
                    c = c + L2[o - kTBW];    // j-1
                    a = a + L1[o + 1];
                    c = c + L2[o + 1];       // k+1
                    a = a + L1[o - 1];
                    c = c + L2[o - 1];       // k-1
                    Own on;
                    T ex, ey;
                    Tr::level(x, y, a, c, oq[l - 1][r], s, sn, on, ex, ey);
                    if (l == D) {
                        if (it >= 0 && rowout[r] && colout)
                            Tr::store(out, po + rowoff[r], x, y, on);
                    } else {
                        nx[r][0] = ex;
                        nx[r][1] = ey;
                        no[r] = on;
                    }
                }
            }
            // rotate level l: q-1 <- q <- q+1 (the value handed up)
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    xm[l - 1][r][q] = x0[l - 1][r][q];
                    x0[l - 1][r][q] = fx[r][q];
                }
                oq[l - 1][r] = fo[r];
            }
            if (l < D) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        fx[r][q] = nx[r][q];
                        wx[l][r][q] = nx[r][q];
                    }
                    fo[r] = no[r];
                }
            }
        }
        __syncthreads();   // every read of this plane's LDS values is done
#pragma unroll
        for (int l = 0; l < D; ++l) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int o = erow[r] * kTBW + lane;
                X[l][0][o] = wx[l][r][0];
                X[l][1][o] = wx[l][r][1];
            }
        }
        __syncthreads();   // the next plane's values are visible
    }
}

template <class Tr, int D, int W, int R, bool PF>
__global__ void __launch_bounds__(64 * W)
marchd_kernel(TBGeom g, typename Tr::In in, typename Tr::Out out,
              LevelScalars<typename Tr::Sc, D> sc)
{
    __shared__ TBLds<typename Tr::T, D, W, R> X;
    uint32_t b = blockIdx.x;
    if (g.remap) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);   // XCD-aware order
    const int tk = (int)(b % (uint32_t)g.nTk);
    b /= (uint32_t)g.nTk;
    const int tj = (int)(b % (uint32_t)g.nTj);
    const int ch = (int)(b / (uint32_t)g.nTj);
    marchd_tile<Tr, D, W, R, PF>(g.N, g.CI, tj, tk, ch, in, out, sc, X);
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
template <class Tr, int D, int W, int R, bool PF>
__global__ void __launch_bounds__(64 * W)
marchd_batch_kernel(const BatchMember *__restrict__ mem, int n_members, uint32_t n_items,
                    LevelScalars<typename Tr::Sc, D> sc)
{
    __shared__ TBLds<typename Tr::T, D, W, R> X;
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
        marchd_tile<Tr, D, W, R, PF>(m.N, m.CI, tj, tk, ch, form_in<Tr>(m.in), form_out<Tr>(m.out),
                                     sc, X);
    }
}

bool march2_supported(const Context *c, int N)
{
    return c->nd == 3 && c->p.stencil_setting == 0 &&
           c->p.laplacian == AXS_LAPLACIAN_STENCIL && N >= 64;
}

// shapes built per form: FP32, and FP64 (double pairs: 8 waves, 2 or 3
// rows, depth 2 or 3 -- with 16 waves the 128 registers per lane spill)
bool multi_shape_built(int D, int W, int R, int prec)
{
    if (prec == AXS_FP64) return W == 8 && (R == 2 || R == 3) && (D == 2 || D == 3);
    switch (W) {
    case 4: return (D == 2 && (R == 4 || R == 6 || R == 8)) || (D == 3 && (R == 4 || R == 6));
    case 8: return D >= 2 && D <= 4 && R >= 2 && R <= 4;
    case 16: return D >= 2 && D <= 4 && (R == 2 || R == 3);
    default: return false;
    }
}

int multi_waves(int flags)
{
    return (flags & kTuneWG1024) ? 16 : ((flags & kTuneWG512) ? 8 : 4);
}

// rows per wave: the configured count, else the default for the shape
static int pick_rows(const Context *c, int N, int D, int W)
{
    const int r = c->tune.rtb;
    if (r && multi_shape_built(D, W, r, c->prec)) return r;
    if (W == 4) return (D == 2 && N >= 512) ? 6 : 4;
    if (c->prec == AXS_FP64) return 2;         // measured best at 1024^3
    return (W == 8) ? 3 : 2;
}

// planes per chunk: the configured count if it divides N, else long enough
// to amortise the 2D halo planes and short enough to give the chip work
// (measured best: 64 at 512^3 and 1024^3, 32 at 256^3, 8 at 128^3)
static int pick_chunk(const Context *c, int N)
{
    int ci = c->tune.ci32;
    if (ci > 0 && N % ci == 0) return ci;
    ci = (N >= 512) ? 64 : ((N >= 256) ? 32 : 8);
    while (N % ci != 0) --ci;
    return ci;
}

// calls f.run<W, R>() for the configured shape of form Tr
template <int D, class Tr, class F>
static int dispatch_shape(int W, int R, F &f)
{
    if constexpr (std::is_same<Tr, F32Form>::value) {
        if (W == 16) {
            if (R == 3) return f.template run<16, 3>();
            return f.template run<16, 2>();
        }
        if (W == 8) {
            if (R == 2) return f.template run<8, 2>();
            if (R == 3) return f.template run<8, 3>();
            return f.template run<8, 4>();
        }
        if constexpr (D == 2) {
            if (R == 8) return f.template run<4, 8>();
            if (R == 4) return f.template run<4, 4>();
            return f.template run<4, 6>();
        } else if constexpr (D == 3) {
            if (R == 4) return f.template run<4, 4>();
            return f.template run<4, 6>();
        }
    } else if constexpr (D == 2 || D == 3) {
        if (W == 8) {
            if (R == 3) return f.template run<8, 3>();
            return f.template run<8, 2>();
        }
    }
    set_error("depth %d with %d waves per work-group is not built for this form", D, W);
    return AXS_ERR_UNSUPPORTED;
}

template <class Tr, int D>
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
        constexpr int TJ = W * R - 2 * D;
        constexpr int TK = kTBW - 2 * D;
        static_assert(TJ > 0 && W * R <= 64, "tile shape");
        TBGeom g;
        g.N = N;
        g.nTj = (N + TJ - 1) / TJ;
        g.nTk = (N + TK - 1) / TK;
        g.CI = pick_chunk(c, N);
        const uint32_t nb = (uint32_t)g.nTj * (uint32_t)g.nTk * (uint32_t)(N / g.CI);
        g.remap = (nb % 8 == 0) ? 1 : 0;
        if (pf)
            hipLaunchKernelGGL((marchd_kernel<Tr, D, W, R, true>), dim3(nb), dim3(64 * W), 0,
                               c->stream, g, in, out, sc);
        else
            hipLaunchKernelGGL((marchd_kernel<Tr, D, W, R, false>), dim3(nb), dim3(64 * W), 0,
                               c->stream, g, in, out, sc);
        return AXS_OK;
    }
};

template <class Tr, int D>
struct BatchLaunch {
    Context *c;
    int src;
    LevelScalars<typename Tr::Sc, D> sc;
    bool pf;
    template <int W, int R>
    int run()
    {
        auto kern = pf ? marchd_batch_kernel<Tr, D, W, R, true>
                       : marchd_batch_kernel<Tr, D, W, R, false>;
        int per_cu = 0;
        AXS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, 0));
        const uint32_t cap = (uint32_t)c->num_cu * (uint32_t)(per_cu > 0 ? per_cu : 1);
        const uint32_t grid = c->mbatch_items < cap ? c->mbatch_items : cap;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * W), 0, c->stream, c->mbatch_tab[src],
                           (int)c->members.size(), c->mbatch_items, sc);
        return AXS_OK;
    }
};

template <class Sc>
static Sc level_scalars(const StepScalars &s)
{
    if constexpr (std::is_same<Sc, StepScalarsF>::value)
        return to_float(s);
    else
        return s;
}

template <class Tr, int D>
static int launch_depth(Context *c, int src, int dst, const StepScalars *s)
{
    LevelScalars<typename Tr::Sc, D> sc;
    for (int l = 0; l < D; ++l) sc.s[l] = level_scalars<typename Tr::Sc>(s[l]);
    const bool pf = (c->tune.flags & kTunePrefetch) != 0;
    const int W = multi_waves(c->tune.flags);
    if (c->members.size() > 1 && c->mbatch_tab[src]) {
        BatchLaunch<Tr, D> f{c, src, sc, pf};
        return dispatch_shape<D, Tr>(W, pick_rows(c, c->members[0].N, D, W), f);
    }
    for (Member &m : c->members) {
        SingleLaunch<Tr, D> f{c, m.N, form_in<Tr>(m.state[src]), form_out<Tr>(m.state[dst]), sc,
                              pf};
        AXS_TRY((dispatch_shape<D, Tr>(W, pick_rows(c, m.N, D, W), f)));
    }
    return AXS_OK;
}

template <class Tr>
static int launch_form(Context *c, int src, int dst, const StepScalars *s, int depth)
{
    if (depth == 4) {
        if constexpr (std::is_same<Tr, F32Form>::value) {
            return launch_depth<Tr, 4>(c, src, dst, s);
        } else {
            set_error("FP64 depth 4 is not built");
            return AXS_ERR_UNSUPPORTED;
        }
    }
    if (depth == 3) return launch_depth<Tr, 3>(c, src, dst, s);
    return launch_depth<Tr, 2>(c, src, dst, s);
}

int multi_step_depth(const Context *c)
{
    if (c->tune.depth < 2 || c->tune.kernel == AXS_KERNEL_SIMPLE) return 1;
    for (const Member &m : c->members)
        if (!(c->nd == 2 ? march2d_supported(c, m.N) : march2_supported(c, m.N))) return 1;
    return c->tune.depth;
}

int prepare_multi_batch(Context *c)
{
    if (c->nd == 2) return prepare_march2d_items(c);
    if (c->members.size() < 2) return AXS_OK;
    const int D = multi_step_depth(c);
    if (D < 2) return AXS_OK;
    const int W = multi_waves(c->tune.flags);
    // one row count for the ensemble: that of its first member's side
    const int R = pick_rows(c, c->members[0].N, D, W);
    const int key = ((D * 100 + W) * 100 + R) * 10000 + c->tune.ci32;
    if (key == c->mbatch_key && c->mbatch_tab[0]) return AXS_OK;
    const int TJ = W * R - 2 * D, TK = kTBW - 2 * D;
    std::vector<BatchMember> tab[2];
    uint32_t items = 0;
    for (const Member &m : c->members) {
        BatchMember b{};
        b.N = m.N;
        b.L = (uint32_t)m.L;
        b.item_begin = items;
        b.nTj = (m.N + TJ - 1) / TJ;
        b.nTk = (m.N + TK - 1) / TK;
        b.CI = pick_chunk(c, m.N);
        items += (uint32_t)b.nTj * b.nTk * (m.N / b.CI);
        for (int h = 0; h < 2; ++h) {
            for (int a = 0; a < 3; ++a) {
                b.in[a] = m.state[h][a];
                b.out[a] = m.state[h ^ 1][a];
            }
            tab[h].push_back(b);
        }
    }
    for (int h = 0; h < 2; ++h) {
        if (!c->mbatch_tab[h])
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

int launch_multi_steps(Context *c, const StepScalars *s, int depth)
{
    const int src = c->cur, dst = c->cur ^ 1;
    if (c->nd == 2)
        AXS_TRY(launch_march2d(c, src, s, depth));
    else if (c->prec == AXS_FP32)
        AXS_TRY(launch_form<F32Form>(c, src, dst, s, depth));
    else if (c->p.eom == AXS_EOM_HAMILTONIAN)
        AXS_TRY(launch_form<F64FormT<true>>(c, src, dst, s, depth));
    else
        AXS_TRY(launch_form<F64FormT<false>>(c, src, dst, s, depth));
    AXS_HIP_TRY(hipGetLastError());
    c->cur = dst;
    return AXS_OK;
}

}  // namespace axs
