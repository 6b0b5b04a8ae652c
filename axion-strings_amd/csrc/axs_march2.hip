// axs_march2.hip -- the 3D multi-step kernels (axs_march2.hpp) for the
// 7-point stencil, and the context-level entry points of the multi-step
// path: eligibility, work-item tables for ensembles, launches.  Radii 2-4
// (stencil_setting 1-3) are instantiated in axs_march2_hi.hip.
#include "axs_march2.hpp"

namespace axs {

bool march2_supported(const Context *c, int N)
{
    return c->nd == 3 && c->p.laplacian == AXS_LAPLACIAN_STENCIL && N >= 64;
}

// shapes built per form and stencil radius (none spills registers).
// Radius 1: FP32 depth 2-4 with 4, 8 or 16 waves, depth 5 with 16 waves of
// 2 rows (115 registers; deeper or taller shapes spill); FP64 (double pairs) 8
// waves, 2 or 3 rows, depth 2 or 3 (16 waves would spill: 128 registers per
// lane).  Radii 2-4: FP32 16 waves of 2 rows, depth 1 or 2 (radius 4: 1);
// FP64 8 waves of 2 rows (radius 2: or 3), depth 1.
bool multi_shape_built(int D, int W, int R, int prec, int RAD)
{
    if (RAD > 1) {
        if (RAD > 4) return false;
        if (prec == AXS_FP64) return D == 1 && W == 8 && (R == 2 || (R == 3 && RAD == 2));
        return W == 16 && R == 2 && D >= 1 && D <= (RAD == 4 ? 1 : 2);
    }
    if (prec == AXS_FP64) return W == 8 && (R == 2 || R == 3) && (D == 2 || D == 3);
    switch (W) {
    case 4: return (D == 2 && (R == 4 || R == 6 || R == 8)) || (D == 3 && (R == 4 || R == 6));
    case 8: return D >= 2 && D <= 4 && R >= 2 && R <= 4;
    case 16: return (D >= 2 && D <= 4 && (R == 2 || R == 3)) || (D == 5 && R == 2);
    default: return false;
    }
}

int multi_waves(int flags)
{
    return (flags & kTuneWG1024) ? 16 : ((flags & kTuneWG512) ? 8 : 4);
}

int stencil_radius_of(const Context *c) { return c->p.stencil_setting + 1; }

// Chunk of planes per work item (FP32 up to 256, FP64 up to 64, at most
// half the smallest side) by a cost model: one work-group per CU, every item
// marches chunk + 2H planes (H = halo), so a pass takes about
// ceil(items / CUs) * (chunk + 2H) plane steps of one work-group; the
// cheapest cap wins (ties: the longer chunk).  It reproduces the measured
// best chunk at every side swept (FP32 depth 4: 256 at 1024^3, 64 at 384^3
// and 256^3, 32 at 192^3, 8 at 128^3 (depth 3); FP64 depth 3: 64 at 1024^3,
// 320^3 and 192^3, 32 at 256^3 and 128^3; 64 for both forms' 48-lattice
// ensemble), where the former rule (fill the chip with at least 3/4
// (FP32) or 4 (FP64) items per CU) took 8 at FP64 192^3 (1.6x slower)
// and 128 at FP32 384^3 (1.2x slower).
int multi_chunk_cap(const Context *c, int TJ, int TK, int H)
{
    const bool f32 = c->prec == AXS_FP32;
    const uint64_t cus = (uint64_t)(c->num_cu > 0 ? c->num_cu : 1);
    int min_side = 1 << 30;
    for (const Member &m : c->members) min_side = m.N < min_side ? m.N : min_side;
    int best = 8;
    uint64_t best_cost = ~0ull;
    for (int cap = f32 ? 256 : 64; cap >= 8; cap >>= 1) {
        if (cap > 8 && 2 * cap > min_side) continue;
        uint64_t items = 0;
        for (const Member &m : c->members)
            items += (uint64_t)((m.N + TJ - 1) / TJ) * (uint64_t)((m.N + TK - 1) / TK) *
                     (uint64_t)(m.N / divisor_below(m.N, cap));
        const uint64_t cost = ((items + cus - 1) / cus) * (uint64_t)(cap + 2 * H);
        if (cost < best_cost) {
            best_cost = cost;
            best = cap;
        }
    }
    return best;
}

// waves per work-group: the configured size; the higher orders, built in
// one shape per form, take that shape when none is configured
int multi_waves_of(const Context *c)
{
    const int f = c->tune.flags;
    if (stencil_radius_of(c) > 1 && !(f & (kTuneWG512 | kTuneWG1024)))
        return (c->prec == AXS_FP64) ? 8 : 16;
    return multi_waves(f);
}

// rows per wave of the higher orders' shapes unless configured: FP64 three
// at radius 2, two beyond (the rings of three rows spill at radius 4)
int high_order_rows(int prec, int RAD)
{
    return (prec == AXS_FP64 && RAD == 2) ? 3 : 2;
}

// rows per wave: the configured count, else the default for the shape
int multi_rows(const Context *c, int N, int D, int W)
{
    const int RAD = stencil_radius_of(c);
    const int r = c->tune.rtb;
    if (r && multi_shape_built(D, W, r, c->prec, RAD)) return r;
    if (RAD > 1) return high_order_rows(c->prec, RAD);
    if (W == 4) return (D == 2 && N >= 512) ? 6 : 4;
    if (c->prec == AXS_FP64) return 2;         // measured best at 1024^3
    return (W == 8) ? 3 : 2;
}

// calls f.run<W, R>() for the built radius-1 shape of form Tr
template <int D, class Tr>
struct Rad1Shapes {
    template <class F>
    int operator()(int W, int R, F &f) const
    {
        if constexpr (std::is_same<Tr, F32Form>::value && D == 5) {
            if (W == 16 && R == 2) return f.template run<16, 2>();
        } else if constexpr (std::is_same<Tr, F32Form>::value) {
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
};

template <class Tr, int D>
static int launch_rad1(Context *c, int src, int dst, const StepScalars *s)
{
    const int W = multi_waves(c->tune.flags);
    return launch_members<Tr, D, 1>(c, src, dst, s, W, multi_rows(c, c->members[0].N, D, W),
                                    multi_rows, Rad1Shapes<D, Tr>{});
}

template <class Tr>
static int launch_form(Context *c, int src, int dst, const StepScalars *s, int depth)
{
    if (depth == 4 || depth == 5) {
        if constexpr (std::is_same<Tr, F32Form>::value) {
            return (depth == 5) ? launch_rad1<Tr, 5>(c, src, dst, s)
                                : launch_rad1<Tr, 4>(c, src, dst, s);
        } else {
            set_error("FP64 depth %d is not built", depth);
            return AXS_ERR_UNSUPPORTED;
        }
    }
    if (depth == 3) return launch_rad1<Tr, 3>(c, src, dst, s);
    return launch_rad1<Tr, 2>(c, src, dst, s);
}

int multi_depth(const Context *c) { return multi_depth_for(c, c->tune.depth); }

int remainder_depth(const Context *c, int r)
{
    const int D = multi_depth(c);
    for (int d = (r < D - 1 ? r : D - 1); d >= 1; --d)
        if (multi_depth_for(c, d) == d) return d;
    return 0;
}

// Depth of the next pass with `left` steps to go: passes of the configured
// depth D; a remainder of one or two steps is taken up by as many passes of
// D + 1 where that depth is built (the FP32 3D pass: a 5-step pass costs
// about what a 4-step pass does per step, a 2-step pass ~1.7 times as much
// and a lone step ~3 times as much), else by shallower passes at the end.
// Every depth gives the bits of one-step launches, so only the time changes.
int next_pass_depth(const Context *c, int64_t left)
{
    const int D = multi_depth(c);
    if (D < 1) return 0;
    const int64_t r = left % D;
    if (r >= 1 && r <= 2 && left >= r * (D + 1) && multi_depth_for(c, D + 1) == D + 1)
        return D + 1;
    return left >= D ? D : remainder_depth(c, (int)left);
}

void invalidate_multi_tables(Context *c)
{
    for (Context::MultiTables &t : c->mtab) t.key = -1;
}

int multi_depth_for(const Context *c, int D)
{
    if (c->tune.kernel == AXS_KERNEL_SIMPLE || D > kMaxDepth) return 0;
    const int RAD = stencil_radius_of(c);
    if (RAD == 1 ? D < 2 : D < 1) return 0;      // radius 1, depth 1: the one-step kernels
    if (c->nd == 2) {
        if (!march2d_depth_built(D, RAD, c->prec)) return 0;
        for (const Member &m : c->members)
            if (!march2d_supported(c, m.N)) return 0;
        return D;
    }
    const int W = multi_waves_of(c);
    if (!multi_shape_built(D, W, multi_rows(c, c->members[0].N, D, W), c->prec, RAD)) return 0;
    for (const Member &m : c->members)
        if (!march2_supported(c, m.N)) return 0;
    return D;
}

int multi_step_depth(const Context *c)
{
    const int D = multi_depth(c);
    return D > 0 ? D : 1;
}

int prepare_multi_batch(Context *c, int D)
{
    if (D < 1) return AXS_OK;
    if (c->nd == 2) return prepare_march2d_items(c, D);
    if (c->members.size() < 2) return AXS_OK;
    Context::MultiTables &mt = c->mtab[D];
    const int RAD = stencil_radius_of(c);
    const int W = multi_waves_of(c);
    // one row count for the ensemble: that of its first member's side
    const int R = multi_rows(c, c->members[0].N, D, W);
    const int key = (((D * 10 + RAD) * 100 + W) * 100 + R) * 10000 + c->tune.ci32;
    if (key == mt.key && mt.tab[0]) return AXS_OK;
    const int TJ = W * R - 2 * D * RAD, TK = tile_cols(D * RAD);
    std::vector<BatchMember> tab[2];
    uint32_t items = 0;
    for (const Member &m : c->members) {
        BatchMember b{};
        b.N = m.N;
        b.L = (uint32_t)m.L;
        b.item_begin = items;
        b.nTj = (m.N + TJ - 1) / TJ;
        b.nTk = (m.N + TK - 1) / TK;
        b.CI = multi_chunk(c, m.N, TJ, TK, D * RAD);
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
        if (!mt.tab[h])
            AXS_HIP_TRY(hipMalloc(&mt.tab[h], c->members.size() * sizeof(BatchMember)));
        AXS_HIP_TRY(hipMemcpyAsync(mt.tab[h], tab[h].data(),
                                   tab[h].size() * sizeof(BatchMember), hipMemcpyHostToDevice,
                                   c->stream));
    }
    AXS_HIP_TRY(hipStreamSynchronize(c->stream));
    mt.items = items;
    mt.key = key;
    return AXS_OK;
}

int launch_multi_steps(Context *c, const StepScalars *s, int depth)
{
    const int src = c->cur, dst = c->cur ^ 1;
    if (c->nd == 2)
        AXS_TRY(launch_march2d(c, src, s, depth));
    else if (stencil_radius_of(c) > 1)
        AXS_TRY(launch_multi_high_order(c, src, dst, s, depth));
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
