// axs_march2d.hip -- the 2D multi-step kernels (axs_march2d.hpp) for the
// 5-point stencil in the FP32 form and the bit-exact FP64 form
// (field-rescaled or Hamiltonian), and the 2D work lists.  Radii 2-4
// (stencil_setting 1-3) are instantiated in axs_march2d_hi.hip.
#include "axs_march2d.hpp"

namespace axs {

bool march2d_supported(const Context *c, int N)
{
    return c->nd == 2 && c->p.laplacian == AXS_LAPLACIAN_STENCIL && N >= kW2;
}

// radius 1: 2, 3, 4, 6 or 8 steps per pass; radii 2-4: 1-4 steps (FP32),
// 1-2 steps (FP64: the register rings grow with the radius)
bool march2d_depth_built(int D, int RAD, int prec)
{
    if (RAD == 1) return D == 2 || D == 3 || D == 4 || D == 6 || D == 8;
    if (RAD > 4) return false;
    return D >= 1 && D <= ((prec == AXS_FP64) ? 2 : 4);
}

// rows per chunk for every member: the configured count if it divides N;
// else the longest chunk of 128, 64 or 32 rows (fewer halo rows per chunk)
// that still deals the chip about 8 waves per SIMD (FP32; 4 in FP64, whose
// kernels hold more registers).  Measured: 128 for the 512-lattice ensemble
// of sides 128..1024 and single lattices from 8192^2 on; 32 (FP32) and 64
// (FP64) at 4096^2, 1.3-1.4x faster there than 128.
static int pick_cap(const Context *c, int TI)
{
    const uint64_t target = (uint64_t)c->num_cu * 4u * ((c->prec == AXS_FP32) ? 8u : 4u);
    for (int cap = 128; cap > 32; cap >>= 1) {
        uint64_t waves = 0;
        for (const Member &m : c->members)
            waves += (uint64_t)((m.N + TI - 1) / TI) * (uint64_t)(m.N / divisor_below(m.N, cap));
        if (waves >= target) return cap;
    }
    return 32;
}

int prepare_march2d_items(Context *c, int D)
{
    if (D < 1) return AXS_OK;
    Context::MultiTables &mt = c->mtab[D];
    const int RAD = stencil_radius_of(c);
    const int key = ((D * 10 + RAD) * 100 + 2) * 10000 + c->tune.ci32;
    if (key == mt.key && mt.tab[0]) return AXS_OK;
    const int TI = kW2 - 2 * D * RAD;
    const int cap = pick_cap(c, TI);
    std::vector<BatchMember> tab[2];
    uint32_t items = 0;
    for (const Member &m : c->members) {
        BatchMember b{};
        b.N = m.N;
        b.L = (uint32_t)m.L;
        b.item_begin = items;
        b.nTj = (m.N + TI - 1) / TI;       // strips
        b.nTk = 1;
        const int cfg = c->tune.ci32;      // rows per chunk
        b.CI = (cfg > 0 && m.N % cfg == 0) ? cfg : divisor_below(m.N, cap);
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

template <class Tr>
static int launch_form2d(Context *c, int src, const StepScalars *s, int depth, bool pf)
{
    switch (depth) {
    case 2: return launch_march2d_pass<Tr, 2, 1>(c, src, s, pf);
    case 3: return launch_march2d_pass<Tr, 3, 1>(c, src, s, pf);
    case 4: return launch_march2d_pass<Tr, 4, 1>(c, src, s, pf);
    case 6: return launch_march2d_pass<Tr, 6, 1>(c, src, s, pf);
    case 8: return launch_march2d_pass<Tr, 8, 1>(c, src, s, pf);
    default:
        set_error("2D depth %d is not built", depth);
        return AXS_ERR_UNSUPPORTED;
    }
}

int launch_march2d(Context *c, int src, const StepScalars *s, int depth)
{
    if (depth < 1 || depth > kMaxDepth || !c->mtab[depth].tab[src]) {
        set_error("2D multi-step work list missing");
        return AXS_ERR_STATE;
    }
    const bool pf = (c->tune.flags & kTunePrefetch) != 0;
    if (stencil_radius_of(c) > 1) return launch_march2d_high_order(c, src, s, depth, pf);
    if (c->prec == AXS_FP32) return launch_form2d<F32Form>(c, src, s, depth, pf);
    if (c->p.eom == AXS_EOM_HAMILTONIAN) return launch_form2d<F64FormT<true>>(c, src, s, depth, pf);
    return launch_form2d<F64FormT<false>>(c, src, s, depth, pf);
}

}  // namespace axs
