// axs_march2d_hi.hip -- the 2D multi-step kernels (axs_march2d.hpp) for the
// higher-order stencils (stencil_setting 1-3: radius 2-4, s_integrate.cpp:
// 9-14): FP32 1-4 steps per pass, FP64 (field-rescaled or Hamiltonian) 1-2.
#include "axs_march2d.hpp"

namespace axs {

template <class Tr, int RAD>
static int launch_rad(Context *c, int src, const StepScalars *s, int depth, bool pf)
{
    switch (depth) {
    case 1: return launch_march2d_pass<Tr, 1, RAD>(c, src, s, pf);
    case 2: return launch_march2d_pass<Tr, 2, RAD>(c, src, s, pf);
    case 3:
    case 4:
        if constexpr (std::is_same<Tr, F32Form>::value) {
            if (depth == 3) return launch_march2d_pass<Tr, 3, RAD>(c, src, s, pf);
            return launch_march2d_pass<Tr, 4, RAD>(c, src, s, pf);
        }
        [[fallthrough]];
    default:
        set_error("2D depth %d is not built for stencil radius %d", depth, RAD);
        return AXS_ERR_UNSUPPORTED;
    }
}

template <class Tr>
static int launch_tr(Context *c, int src, const StepScalars *s, int depth, bool pf)
{
    switch (stencil_radius_of(c)) {
    case 2: return launch_rad<Tr, 2>(c, src, s, depth, pf);
    case 3: return launch_rad<Tr, 3>(c, src, s, depth, pf);
    case 4: return launch_rad<Tr, 4>(c, src, s, depth, pf);
    default:
        set_error("stencil radius %d has no multi-step kernel", stencil_radius_of(c));
        return AXS_ERR_UNSUPPORTED;
    }
}

int launch_march2d_high_order(Context *c, int src, const StepScalars *s, int depth, bool pf)
{
    if (c->prec == AXS_FP32) return launch_tr<F32Form>(c, src, s, depth, pf);
    if (c->p.eom == AXS_EOM_HAMILTONIAN) return launch_tr<F64FormT<true>>(c, src, s, depth, pf);
    return launch_tr<F64FormT<false>>(c, src, s, depth, pf);
}

}  // namespace axs
