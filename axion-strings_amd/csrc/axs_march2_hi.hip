// axs_march2_hi.hip -- the 3D multi-step kernels (axs_march2.hpp) for the
// higher-order stencils (stencil_setting 1-3: radius 2-4, s_integrate.cpp:
// 9-14): FP32 one or two steps per pass (radius 4: one) with 16 waves of
// 2 rows; FP64 (field-rescaled or Hamiltonian) one step with 8 waves of 2
// rows (radius 2: or 3).
#include "axs_march2.hpp"

namespace axs {

template <class Tr, int RAD>
struct HiShapes {
    template <class F>
    int operator()(int W, int R, F &f) const
    {
        if constexpr (std::is_same<Tr, F32Form>::value) {
            if (W == 16 && R == 2) return f.template run<16, 2>();
        } else {
            if constexpr (RAD == 2)
                if (W == 8 && R == 3) return f.template run<8, 3>();
            if (W == 8 && R == 2) return f.template run<8, 2>();
        }
        set_error("this work-group shape is not built for the higher-order stencils");
        return AXS_ERR_UNSUPPORTED;
    }
};

template <class Tr, int RAD>
static int launch_rad(Context *c, int src, int dst, const StepScalars *s, int depth)
{
    const int W = multi_waves_of(c);
    const int R = multi_rows(c, c->members[0].N, depth, W);
    if (!multi_shape_built(depth, W, R, c->prec, RAD)) {
        set_error("depth %d is not built for stencil radius %d in this form", depth, RAD);
        return AXS_ERR_UNSUPPORTED;
    }
    if constexpr (std::is_same<Tr, F32Form>::value && RAD < 4) {
        if (depth == 2)
            return launch_members<Tr, 2, RAD>(c, src, dst, s, W, R, multi_rows, HiShapes<Tr, RAD>{});
    }
    return launch_members<Tr, 1, RAD>(c, src, dst, s, W, R, multi_rows, HiShapes<Tr, RAD>{});
}

template <class Tr>
static int launch_tr(Context *c, int src, int dst, const StepScalars *s, int depth)
{
    switch (stencil_radius_of(c)) {
    case 2: return launch_rad<Tr, 2>(c, src, dst, s, depth);
    case 3: return launch_rad<Tr, 3>(c, src, dst, s, depth);
    case 4: return launch_rad<Tr, 4>(c, src, dst, s, depth);
    default:
        set_error("stencil radius %d has no multi-step kernel", stencil_radius_of(c));
        return AXS_ERR_UNSUPPORTED;
    }
}

int launch_multi_high_order(Context *c, int src, int dst, const StepScalars *s, int depth)
{
    if (c->prec == AXS_FP32) return launch_tr<F32Form>(c, src, dst, s, depth);
    if (c->p.eom == AXS_EOM_HAMILTONIAN) return launch_tr<F64FormT<true>>(c, src, dst, s, depth);
    return launch_tr<F64FormT<false>>(c, src, dst, s, depth);
}

}  // namespace axs
