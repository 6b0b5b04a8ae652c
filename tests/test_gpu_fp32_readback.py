"""The FP32 form's read-back ker (one reconstruction pass per state, the
slow-axis march of canonical_fp32_kernel) bit for bit against a float32
emulation of the kernels' arithmetic: each fused multiply-add is a long-double
product (exact: 24 x 24 bits) plus a long-double sum, rounded once to float.
The stencil coefficients are the reference's float-rounded ones
(s_integrate.cpp:9-14), the scalars those of the field-rescaled form
(s_integrate.cpp:305-348) rounded to float as the kernels take them."""
import numpy as np
import pytest

import axs
from fields import random_state

pytestmark = pytest.mark.gpu

f32 = np.float32


def _fma(a, b, c):
    ld = np.longdouble
    return (np.asarray(a, ld) * np.asarray(b, ld) + np.asarray(c, ld)).astype(np.float32)


def _coefs(stencil, nd):
    n = f32(nd)
    table = {
        0: [f32(-2.0) * n, f32(1.0)],
        1: [f32(-5.0) / f32(2.0) * n, f32(4.0) / f32(3.0), f32(-1.0) / f32(12.0)],
        2: [f32(-49.0) / f32(18.0) * n, f32(3.0) / f32(2.0), f32(-3.0) / f32(20.0),
            f32(1.0) / f32(90.0)],
        3: [f32(-205.0) / f32(72.0) * n, f32(8.0) / f32(5.0), f32(-1.0) / f32(5.0),
            f32(8.0) / f32(315.0), f32(-1.0) / f32(560.0)],
    }
    return [float(c) for c in table[stencil]]


def _ker_emulated(x, y, nd, stencil, tau, dx=0.5, lam=1.0, T0=4.0):
    c = _coefs(stencil, nd)
    alpha = 1.0 / (float(f32(dx)) ** 2)
    cl = f32((-1.0 / (tau * tau)) * float(f32(lam)))
    pot0 = f32((T0 * T0) / float(f32(3.0)) - tau * tau)
    ac0 = f32(alpha * c[0])
    alpha = f32(alpha)
    # numpy axes of the reference layout: 3D [i, j, k] (axis 0 = i);
    # 2D [j, i].  The FP32 form sums 3D i, j, k and 2D j, then i (numpy
    # axis 0, then 1)
    axes = [0, 1, 2] if nd == 3 else [0, 1]
    s1 = np.full(x.shape, -0.0, np.float32)
    s2 = np.full(x.shape, -0.0, np.float32)
    for l in range(1, len(c)):
        cl_ = f32(c[l])
        for ax in axes:
            for sh in (-l, l):          # +l neighbour first, then -l
                s1 = _fma(cl_, np.roll(x, sh, axis=ax), s1)
                s2 = _fma(cl_, np.roll(y, sh, axis=ax), s2)
    S = _fma(x, x, _fma(y, y, pot0))
    t = _fma(cl, S, ac0)
    k1 = _fma(alpha, s1, (x * t).astype(np.float32))
    k2 = _fma(alpha, s2, (y * t).astype(np.float32))
    return k1, k2


@pytest.mark.parametrize("nd,N,stencil", [(3, 20, 0), (3, 20, 1), (3, 24, 2), (3, 24, 3),
                                          (3, 66, 0), (2, 40, 0), (2, 40, 2), (2, 64, 3)])
def test_fp32_readback_ker_bit_exact(nd, N, stencil):
    lat = axs.Lattice(nd, N, precision="fp32", stencil_setting=stencil, T_initial=4.0)
    p1, p2, v1, v2, _, _ = random_state(nd, N, seed=9)
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    lat.step(3)
    g = lat.get_state()
    shape = (N,) * nd
    x = g["phi1"].astype(np.float32).reshape(shape)
    y = g["phi2"].astype(np.float32).reshape(shape)
    k1, k2 = _ker_emulated(x, y, nd, stencil, lat.tau)
    np.testing.assert_array_equal(g["ker1"].reshape(shape), k1.astype(np.float64))
    np.testing.assert_array_equal(g["ker2"].reshape(shape), k2.astype(np.float64))
