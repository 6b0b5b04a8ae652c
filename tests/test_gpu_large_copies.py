"""Results larger than one 64 MiB staging chunk reach the caller's buffer
through the context's pinned staging buffer in overlapped chunks: a
coordinate list of millions of strings (random phases at 256^3 flag most
sites) and the 256^3 axion field, both against the oracle, and a
capacity-limited list that ends inside a chunk."""
import ctypes

import numpy as np
import pytest

import axs
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_large_string_list_and_axion_field(prec):
    N = 256
    rng = np.random.default_rng(77)
    p1, p2 = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3)
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N, precision=prec)
    lat.set_state(p1, p2, z, z, tau=1.0)
    g = lat.get_state(fields=("phi1", "phi2"))
    ax = lat.axion()                                       # 134 MB: three chunks
    np.testing.assert_array_equal(ax, O.atan2_cr(g["phi1"], g["phi2"]))
    n_ref, c_ref = O.cores(ax, 3, N)
    assert n_ref * 3 * 4 > (64 << 20)                      # the list spans several chunks
    n, c = lat.strings()
    assert n == n_ref
    np.testing.assert_array_equal(c, c_ref)
    # a list cut at a capacity that ends inside the second chunk
    cap = ((64 << 20) // 12) + 12345
    buf = np.full((cap, 3), -1, np.int32)
    cnt = ctypes.c_int64()
    rc = lat.lib.axs_member_strings(lat._ctx, 0, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    cap, ctypes.byref(cnt))
    assert rc == axs.AXS_ERR_CAPACITY and cnt.value == n_ref
    np.testing.assert_array_equal(buf, c_ref[:cap])
