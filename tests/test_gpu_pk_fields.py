"""GPU parity of output_powerspec(re, im) for every field selector
(pk_estimation.cpp:37-214: any real array and an optional imaginary one):
phi1, phi2, phidot1, phidot2 and the complex phi1 + i phi2, phidot1 + i
phidot2, in linear and log bins, both forms, against the oracle's complex
transform of the read-back fields."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import random_state

pytestmark = pytest.mark.gpu

PARTS = {"phi2": ("phi2", None), "phi1": ("phi1", None), "phi": ("phi1", "phi2"),
         "phidot2": ("phidot2", None), "phidot1": ("phidot1", None),
         "phidot": ("phidot1", "phidot2")}


@pytest.mark.parametrize("nd,N,prec", [(2, 64, "fp64"), (2, 96, "fp32"), (3, 32, "fp32"),
                                       (3, 48, "fp64")])
@pytest.mark.parametrize("binning", ["linear", "log"])
def test_powerspec_fields_against_oracle(nd, N, prec, binning):
    lat = axs.Lattice(nd, N, precision=prec)
    p1, p2, v1, v2, _, _ = random_state(nd, N, seed=N + nd, vel_scale=0.5)
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    lat.step(5)
    g = lat.get_state()
    for field, (re, im) in PARTS.items():
        ks, pk, cnt = lat.powerspec(binning=binning, nbins=20, field=field)
        a, b = g[re], (g[im] if im else None)
        if binning == "log":
            ks_r, pk_r, cnt_r = O.powerspec_log(a, nd, N, 0.5, 20, field_imag=b)
        else:
            ks_r, pk_r, cnt_r = O.powerspec(a, nd, N, 0.5, field_imag=b)
        np.testing.assert_array_equal(cnt, cnt_r, err_msg=field)
        sel = cnt > 0
        np.testing.assert_allclose(ks[sel], ks_r[sel], rtol=1e-13, err_msg=field)
        np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max(),
                                   err_msg=field)
    # the phi2 selector is the reference's call, bit for bit
    a = lat.powerspec(binning=binning, nbins=20)
    raw = (np.zeros_like(a[0]), np.zeros_like(a[1]), np.zeros_like(a[2]))
    rc = lat.lib.axs_member_powerspec_field(lat._ctx, 0, axs.AXS_FIELD_PHI2,
                                            axs.AXS_BIN_LINEAR if binning == "linear"
                                            else axs.AXS_BIN_LOG, 20,
                                            raw[0].ctypes.data_as(axs.POINTER(axs.c_double)),
                                            raw[1].ctypes.data_as(axs.POINTER(axs.c_double)),
                                            raw[2].ctypes.data_as(axs.POINTER(axs.c_int32)))
    assert rc == axs.AXS_OK
    for x, y in zip(a, raw):
        np.testing.assert_array_equal(x, y)


def test_powerspec_field_hamiltonian_pi_and_errors():
    """in the Hamiltonian form the velocity arrays hold pi; unknown field
    selectors are rejected"""
    nd, N = 2, 64
    lat = axs.Lattice(nd, N, eom="hamiltonian")
    p1, p2, v1, v2, _, _ = random_state(nd, N, seed=3, vel_scale=0.5)
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    lat.step(4)
    g = lat.get_state()
    ks, pk, cnt = lat.powerspec(binning="linear", field="phidot")
    _, pk_r, cnt_r = O.powerspec(g["phidot1"], nd, N, 0.5, field_imag=g["phidot2"])
    np.testing.assert_array_equal(cnt, cnt_r)
    sel = cnt > 0
    np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max())
    out = np.zeros(N // 2)
    cnt = np.zeros(N // 2, np.int32)
    for bad in (-1, 6):
        rc = lat.lib.axs_member_powerspec_field(lat._ctx, 0, bad, axs.AXS_BIN_LINEAR, 0,
                                                out.ctypes.data_as(axs.POINTER(axs.c_double)),
                                                out.ctypes.data_as(axs.POINTER(axs.c_double)),
                                                cnt.ctypes.data_as(axs.POINTER(axs.c_int32)))
        assert rc == axs.AXS_ERR_INVALID_ARG
