"""axs_member_energy_noalloc: the energy entry point that allocates nothing
(partial sums in a fixed per-GPU device workspace) gives axs_member_energy's
result bit for bit, on two calls in a row, the second on a smaller lattice,
in both forms; the FP64 result also matches the oracle (the energy
functional of the field-rescaled form)."""
import ctypes

import numpy as np
import pytest

import axs
import oracle as O
from fields import random_state

pytestmark = pytest.mark.gpu


def _energy_noalloc_abi(lat, member=0):
    # straight through the C-ABI
    out = np.zeros(3, np.float64)
    rc = lat.lib.axs_member_energy_noalloc(lat._ctx, member,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == axs.AXS_OK, lat.lib.axs_last_error().decode()
    return out


def _pair(nd, N, prec, seed, nsteps):
    """two contexts with the same state: one for each entry point (the FP32
    form keeps its sums per state, so one context would answer the second
    call from the first one's result)"""
    state = random_state(nd, N, seed=seed)
    out = []
    for _ in range(2):
        lat = axs.Lattice(nd, N, precision=prec)
        lat.set_state(*state, tau=1.0)
        lat.step(nsteps)
        out.append(lat)
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nd,sizes", [(3, (96, 40)), (2, (512, 130))])
def test_noalloc_energy_equals_energy(prec, nd, sizes):
    for N in sizes:                            # the second call on a smaller lattice
        ref_lat, lat = _pair(nd, N, prec, seed=N, nsteps=7)
        want = ref_lat.energy()
        got = _energy_noalloc_abi(lat)
        np.testing.assert_array_equal(got, want)
        if prec == "fp64":
            g = lat.get_state()
            ref = O.energy(O.params(nd, N), lat.tau, g["phi1"], g["phi2"], g["phidot1"],
                           g["phidot2"])
            np.testing.assert_allclose(got, ref, rtol=1e-11)
        # again on a new state of the same context
        lat.step(3)
        ref_lat.step(3)
        np.testing.assert_array_equal(_energy_noalloc_abi(lat), ref_lat.energy())
        ref_lat.close()
        lat.close()


def test_noalloc_energy_ensemble_members_and_errors():
    Ns = [64, 32, 48]
    a = axs.Lattice(3, members=Ns)
    b = axs.Lattice(3, members=Ns)
    for r, N in enumerate(Ns):
        st = random_state(3, N, seed=50 + r)
        a.set_state(*st, member=r)
        b.set_state(*st, member=r)
    a.tau = b.tau = 1.0
    a.step(4)
    b.step(4)
    for r in (0, 2, 1):
        np.testing.assert_array_equal(_energy_noalloc_abi(b, r), a.energy(member=r))
    out = np.zeros(3)
    assert b.lib.axs_member_energy_noalloc(b._ctx, 3, out.ctypes.data_as(
        ctypes.POINTER(ctypes.c_double))) == axs.AXS_ERR_INVALID_ARG
    assert b.lib.axs_member_energy_noalloc(b._ctx, 0, None) == axs.AXS_ERR_INVALID_ARG
