"""Ensembles: every member of a batched context evolves exactly as the same
lattice alone (FP64 bit-exact against the oracle), with per-member sampling."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import kinv_field

pytestmark = pytest.mark.gpu


def test_uneven_2d_batch_bit_exact():
    Ns = [32, 64, 128, 64, 32]
    lat = axs.Lattice(2, members=Ns)
    states = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(2, N, seed=1998 + r)
        z = np.zeros(N * N)
        lat.set_state(p1, p2, z, z, member=r)
        states.append((p1, p2, z, z))
    lat.tau = 1.0
    lat.step(60)
    npl, ncores = lat.string_counts()
    for r, N in enumerate(Ns):
        st = O.State(*states[r])
        st.run(O.params(2, N), O.Matrix(2, N), 60)
        g = lat.get_state(member=r)
        for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
            np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=f"member {r} {k}")
        n_ref, c_ref = O.cores(O.axion(st.phi1, st.phi2), 2, N)
        assert npl[r] == n_ref
        assert ncores[r] == O.num_cores2(c_ref)


def test_uneven_3d_batch_fp32_members_match_single():
    Ns = [32, 48, 32]
    lat = axs.Lattice(3, members=Ns, precision="fp32")
    singles = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(3, N, seed=2000 + r)
        z = np.zeros(N ** 3)
        lat.set_state(p1, p2, z, z, member=r)
        s = axs.Lattice(3, N, precision="fp32")
        s.set_state(p1, p2, z, z, tau=1.0)
        singles.append(s)
    lat.tau = 1.0
    lat.step(40)
    for r, s in enumerate(singles):
        s.step(40)
        a, b = lat.get_state(member=r), s.get_state()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"member {r} {k}")
