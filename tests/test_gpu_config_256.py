"""The 3D 256^3 FP64 configuration at its full size against the CPU oracle:
the k^-1 Gaussian random field from the host generator (seed 1998), lambda =
1, dx = 0.5, dt = 0.1, T_initial = 4, 7-point Laplacian, 1000 steps; the
string finder (bit-exact: same plaquette set, same order, same count) and
the energy reduction every 20 steps, the fields bit for bit at the end."""
import numpy as np
import pytest

import axs
import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def test_256_cubed_fp64_1000_steps_sampled_every_20():
    N, nd, steps, every = 256, 3, 1000, 20
    p1, p2, v1, v2 = axs.host_init("kinv", nd, N, T_initial=4.0, seed=1998)
    lat = axs.Lattice(nd, N, precision="fp64", T_initial=4.0)
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    st = O.State(p1, p2, v1, v2)
    prm = O.params(nd, N, T_initial=4.0)
    A = O.Matrix(nd, N)
    total_strings = 0
    for t in range(0, steps, every):
        # sampling happens before the step, as in run_standard (s_run.cpp:85-155)
        n_gpu, c_gpu = lat.strings()
        n_ref, c_ref = O.cores(O.axion(st.phi1, st.phi2), nd, N)
        assert n_gpu == n_ref, t
        np.testing.assert_array_equal(c_gpu, c_ref, err_msg=f"step {t}")
        total_strings += n_gpu
        e_gpu = lat.energy()
        e_ref = O.energy(prm, st.tau, st.phi1, st.phi2, st.phidot1, st.phidot2)
        np.testing.assert_allclose(e_gpu, e_ref, rtol=1e-11, err_msg=f"step {t}")
        lat.step(every)
        st.run(prm, A, every)
    assert lat.tau == st.tau
    g = lat.get_state()
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=k)
    assert total_strings > 0          # the run forms strings
