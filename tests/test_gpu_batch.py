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


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_3d_batch_on_march_kernel_matches_single(precision):
    # sides that the persistent march kernel takes (multiples of 64), one
    # launch per step for the ensemble; every member equals its lone run
    Ns = [64, 128, 192, 64]
    lat = axs.Lattice(3, members=Ns, precision=precision)
    lat.set_kernel("march")
    singles = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(3, N, seed=3000 + r)
        z = np.zeros(N ** 3)
        lat.set_state(p1, p2, z, z, member=r)
        s = axs.Lattice(3, N, precision=precision)
        s.set_kernel("simple")
        s.set_state(p1, p2, z, z, tau=1.0)
        singles.append(s)
    lat.tau = 1.0
    lat.step(6)
    npl, _ = lat.string_counts()
    for r, s in enumerate(singles):
        s.step(6)
        a, b = lat.get_state(member=r), s.get_state()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"member {r} {k}")
        assert npl[r] == s.strings(coords=False)[0]


@pytest.mark.parametrize("nd,Ns", [(2, [64, 128, 32]), (3, [64, 64])])
def test_graph_captured_steps_match_direct(nd, Ns):
    runs = []
    for graph in (False, True):
        lat = axs.Lattice(nd, members=Ns)
        flags = axs.AXS_MARCH_PREFETCH | axs.AXS_MARCH_NT_STORE | (axs.AXS_STEP_GRAPH if graph else 0)
        lat.set_kernel("auto", flags=flags)
        for r, N in enumerate(Ns):
            p1, p2 = kinv_field(nd, N, seed=40 + r)
            z = np.zeros(N ** nd)
            lat.set_state(p1, p2, z, z, member=r)
        lat.tau = 1.0
        lat.step(7)
        lat.step(5)
        runs.append([lat.get_state(member=r) for r in range(len(Ns))])
        assert lat.tau == pytest.approx(1.0 + 12 * float(np.float32(0.1)))
    for a, b in zip(*runs):
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.parametrize("nd,Ns,precision", [(3, [128], "fp32"), (3, [64, 128], "fp64"),
                                             (2, [128, 256], "fp32")])
def test_autotune_keeps_state_and_results(nd, Ns, precision):
    lat = axs.Lattice(nd, members=Ns, precision=precision)
    ref = axs.Lattice(nd, members=Ns, precision=precision)
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(nd, N, seed=50 + r)
        z = np.zeros(N ** nd)
        lat.set_state(p1, p2, z, z, member=r)
        ref.set_state(p1, p2, z, z, member=r)
    lat.tau = ref.tau = 1.0
    before = [lat.get_state(member=r) for r in range(len(Ns))]
    cfg, ms = lat.autotune(trials=3)
    assert ms > 0 and cfg["kernel"] in (1, 2)
    after = [lat.get_state(member=r) for r in range(len(Ns))]
    for a, b in zip(before, after):
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])
    assert lat.tau == 1.0
    lat.step(10)
    ref.step(10)
    for r in range(len(Ns)):
        a, b = lat.get_state(member=r), ref.get_state(member=r)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.parametrize("depth,flags,rows", [(2, 3, 0), (3, 35, 0), (4, 19, 3), (3, 2, 4),
                                               (5, 35, 0)])
def test_multi_step_kernel_on_ensemble_matches_single(depth, flags, rows):
    # several steps per launch on an uneven FP32 ensemble (one persistent
    # launch for all members): every member equals its lone one-step run
    Ns = [64, 128, 96]
    lat = axs.Lattice(3, members=Ns, precision="fp32")
    lat.set_kernel("march", depth=depth, flags=flags, rows=rows)
    singles = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(3, N, seed=6000 + r)
        z = np.zeros(N ** 3)
        lat.set_state(p1, p2, z, z, member=r)
        s = axs.Lattice(3, N, precision="fp32")
        s.set_kernel("simple")
        s.set_state(p1, p2, z, z, tau=1.0)
        singles.append(s)
    lat.tau = 1.0
    lat.step(9)
    for r, s in enumerate(singles):
        s.step(9)
        a, b = lat.get_state(member=r), s.get_state()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"member {r} {k}")


@pytest.mark.parametrize("depth,flags", [(4, 3), (8, 11), (3, 2)])
def test_2d_multi_step_ensemble_matches_single(depth, flags):
    # the ensemble form of the 2D multi-step kernel: one persistent launch
    # per D steps for all members, each member equal to its lone run
    Ns = [128, 256, 64, 100]
    lat = axs.Lattice(2, members=Ns, precision="fp32")
    lat.set_kernel("march", depth=depth, flags=flags)
    singles = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(2, N, seed=7000 + r)
        z = np.zeros(N ** 2)
        lat.set_state(p1, p2, z, z, member=r)
        s = axs.Lattice(2, N, precision="fp32")
        s.set_kernel("simple")
        s.set_state(p1, p2, z, z, tau=1.0)
        singles.append(s)
    lat.tau = 1.0
    lat.step(13)
    npl, _ = lat.string_counts()
    for r, s in enumerate(singles):
        s.step(13)
        a, b = lat.get_state(member=r), s.get_state()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"member {r} {k}")
        assert npl[r] == s.strings(coords=False)[0]


def test_fp64_multi_step_ensemble_bit_exact():
    Ns = [64, 128, 96]
    lat = axs.Lattice(3, members=Ns, precision="fp64")
    lat.set_kernel("march", depth=2, flags=19)
    states = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(3, N, seed=8000 + r)
        z = np.zeros(N ** 3)
        lat.set_state(p1, p2, z, z, member=r)
        states.append((p1, p2, z, z))
    lat.tau = 1.0
    lat.step(7)
    for r, N in enumerate(Ns):
        st = O.State(*states[r])
        st.run(O.params(3, N), O.Matrix(3, N), 7)
        g = lat.get_state(member=r)
        for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
            np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=f"member {r} {k}")


def test_2d_fp64_multi_step_ensemble_bit_exact():
    Ns = [128, 64, 256]
    lat = axs.Lattice(2, members=Ns, precision="fp64")
    lat.set_kernel("march", depth=4, flags=3)
    states = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(2, N, seed=9000 + r)
        z = np.zeros(N ** 2)
        lat.set_state(p1, p2, z, z, member=r)
        states.append((p1, p2, z, z))
    lat.tau = 1.0
    lat.step(11)
    for r, N in enumerate(Ns):
        st = O.State(*states[r])
        st.run(O.params(2, N), O.Matrix(2, N), 11)
        g = lat.get_state(member=r)
        for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
            np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=f"member {r} {k}")
