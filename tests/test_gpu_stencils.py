"""GPU parity of the higher-order stencils (stencil_setting 1-3: 4th, 6th and
8th-order Laplacians of radius 2-4, s_integrate.cpp:9-14) on the multi-step
kernels: one to four steps per pass over HBM, bit-identical to one-step
launches in the FP32 form and to the oracle in the FP64 form, single
lattices (also sides the tiles do not divide) and uneven ensembles."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import kinv_field, random_state, rel_l2

pytestmark = pytest.mark.gpu

PF, NT = axs.AXS_MARCH_PREFETCH, axs.AXS_MARCH_NT_STORE
WG512, WG1024 = axs.AXS_MARCH_WG512, axs.AXS_MARCH_WG1024


def _oracle_run(nd, N, state, nsteps, stencil):
    p = O.params(nd, N, stencil_setting=stencil)
    st = O.State(*state)
    st.run(p, O.Matrix(nd, N, stencil), nsteps)
    return st


def _assert_oracle(lat, st, member=0):
    g = lat.get_state(member=member)
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=k)
    assert lat.tau == st.tau


def _simple(nd, Ns, states, nsteps, stencil, precision):
    lat = axs.Lattice(nd, members=Ns, precision=precision, stencil_setting=stencil)
    lat.set_kernel("simple")
    for r, s in enumerate(states):
        lat.set_state(*s, member=r)
    lat.tau = 1.0
    lat.step(nsteps)
    return lat


def _assert_same(a, b, n_members=1):
    for r in range(n_members):
        ga, gb = a.get_state(member=r), b.get_state(member=r)
        for k in ga:
            np.testing.assert_array_equal(ga[k], gb[k], err_msg=f"member {r} {k}")
    assert a.tau == b.tau


@pytest.mark.parametrize("N,stencil,depth,chunk,flags,nsteps", [
    (64, 1, 1, 0, PF | NT | WG1024, 5), (64, 1, 2, 0, PF | NT | WG1024, 7),
    (80, 2, 2, 16, NT | WG1024, 5), (64, 2, 1, 0, NT, 4), (100, 3, 1, 0, PF | NT | WG1024, 4),
    (128, 1, 2, 32, PF | NT | WG1024, 6), (96, 2, 2, 0, PF | NT, 9)])
def test_3d_fp32_multi_step_matches_one_step(N, stencil, depth, chunk, flags, nsteps):
    # no work-group size configured (flags without WG1024): the built one
    state = random_state(3, N, seed=40 + stencil, ker=True)
    a = _simple(3, [N], [state], nsteps, stencil, "fp32")
    b = axs.Lattice(3, N, precision="fp32", stencil_setting=stencil)
    b.set_kernel("march", chunk=chunk, depth=depth, flags=flags)
    assert b.kernel_config()["depth"] == depth
    assert b.bytes_per_site() == (32.0 / depth, 32.0)
    b.set_state(*state, tau=1.0)
    b.step(nsteps)
    _assert_same(a, b)


@pytest.mark.parametrize("N,stencil,rows,chunk,flags,nsteps", [
    (64, 1, 0, 0, PF | NT | WG512, 4), (64, 1, 2, 16, NT | WG512, 3), (80, 2, 0, 0, PF | NT, 3),
    (64, 3, 2, 0, PF | NT | WG512, 3), (96, 1, 3, 32, PF | NT | WG512, 3)])
def test_3d_fp64_multi_step_bit_exact_vs_oracle(N, stencil, rows, chunk, flags, nsteps):
    state = random_state(3, N, seed=50 + stencil, ker=True)
    st = _oracle_run(3, N, state, nsteps, stencil)
    lat = axs.Lattice(3, N, stencil_setting=stencil)
    lat.set_kernel("march", rows=rows, chunk=chunk, depth=1, flags=flags)
    assert lat.kernel_config()["depth"] == 1
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    _assert_oracle(lat, st)


@pytest.mark.parametrize("N,stencil,depth,chunk,flags,nsteps", [
    (64, 1, 1, 0, PF | NT, 6), (64, 1, 4, 0, PF | NT, 9), (100, 2, 3, 0, NT, 10),
    (128, 3, 2, 32, PF | NT, 7), (128, 3, 4, 0, PF | NT, 8), (96, 2, 4, 0, PF | NT | 8, 13)])
def test_2d_fp32_multi_step_matches_one_step(N, stencil, depth, chunk, flags, nsteps):
    state = random_state(2, N, seed=60 + stencil, ker=True)
    a = _simple(2, [N], [state], nsteps, stencil, "fp32")
    b = axs.Lattice(2, N, precision="fp32", stencil_setting=stencil)
    b.set_kernel("march", chunk=chunk, depth=depth, flags=flags)
    assert b.kernel_config()["depth"] == depth
    b.set_state(*state, tau=1.0)
    b.step(nsteps)
    _assert_same(a, b)


@pytest.mark.parametrize("N,stencil,depth,chunk,nsteps", [
    (64, 1, 1, 0, 5), (64, 2, 2, 0, 9), (100, 3, 2, 0, 7), (128, 3, 1, 32, 4),
    (128, 1, 2, 64, 6)])
def test_2d_fp64_multi_step_bit_exact_vs_oracle(N, stencil, depth, chunk, nsteps):
    state = random_state(2, N, seed=70 + stencil, ker=True)
    st = _oracle_run(2, N, state, nsteps, stencil)
    lat = axs.Lattice(2, N, stencil_setting=stencil)
    lat.set_kernel("march", chunk=chunk, depth=depth, flags=PF | NT)
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    _assert_oracle(lat, st)


@pytest.mark.parametrize("nd,Ns,stencil,precision,depth,nsteps", [
    (3, [64, 96, 64], 2, "fp32", 2, 5), (3, [80, 64], 1, "fp64", 1, 3),
    (3, [64, 128], 3, "fp32", 1, 3),
    (2, [64, 128, 96], 3, "fp32", 3, 7), (2, [128, 64, 100], 2, "fp64", 2, 5)])
def test_multi_step_ensembles_match_one_step(nd, Ns, stencil, precision, depth, nsteps):
    # one persistent launch per pass over every member of an uneven ensemble
    states = [random_state(nd, N, seed=80 + r, ker=True) for r, N in enumerate(Ns)]
    a = _simple(nd, Ns, states, nsteps, stencil, precision)
    b = axs.Lattice(nd, members=Ns, precision=precision, stencil_setting=stencil)
    b.set_kernel("march", depth=depth, flags=PF | NT)
    for r, s in enumerate(states):
        b.set_state(*s, member=r)
    b.tau = 1.0
    b.step(nsteps)
    _assert_same(a, b, len(Ns))


@pytest.mark.parametrize("nd,N,stencil,precision,depth,flags", [
    (3, 64, 1, "fp32", 2, PF | NT | WG1024), (3, 64, 3, "fp32", 1, PF | NT | WG1024),
    (3, 64, 1, "fp64", 1, PF | NT | WG512), (3, 64, 3, "fp64", 1, PF | NT | WG512),
    (2, 64, 1, "fp32", 4, PF | NT), (2, 64, 2, "fp32", 3, PF | NT), (2, 64, 3, "fp32", 2, PF | NT),
    (2, 64, 3, "fp64", 2, PF | NT),
    (3, 48, 1, "fp32", 1, PF | NT), (2, 48, 2, "fp64", 1, PF | NT)])
def test_default_configuration(nd, N, stencil, precision, depth, flags):
    # the multi-step kernel from side 64 on, in its fastest measured depth
    lat = axs.Lattice(nd, N, precision=precision, stencil_setting=stencil)
    cfg = lat.kernel_config()
    assert cfg["depth"] == depth and cfg["flags"] == flags


@pytest.mark.parametrize("nd,stencil,precision,depth,rows,flags,kw", [
    (3, 1, "fp64", 2, 0, PF | NT | WG512, {}),        # FP64: one step per pass
    (3, 3, "fp32", 2, 0, PF | NT | WG1024, {}),       # radius 4: one step
    (3, 2, "fp32", 1, 0, PF | NT | WG512, {}),        # FP32: 16 waves
    (3, 3, "fp64", 1, 3, PF | NT | WG512, {}),        # 3 rows spill beyond radius 2
    (3, 1, "fp32", 3, 0, PF | NT | WG1024, {}),
    (2, 2, "fp64", 3, 0, PF | NT, {}),
    (2, 1, "fp32", 5, 0, PF | NT, {}),
    (2, 1, "fp32", 1, 0, PF | NT | WG512, {}),
    (3, 1, "fp64", 2, 0, PF | NT | WG512, {"eom": "hamiltonian"}),
    (2, 2, "fp64", 3, 0, PF | NT, {"eom": "hamiltonian"})])
def test_unbuilt_configurations_rejected(nd, stencil, precision, depth, rows, flags, kw):
    lat = axs.Lattice(nd, 64, precision=precision, stencil_setting=stencil, **kw)
    before = lat.kernel_config()
    with pytest.raises(axs.AxsError):
        lat.set_kernel("march", rows=rows, depth=depth, flags=flags)
    assert lat.kernel_config() == before


@pytest.mark.parametrize("nd,N,stencil,depth,nsteps", [(3, 64, 1, 1, 3), (3, 80, 3, 1, 2),
                                                   (2, 128, 2, 2, 7), (2, 64, 3, 1, 5)])
def test_hamiltonian_form_higher_order_multi_step_bit_exact(nd, N, stencil, depth, nsteps):
    # the Hamiltonian form (vvsl_hamiltonian_form, s_integrate.cpp:358-395)
    # with the higher-order stencils on the multi-step kernels, bit for bit
    state = random_state(nd, N, seed=90, ker=True)
    lat = axs.Lattice(nd, N, eom="hamiltonian", stencil_setting=stencil)
    lat.set_kernel("march", depth=depth, flags=PF | NT | (WG512 if nd == 3 else 0))
    assert lat.kernel_config()["depth"] == depth
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    p = O.params(nd, N, stencil_setting=stencil)
    st = O.State(*state)
    st.run_form(p, O.Matrix(nd, N, stencil), nsteps, eom=1, laplacian=0)
    _assert_oracle(lat, st)


@pytest.mark.parametrize("nd,N,precision", [(3, 64, "fp32"), (2, 128, "fp64")])
def test_autotune_higher_order(nd, N, precision):
    # the tuner's choice runs and keeps the bits of the one-step kernels
    stencil, nsteps = 2, 4
    state = random_state(nd, N, seed=95, ker=True)
    a = _simple(nd, [N], [state], nsteps, stencil, precision)
    b = axs.Lattice(nd, N, precision=precision, stencil_setting=stencil)
    b.set_state(*state, tau=1.0)
    cfg, ms = b.autotune(trials=2)
    assert ms > 0
    b.set_state(*state, tau=1.0)
    b.step(nsteps)
    _assert_same(a, b)


@pytest.mark.parametrize("nd,N,stencil,nsteps", [(3, 64, 2, 200), (2, 256, 3, 300)])
def test_fp32_form_accuracy_higher_order(nd, N, stencil, nsteps):
    # the FP32 form (fused multiply-adds, centre term in the potential) on
    # the multi-step kernels against the FP64 form on the same field
    p1, p2 = kinv_field(nd, N, seed=1998)
    z = np.zeros(N ** nd)
    out = {}
    for prec in ("fp32", "fp64"):
        lat = axs.Lattice(nd, N, precision=prec, stencil_setting=stencil)
        assert lat.kernel_config()["depth"] >= 1
        lat.set_state(p1, p2, z, z, tau=1.0)
        lat.step(nsteps)
        out[prec] = (lat.get_state(), lat.strings(coords=False)[0])
    for k in ("phi1", "phi2"):
        assert rel_l2(out["fp32"][0][k], out["fp64"][0][k]) < 1e-4, k
    n32, n64 = out["fp32"][1], out["fp64"][1]
    assert abs(n32 - n64) <= max(1, 0.01 * n64)
