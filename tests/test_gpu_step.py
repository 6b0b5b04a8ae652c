"""GPU parity of the step (vvsl_field_rescaled, s_integrate.cpp:305-348):
the FP64 form against the oracle, bit for bit; the FP32 form against FP64
within the stated tolerance."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import kinv_field, random_state, rel_l2

pytestmark = pytest.mark.gpu


def _oracle_run(nd, N, state, nsteps, **kw):
    p = O.params(nd, N, **kw)
    st = O.State(*state)
    st.run(p, O.Matrix(nd, N, kw.get("stencil_setting", 0)), nsteps)
    return st


def _gpu_run(nd, N, state, nsteps, precision="fp64", **kw):
    lat = axs.Lattice(nd, N, precision=precision, **kw)
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    return lat


def _assert_same(lat, st):
    g = lat.get_state()
    for k in ("phi1", "phi2", "phidot1", "phidot2"):
        np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=k)
    np.testing.assert_array_equal(g["ker1"], st.ker1)
    np.testing.assert_array_equal(g["ker2"], st.ker2)
    assert lat.tau == st.tau


def test_fp64_2d_256_500_steps_bit_exact():
    # 2D 256^2, k^-1 field, phidot = 0, ker = 0, lambda=1, dx=0.5, dt=0.1
    N = 256
    p1, p2 = kinv_field(2, N, seed=1998)
    z = np.zeros(N * N)
    state = (p1, p2, z, z)
    st = _oracle_run(2, N, state, 500)
    lat = _gpu_run(2, N, state, 500)
    _assert_same(lat, st)


@pytest.mark.parametrize("N,nsteps", [(64, 120), (48, 60)])
def test_fp64_3d_bit_exact_with_velocity_and_ker(N, nsteps):
    state = random_state(3, N, seed=N, ker=True)
    st = _oracle_run(3, N, state, nsteps)
    lat = _gpu_run(3, N, state, nsteps)
    _assert_same(lat, st)


def test_fp64_3d_256_first_steps_bit_exact():
    # the 3D 256^3 lattice itself, 8 steps from the k^-1 field
    N = 256
    p1, p2 = kinv_field(3, N, seed=1998)
    z = np.zeros(N ** 3)
    st = _oracle_run(3, N, (p1, p2, z, z), 8)
    lat = _gpu_run(3, N, (p1, p2, z, z), 8)
    _assert_same(lat, st)


@pytest.mark.parametrize("nd,N,stencil", [(2, 64, 1), (2, 64, 2), (2, 64, 3), (3, 24, 1), (3, 24, 3)])
def test_fp64_higher_order_stencils_bit_exact(nd, N, stencil):
    rng = np.random.default_rng(stencil)
    L = N ** nd
    state = (rng.standard_normal(L), rng.standard_normal(L), 0.1 * rng.standard_normal(L),
             0.1 * rng.standard_normal(L))
    st = _oracle_run(nd, N, state, 25, stencil_setting=stencil)
    lat = _gpu_run(nd, N, state, 25, stencil_setting=stencil)
    _assert_same(lat, st)


def test_split_calls_equal_one_call():
    # step(a); step(b) == step(a+b): tau and state carry across calls
    N = 64
    state = random_state(2, N, seed=3)
    a = _gpu_run(2, N, state, 70)
    b = _gpu_run(2, N, state, 30)
    b.step(40)
    ga, gb = a.get_state(), b.get_state()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k])
    assert a.tau == b.tau


@pytest.mark.parametrize("nd,N,nsteps", [(2, 256, 500), (3, 64, 300)])
def test_fp32_form_against_fp64(nd, N, nsteps):
    # FP32 phi / velocity vs the FP64 form on the same input: field rel-L2
    # <= 1e-4 (the precision form's accuracy target)
    p1, p2 = kinv_field(nd, N, seed=1998)
    z = np.zeros(N ** nd)
    f64 = _gpu_run(nd, N, (p1, p2, z, z), nsteps, precision="fp64").get_state()
    f32 = _gpu_run(nd, N, (p1, p2, z, z), nsteps, precision="fp32").get_state()
    err = max(rel_l2(f32["phi1"], f64["phi1"]), rel_l2(f32["phi2"], f64["phi2"]))
    verr = max(rel_l2(f32["phidot1"], f64["phidot1"]), rel_l2(f32["phidot2"], f64["phidot2"]))
    print(f"fp32 vs fp64 {nd}D N={N} steps={nsteps}: field rel-L2 {err:.3e}, velocity {verr:.3e}")
    assert err <= 1e-4
    assert verr <= 1e-3


def test_fp32_state_round_trip():
    N = 32
    state = random_state(2, N, seed=4, ker=True)
    lat = axs.Lattice(2, N, precision="fp32")
    lat.set_state(*state, tau=1.0)
    g = lat.get_state()
    for k, ref in zip(("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"), state):
        np.testing.assert_allclose(g[k], ref, rtol=0, atol=2e-6 * max(1.0, np.abs(ref).max()))


def test_step_timing_reported():
    # the HIP-event time per step brackets the step kernels: on a lattice
    # where they dominate (512^3 FP64, ~2.4 ms per one-step launch, 40
    # launches) it is the host's wall time less the call's own set-up and
    # synchronisation -- within 10% below and never above it
    import time
    N = 512
    lat = axs.Lattice(3, N, precision="fp64")
    lat.init("kinv", seed=3)
    lat.tau = 1.0
    for graph in (False, True):
        flags = axs.AXS_MARCH_PREFETCH | axs.AXS_MARCH_NT_STORE | (axs.AXS_STEP_GRAPH if graph else 0)
        lat.set_kernel("auto", flags=flags)
        lat.step(2)
        t0 = time.perf_counter()
        lat.step(40)
        wall_ms = (time.perf_counter() - t0) * 1e3 / 40
        assert 0.9 * wall_ms <= lat.last_step_ms <= wall_ms, (graph, lat.last_step_ms, wall_ms)
    s, a = lat.bytes_per_site()
    assert (s, a) == (96.0, 64.0)


@pytest.mark.parametrize("precision,rows,chunk,flags", [
    ("fp64", 4, 0, 1), ("fp64", 4, 16, 0), ("fp64", 4, 128, 3), ("fp64", 2, 32, 1), ("fp64", 2, 0, 7),
    ("fp32", 4, 0, 1), ("fp32", 8, 32, 0), ("fp32", 8, 128, 3), ("fp32", 4, 64, 5), ("fp32", 4, 16, 2)])
def test_march_kernel_matches_simple(precision, rows, chunk, flags):
    # the 2.5D march kernel and the one-site-per-work-item kernel compute
    # the same expression trees: identical bits for every launch shape
    N, nsteps = 128, 12
    state = random_state(3, N, seed=5, ker=True)
    a = axs.Lattice(3, N, precision=precision)
    a.set_kernel("simple")
    a.set_state(*state, tau=1.0)
    a.step(nsteps)
    b = axs.Lattice(3, N, precision=precision)
    b.set_kernel("march", rows=rows, chunk=chunk, flags=flags)
    b.set_state(*state, tau=1.0)
    b.step(nsteps)
    ga, gb = a.get_state(), b.get_state()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_march_kernel_fp64_oracle_192():
    # a side that is a multiple of 64 but not a power of two (the 192^3
    # ensemble members), march kernel against the oracle
    N, nsteps = 192, 4
    state = random_state(3, N, seed=9)
    st = _oracle_run(3, N, state, nsteps)
    lat = axs.Lattice(3, N)
    lat.set_kernel("march")
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    _assert_same(lat, st)


@pytest.mark.parametrize("N,depth,rows,chunk,flags,nsteps", [
    (64, 2, 4, 0, 3, 8), (64, 2, 8, 16, 2, 7), (100, 2, 6, 0, 3, 9), (128, 2, 6, 32, 3, 12),
    (192, 2, 4, 64, 3, 5), (128, 2, 6, 0, 11, 10), (128, 2, 3, 0, 19, 6), (96, 2, 2, 0, 18, 5),
    (64, 3, 4, 0, 3, 9), (100, 3, 6, 0, 3, 10), (128, 3, 4, 32, 2, 8), (128, 3, 4, 0, 19, 7),
    (192, 3, 3, 0, 27, 6), (128, 3, 6, 16, 3, 3), (80, 3, 2, 0, 18, 11), (128, 3, 2, 0, 34, 7),
    (128, 4, 3, 0, 19, 9), (100, 4, 2, 0, 35, 8), (192, 4, 4, 32, 18, 7), (64, 2, 3, 16, 35, 6),
    (64, 4, 3, 0, 34, 12), (128, 5, 2, 0, 35, 11), (100, 5, 2, 0, 35, 13),
    (192, 5, 2, 32, 34, 10), (64, 5, 2, 16, 35, 7), (66, 4, 2, 0, 35, 9), (70, 5, 2, 0, 35, 7),
    (130, 4, 2, 0, 35, 6), (122, 3, 3, 0, 19, 5)])
def test_multi_step_kernel_matches_one_step(N, depth, rows, chunk, flags, nsteps):
    # two to five time steps per pass (temporal blocking) evaluate the same
    # expressions as one-step launches: identical bits on every side (also
    # sides the tiles do not divide), launch shape (4, 8 or 16 waves per
    # work-group: AXS_MARCH_WG512 / AXS_MARCH_WG1024) and step count (a
    # count that is not a multiple of the depth ends with shallower passes)
    state = random_state(3, N, seed=11, ker=True)
    a = axs.Lattice(3, N, precision="fp32")
    a.set_kernel("simple")
    a.set_state(*state, tau=1.0)
    a.step(nsteps)
    b = axs.Lattice(3, N, precision="fp32")
    b.set_kernel("march", rows=rows, chunk=chunk, depth=depth, flags=flags)
    cfg = b.kernel_config()
    assert cfg["depth"] == depth and cfg["rows"] == rows
    assert b.bytes_per_site() == (32.0 / depth, 32.0)
    b.set_state(*state, tau=1.0)
    b.step(nsteps)
    ga, gb = a.get_state(), b.get_state()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
    assert a.tau == b.tau
    na, _ = a.strings(coords=False)
    nb, _ = b.strings(coords=False)
    assert na == nb


@pytest.mark.parametrize("nd,N,precision,kernel", [
    (3, 64, "fp64", "march"), (2, 48, "fp32", "auto"), (3, 48, "fp32", "march"),
    (3, 64, "fp32", "simple")])
def test_two_step_kernel_rejects_other_forms(nd, N, precision, kernel):
    # a rejected configuration leaves the context's own one in place (a 3D
    # context below side 256 starts at depth 3, FP32 and FP64 alike)
    lat = axs.Lattice(nd, N, precision=precision)
    before = lat.kernel_config()
    assert before["depth"] == (3 if (nd, N) == (3, 64) else 1)
    with pytest.raises(axs.AxsError, match="depth 2"):
        lat.set_kernel(kernel, depth=2)
    assert lat.kernel_config() == before


@pytest.mark.parametrize("depth,rows,flags", [(4, 0, 3), (5, 0, 19), (3, 8, 3), (2, 6, 19),
                                              (2, 5, 3), (3, 4, 35), (2, 2, 51), (5, 3, 35),
                                              (6, 0, 35)])
def test_multi_step_kernel_rejects_unbuilt_shapes(depth, rows, flags):
    lat = axs.Lattice(3, 64, precision="fp32")
    before = lat.kernel_config()
    with pytest.raises(axs.AxsError):
        lat.set_kernel("march", rows=rows, depth=depth, flags=flags)
    assert lat.kernel_config() == before


@pytest.mark.parametrize("N,depth,chunk,flags,nsteps", [
    (64, 2, 0, 3, 8), (64, 8, 16, 2, 19), (96, 3, 0, 3, 10), (100, 4, 0, 3, 9),
    (128, 6, 32, 3, 13), (256, 8, 0, 11, 17), (128, 4, 64, 2, 7)])
def test_2d_multi_step_kernel_matches_one_step(N, depth, chunk, flags, nsteps):
    # 2D strips (one wave per 64 columns, i neighbours by lane shuffles):
    # D steps per pass are bit-identical to one-step launches
    state = random_state(2, N, seed=12, ker=True)
    a = axs.Lattice(2, N, precision="fp32")
    a.set_kernel("simple")
    a.set_state(*state, tau=1.0)
    a.step(nsteps)
    b = axs.Lattice(2, N, precision="fp32")
    b.set_kernel("march", chunk=chunk, depth=depth, flags=flags)
    assert b.kernel_config()["depth"] == depth
    assert b.bytes_per_site() == (32.0 / depth, 32.0)
    b.set_state(*state, tau=1.0)
    b.step(nsteps)
    ga, gb = a.get_state(), b.get_state()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
    assert a.tau == b.tau


@pytest.mark.parametrize("depth,rows,flags", [(5, 0, 3), (4, 2, 3), (4, 0, 19), (10, 0, 3)])
def test_2d_multi_step_rejects_unbuilt_shapes(depth, rows, flags):
    lat = axs.Lattice(2, 128, precision="fp32")
    before = lat.kernel_config()
    with pytest.raises(axs.AxsError):
        lat.set_kernel("march", rows=rows, depth=depth, flags=flags)
    assert lat.kernel_config() == before


@pytest.mark.parametrize("N,depth,rows,chunk,flags,nsteps", [
    (64, 2, 2, 0, 19, 8), (64, 3, 3, 16, 18, 7), (100, 2, 3, 0, 19, 9), (128, 3, 2, 32, 27, 10),
    (192, 2, 2, 0, 18, 5), (66, 3, 2, 0, 19, 6), (70, 2, 3, 0, 19, 5)])
def test_fp64_multi_step_bit_exact_vs_oracle(N, depth, rows, chunk, flags, nsteps):
    # the plain FP64 form, several steps per pass (8 waves per work-group):
    # the reference's velocity-Verlet bit for bit, ker included
    state = random_state(3, N, seed=13, ker=True)
    st = _oracle_run(3, N, state, nsteps)
    lat = axs.Lattice(3, N)
    lat.set_kernel("march", rows=rows, chunk=chunk, depth=depth, flags=flags)
    assert lat.kernel_config()["depth"] == depth
    assert lat.bytes_per_site()[0] == 96.0 / depth
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    _assert_same(lat, st)


@pytest.mark.parametrize("N,depth,chunk,flags,nsteps", [
    (64, 2, 0, 3, 8), (96, 4, 0, 2, 11), (128, 6, 32, 3, 13), (100, 3, 0, 11, 7)])
def test_2d_fp64_multi_step_bit_exact_vs_oracle(N, depth, chunk, flags, nsteps):
    state = random_state(2, N, seed=14, ker=True)
    st = _oracle_run(2, N, state, nsteps)
    lat = axs.Lattice(2, N)
    lat.set_kernel("march", chunk=chunk, depth=depth, flags=flags)
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    _assert_same(lat, st)


@pytest.mark.parametrize("Ns,depth", [([128], 3), ([192], 4), ([256], 4), ([128] * 4 + [192] * 4, 4),
                                      ([1024], 4)])
def test_fp32_default_depth_follows_the_work(Ns, depth):
    # four steps per pass when the four-step tiles still fill the chip in
    # chunks of 32 planes or more (192^3 and up, ensembles), else three
    lat = axs.Lattice(3, members=Ns, precision="fp32")
    assert lat.kernel_config()["depth"] == depth


@pytest.mark.parametrize("Ns,precision,depth", [([4096], "fp32", 6), ([8192], "fp32", 8),
                                                ([8192], "fp64", 6), ([1024] * 32, "fp32", 8)])
def test_2d_default_depth(Ns, precision, depth):
    # FP32 2D contexts of 2^25 sites and more start at eight steps per pass
    lat = axs.Lattice(2, members=Ns, precision=precision)
    assert lat.kernel_config()["depth"] == depth

