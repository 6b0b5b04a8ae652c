"""GPU parity of the sampled quantities: string finder (bit-exact set and
count, string_finding.cpp:3-86), core count / xi (s_run.cpp:96-132),
background sums (s_run.cpp:135-150), energy, power spectrum
(pk_estimation.cpp:37-214)."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import kinv_field, random_state, vortex2d

pytestmark = pytest.mark.gpu


def _evolved(nd, N, nsteps, precision="fp64", seed=1998, **kw):
    p1, p2 = kinv_field(nd, N, seed=seed)
    z = np.zeros(N ** nd)
    lat = axs.Lattice(nd, N, precision=precision, **kw)
    lat.set_state(p1, p2, z, z, tau=1.0)
    lat.step(nsteps)
    return lat


@pytest.mark.parametrize("nd,N,nsteps", [(2, 256, 0), (2, 256, 50), (2, 256, 250), (2, 256, 499),
                                         (3, 64, 0), (3, 64, 40), (3, 96, 100), (3, 288, 2)])
def test_strings_bit_exact(nd, N, nsteps):
    lat = _evolved(nd, N, nsteps)
    g = lat.get_state()
    n_ref, c_ref = O.cores(O.axion(g["phi1"], g["phi2"]), nd, N, thr=1)
    n, c = lat.strings()
    assert n == n_ref
    np.testing.assert_array_equal(c, c_ref)
    npl, ncores = lat.string_counts()
    assert npl[0] == n_ref
    if nd == 2:
        assert ncores[0] == O.num_cores2(c_ref)
    else:
        assert ncores[0] == n_ref
    print(f"{nd}D N={N} step {nsteps}: {n} flagged sites")


@pytest.mark.parametrize("N,prec", [(8, "fp64"), (30, "fp64"), (32, "fp32"), (34, "fp64"),
                                    (64, "fp32"), (66, "fp32"), (100, "fp64"), (130, "fp32")])
def test_strings_3d_sides(N, prec):
    """The 3D one-pass finder on random phases (dense strings) at sides below,
    across and off its 64-column and 30-row tiles: the oracle's set in the
    oracle's order, in both storage forms (the FP32 form against the oracle
    on its own widened state)."""
    rng = np.random.default_rng(N)
    p1, p2 = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3)
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N, precision=prec)
    lat.set_state(p1, p2, z, z, tau=1.0)
    g = lat.get_state()
    n_ref, c_ref = O.cores(O.axion(g["phi1"], g["phi2"]), 3, N, thr=1)
    n, c = lat.strings()
    assert n == n_ref and n > 0
    np.testing.assert_array_equal(c, c_ref)
    assert lat.string_counts()[0][0] == n_ref


@pytest.mark.parametrize("N,prec", [(64, "fp64"), (66, "fp64"), (64, "fp32"), (66, "fp32")])
def test_strings_3d_threshold_straddle(N, prec):
    """Every plaquette sum at the acceptance threshold: the phase turns by
    accept (float 0.495, thr = 1) of a circle per site along k, plus noise
    of 1e-9 turns per site, so plaquettes b and c of every site sum to
    accept within 1e-8 (with a slight upward bias from the two short edges):
    most sites are flagged and some are not.  The one-pass finder's float
    screen cannot decide any of them: each site goes through its FP64
    fallback, which must give the oracle's set."""
    rng = np.random.default_rng(5)
    k = np.arange(N ** 3) % N
    accept = float(np.float32(0.495))
    turns = accept * k + 1e-9 * rng.standard_normal(N ** 3)
    p1, p2 = np.cos(2 * np.pi * turns), np.sin(2 * np.pi * turns)
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N, precision=prec)
    lat.set_state(p1, p2, z, z, tau=1.0)
    g = lat.get_state()
    n_ref, c_ref = O.cores(O.axion(g["phi1"], g["phi2"]), 3, N, thr=1)
    n, c = lat.strings()
    assert 0.05 * N ** 3 < n_ref < 0.99 * N ** 3
    assert n == n_ref
    np.testing.assert_array_equal(c, c_ref)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_strings_3d_zero_tiny_and_nan_inputs(prec):
    """Sites whose float angle is not trusted (exact zeros of either sign,
    components below 1e-30, NaN in FP64) take the FP64 fallback: the set
    is still the oracle's."""
    N = 64
    rng = np.random.default_rng(11)
    p1, p2 = kinv_field(3, N, seed=3)
    L = N ** 3
    pick = rng.choice(L, 4000, replace=False)
    p1[pick[:1000]], p2[pick[:1000]] = 0.0, 0.0
    p1[pick[1000:2000]], p2[pick[1000:2000]] = -0.0, -0.0
    p1[pick[2000:3000]] = 1e-40 * rng.standard_normal(1000)
    p2[pick[2000:3000]] = 1e-40 * rng.standard_normal(1000)
    if prec == "fp64":
        p1[pick[3000:3100]] = np.nan
        p1[pick[3100:3200]], p2[pick[3100:3200]] = 1e-300, -3e-310
    z = np.zeros(L)
    lat = axs.Lattice(3, N, precision=prec)
    lat.set_state(p1, p2, z, z, tau=1.0)
    g = lat.get_state()
    n_ref, c_ref = O.cores(O.axion(g["phi1"], g["phi2"]), 3, N, thr=1)
    n, c = lat.strings()
    assert n == n_ref and n > 0
    np.testing.assert_array_equal(c, c_ref)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("N", [64, 40])
def test_strings_3d_straight_lines(axis, N):
    """A straight string along each axis (winding in the plane of the other
    two), which exercises the three plaquettes separately."""
    idx = np.indices((N, N, N), dtype=np.float64)
    u, v = [idx[a] for a in range(3) if a != axis]
    phase = np.arctan2(u - 10.5, v - 17.5)
    p1, p2 = np.cos(phase).ravel(), np.sin(phase).ravel()
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N)
    lat.set_state(p1, p2, z, z, tau=1.0)
    n_ref, c_ref = O.cores(O.axion(p1, p2), 3, N, thr=1)
    n, c = lat.strings()
    assert n == n_ref and n >= N
    np.testing.assert_array_equal(c, c_ref)


@pytest.mark.parametrize("nd,Ns,prec", [(2, [64, 66, 130, 256, 64], "fp64"), (2, [8], "fp64"),
                                        (2, [8, 66, 4, 2048], "fp64"), (2, [130, 64, 96], "fp32"),
                                        (3, [64, 34, 96], "fp64"), (3, [66, 64], "fp32")])
def test_string_counts_ensemble(nd, Ns, prec):
    """axs_string_counts over an ensemble of uneven sides, dense (random
    phases) and smooth members: flagged sites and, in 2D, the core count
    n - (list-adjacent pairs) taken from the ballot words, against the
    oracle's list of each member (on the member's own, possibly
    float-rounded, state)."""
    lat = axs.Lattice(nd, members=Ns, precision=prec)
    fields = []
    for r, N in enumerate(Ns):
        if r % 2 == 0:
            rng = np.random.default_rng(100 + r)
            p1, p2 = rng.standard_normal(N ** nd), rng.standard_normal(N ** nd)
        else:
            p1, p2 = kinv_field(nd, N, seed=200 + r)
        z = np.zeros(N ** nd)
        lat.set_state(p1, p2, z, z, member=r)
        g = lat.get_state(member=r, fields=("phi1", "phi2"))
        fields.append((g["phi1"], g["phi2"]))
    npl, ncores = lat.string_counts()
    for r, (N, (p1, p2)) in enumerate(zip(Ns, fields)):
        n_ref, c_ref = O.cores(O.axion(p1, p2), nd, N, thr=1)
        assert npl[r] == n_ref, (r, N)
        assert ncores[r] == (O.num_cores2(c_ref) if nd == 2 else n_ref), (r, N)
    assert npl.sum() > 0


@pytest.mark.parametrize("nd,Ns,binning,prec", [(2, [64, 128, 64, 96, 128], "log", "fp64"),
                                                (2, [64, 128, 64], "linear", "fp32"),
                                                (3, [32, 48, 32], "log", "fp32"),
                                                (3, [32, 48, 32, 32], "linear", "fp64")])
def test_powerspec_all(nd, Ns, binning, prec):
    """Every member's spectrum from one call (batched transforms per side)
    against the oracle on the member's own state, and against the
    one-member call."""
    lat = axs.Lattice(nd, members=Ns, precision=prec)
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(nd, N, seed=300 + r)
        z = np.zeros(N ** nd)
        lat.set_state(p1, p2, z, z, member=r)
    out = lat.powerspec_all(binning=binning, nbins=20)
    assert len(out) == len(Ns)
    for r, N in enumerate(Ns):
        ks, pk, cnt = out[r]
        phi2 = lat.get_state(member=r)["phi2"]
        if binning == "log":
            ks_r, pk_r, cnt_r = O.powerspec_log(phi2, nd, N, 0.5, 20)
        else:
            ks_r, pk_r, cnt_r = O.powerspec(phi2, nd, N, 0.5)
        np.testing.assert_array_equal(cnt, cnt_r)
        sel = cnt > 0
        np.testing.assert_allclose(ks[sel], ks_r[sel], rtol=1e-13)
        np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max())
        ks1, pk1, cnt1 = lat.powerspec(member=r, binning=binning, nbins=20)
        np.testing.assert_array_equal(cnt, cnt1)
        np.testing.assert_allclose(pk, pk1, rtol=1e-12, atol=1e-14 * pk1.max())


@pytest.mark.parametrize("nd,Ns", [(3, [96, 64]), (3, [64, 130]), (2, [256, 64]), (2, [66])])
def test_string_list_after_counts(nd, Ns):
    """A coordinate list right after axs_string_counts reuses the last
    member's search (only the offsets are computed) and searches any other
    member afresh: every list is the oracle's."""
    lat = axs.Lattice(nd, members=Ns)
    fields = []
    for r, N in enumerate(Ns):
        rng = np.random.default_rng(40 + r)
        p1, p2 = rng.standard_normal(N ** nd), rng.standard_normal(N ** nd)
        z = np.zeros(N ** nd)
        lat.set_state(p1, p2, z, z, member=r)
        fields.append((p1, p2))
    npl, _ = lat.string_counts()
    for r in [len(Ns) - 1] + list(range(len(Ns) - 1)):
        N = Ns[r]
        n_ref, c_ref = O.cores(O.axion(*fields[r]), nd, N, thr=1)
        n, c = lat.strings(member=r)
        assert n == n_ref == npl[r]
        np.testing.assert_array_equal(c, c_ref)


def test_strings_hand_built_vortices():
    N = 64
    p1, p2 = vortex2d(N, [(10.5, 12.5), (40.5, 30.5), (20.5, 50.5)], [1, -1, 1])
    z = np.zeros(N * N)
    lat = axs.Lattice(2, N)
    lat.set_state(p1, p2, z, z, tau=1.0)
    n_ref, c_ref = O.cores(O.axion(p1, p2), 2, N)
    n, c = lat.strings()
    assert n == n_ref and n > 0
    np.testing.assert_array_equal(c, c_ref)


def test_strings_fp32_form_count_close():
    nd, N, nsteps = 3, 96, 150
    f64 = _evolved(nd, N, nsteps, "fp64").strings(coords=False)[0]
    f32 = _evolved(nd, N, nsteps, "fp32").strings(coords=False)[0]
    print(f"string count fp64 {f64} fp32 {f32}")
    assert abs(f32 - f64) <= 0.01 * f64 + 2


def test_strings_capacity_error():
    lat = _evolved(2, 128, 20)
    n, _ = lat.strings(coords=False)
    assert n > 2
    buf = np.zeros((1, 2), np.int32)
    cnt = axs.ctypes.c_int64()
    rc = lat.lib.axs_member_strings(lat._ctx, 0, buf.ctypes.data_as(axs.POINTER(axs.c_int32)), 1,
                                    axs.ctypes.byref(cnt))
    assert rc == axs.AXS_ERR_CAPACITY and cnt.value == n


@pytest.mark.parametrize("nd,N,nsteps,prec", [(2, 256, 100, "fp64"), (3, 64, 50, "fp64"),
                                              (3, 64, 50, "fp32")])
def test_background(nd, N, nsteps, prec):
    lat = _evolved(nd, N, nsteps, prec)
    g = lat.get_state()
    ref = O.background(g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    got = lat.background()
    scale = np.array([np.abs(g["phi1"]).sum(), np.abs(g["phi2"]).sum(),
                      np.abs(g["phidot1"]).sum(), np.abs(g["phidot2"]).sum(), ref[4], ref[5]])
    # same fields in both (the FP32 state read back as doubles): sums agree
    # to FP64 rounding of the sum of magnitudes
    tol = 1e-12 if prec == "fp64" else 1e-6
    assert np.all(np.abs(got - ref) <= tol * scale + 1e-300), (got, ref)


@pytest.mark.parametrize("nd,N,nsteps,prec", [(2, 128, 60, "fp64"), (3, 48, 30, "fp64"),
                                              (3, 48, 30, "fp32")])
def test_energy(nd, N, nsteps, prec):
    lat = _evolved(nd, N, nsteps, prec)
    g = lat.get_state()
    p = O.params(nd, N)
    ref = O.energy(p, lat.tau, g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    got = lat.energy()
    tol = 1e-11 if prec == "fp64" else 1e-5
    np.testing.assert_allclose(got, ref, rtol=tol)


@pytest.mark.parametrize("nd,N,stencil", [(3, 48, 0), (3, 40, 2), (3, 34, 3), (2, 128, 0),
                                          (2, 96, 1), (2, 66, 3)])
def test_fp32_sums_before_read_back(nd, N, stencil):
    """FP32 form: background and energy taken straight after steps rebuild
    phidot inside the summing pass (ker recomputed from phi); taken after a
    read-back they read ker from the read-back buffer.  Both must be bit for
    bit the same, match the oracle's sums over the read-back fields, and a
    second call (served from the per-state cache) must return them again."""
    def evolved():
        lat = axs.Lattice(nd, N, precision="fp32", stencil_setting=stencil)
        p1, p2, v1, v2, _, _ = random_state(nd, N, seed=5)
        lat.set_state(p1, p2, v1, v2, tau=1.0)
        lat.step(7)
        return lat
    a = evolved()
    bg_a, e_a = a.background(), a.energy()
    np.testing.assert_array_equal(a.background(), bg_a)
    np.testing.assert_array_equal(a.energy(), e_a)
    b = evolved()
    g = b.get_state()
    np.testing.assert_array_equal(b.background(), bg_a)
    np.testing.assert_array_equal(b.energy(), e_a)
    ref = O.background(g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    scale = np.array([np.abs(g[k]).sum() for k in ("phi1", "phi2", "phidot1", "phidot2")] +
                     [ref[4], ref[5]])
    assert np.all(np.abs(bg_a - ref) <= 1e-12 * scale + 1e-300), (bg_a, ref)
    e_ref = O.energy(O.params(nd, N, stencil_setting=stencil), a.tau, g["phi1"], g["phi2"],
                     g["phidot1"], g["phidot2"])
    np.testing.assert_allclose(e_a, e_ref, rtol=1e-11)


def test_fp32_sums_follow_tau_and_state():
    # the cached sums are keyed by state and tau: a new tau (set_tau) or new
    # state (set_state with an uploaded ker) gives the oracle's new values
    nd, N = 3, 32
    lat = axs.Lattice(nd, N, precision="fp32")
    p1, p2, v1, v2, k1, k2 = random_state(nd, N, seed=8, ker=True)
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    lat.step(3)
    e0 = lat.energy()
    lat.tau = 2.5
    e1 = lat.energy()
    g = lat.get_state()
    ref = O.energy(O.params(nd, N), 2.5, g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    np.testing.assert_allclose(e1, ref, rtol=1e-11)
    assert e1[2] != e0[2]
    lat.set_state(p1, p2, v1, v2, ker1=k1, ker2=k2, tau=1.0)
    g = lat.get_state()
    ref = O.background(g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    np.testing.assert_allclose(lat.background(), ref, rtol=1e-12, atol=1e-9)
    ref = O.energy(O.params(nd, N), 1.0, g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    np.testing.assert_allclose(lat.energy(), ref, rtol=1e-11)


def test_energy_free_field_conserved():
    # lambda = 0, T0 = 0: H = K + G is conserved by velocity-Verlet up to a
    # bounded O((w dt)^2) oscillation.  With smooth data (modes |n| < N/8,
    # w dt <= 0.16) the energy stays within 1e-2 over 800 steps and does not
    # drift: the change over the second 400 steps is no larger than the first.
    nd, N = 3, 32
    p1, p2 = kinv_field(nd, N, seed=21)
    v1, v2 = kinv_field(nd, N, seed=22)
    lat = axs.Lattice(nd, N, lam=0.0, T_initial=0.0)
    lat.set_state(p1, p2, 0.5 * v1, 0.5 * v2, tau=1.0)
    lat.step(1)
    e0 = lat.energy()
    lat.step(400)
    e1 = lat.energy()
    lat.step(400)
    e2 = lat.energy()
    h0, h1, h2 = e0[0] + e0[1], e1[0] + e1[1], e2[0] + e2[1]
    assert abs(h1 - h0) <= 1e-2 * h0 and abs(h2 - h0) <= 1e-2 * h0
    assert e1[2] == 0.0 and e2[2] == 0.0


@pytest.mark.parametrize("nd,N,nsteps,prec", [(2, 256, 50, "fp64"), (3, 64, 20, "fp64"),
                                              (3, 48, 10, "fp64"), (3, 64, 20, "fp32"),
                                              (2, 2048, 5, "fp32")])
def test_powerspec_linear_bins(nd, N, nsteps, prec):
    """Row-order binning (bin tables that fit the work-group's memory) and,
    for the 1024 linear bins of 2048^2, the per-bin binning."""
    lat = _evolved(nd, N, nsteps, prec)
    g = lat.get_state()
    ks_r, pk_r, cnt_r = O.powerspec(g["phi2"], nd, N, 0.5)
    ks, pk, cnt = lat.powerspec(binning="linear")
    np.testing.assert_array_equal(cnt, cnt_r)
    sel = cnt > 0
    np.testing.assert_allclose(ks[sel], ks_r[sel], rtol=1e-13)
    np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max())


@pytest.mark.parametrize("nd,N,nbins", [(2, 256, 64), (3, 64, 64), (3, 48, 20), (2, 256, 500),
                                         (2, 256, 700), (3, 64, 1000)])
def test_powerspec_log_bins(nd, N, nbins):
    """64 and 20 log bins; 500 (row-order binning, most bins narrower than a
    unit shell); 700 and 1000 (too many bins for the row-order binning's
    work-group memory: the per-bin binning)."""
    lat = _evolved(nd, N, 10)
    g = lat.get_state()
    ks_r, pk_r, cnt_r = O.powerspec_log(g["phi2"], nd, N, 0.5, nbins)
    ks, pk, cnt = lat.powerspec(binning="log", nbins=nbins)
    np.testing.assert_array_equal(cnt, cnt_r)
    sel = cnt > 0
    np.testing.assert_allclose(ks[sel], ks_r[sel], rtol=1e-13)
    np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max())


@pytest.mark.parametrize("nd,N,binning", [(3, 64, "log"), (3, 64, "linear"), (2, 512, "log")])
def test_powerspec_repeatable(nd, N, binning):
    """Same state, same bits: the first call also builds the bin table's
    |k| sums and counts, later calls take them from the context."""
    lat = _evolved(nd, N, 10)
    a = lat.powerspec(binning=binning, nbins=64)
    for _ in range(2):
        b = lat.powerspec(binning=binning, nbins=64)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("nd,N,prec,binning", [(3, 128, "fp64", "log"), (3, 128, "fp32", "linear"),
                                               (3, 256, "fp64", "linear"), (2, 128, "fp32", "log"),
                                               (2, 1024, "fp64", "linear"),
                                               (2, 2048, "fp64", "log"), (2, 4096, "fp32", "log")])
def test_powerspec_own_transform(nd, N, prec, binning, monkeypatch):
    """Power-of-two sides from 128 take the transform of our own (passes that
    read phi from the state and bin in the last pass): against the oracle on
    the read-back field, and against the rocFFT path on the same state."""
    lat = _evolved(nd, N, 5, prec)
    g = lat.get_state()
    st = lat.get_state(fields=("phi1", "phi2", "phidot1", "phidot2"))
    if binning == "log":
        ks_r, pk_r, cnt_r = O.powerspec_log(g["phi2"], nd, N, 0.5, 64)
    else:
        ks_r, pk_r, cnt_r = O.powerspec(g["phi2"], nd, N, 0.5)
    ks, pk, cnt = lat.powerspec(binning=binning, nbins=64)
    np.testing.assert_array_equal(cnt, cnt_r)
    sel = cnt > 0
    # the oracle sums |k| plainly over up to ~2e5 modes per bin (as the
    # reference does): its own rounding reaches a few 1e-13 here
    np.testing.assert_allclose(ks[sel], ks_r[sel], rtol=1e-11)
    np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max())
    monkeypatch.setenv("AXS_PK_ROCFFT", "1")
    ref = axs.Lattice(nd, N, precision=prec)
    ref.set_state(st["phi1"], st["phi2"], st["phidot1"], st["phidot2"], tau=lat.tau)
    ks2, pk2, cnt2 = ref.powerspec(binning=binning, nbins=64)
    np.testing.assert_array_equal(cnt, cnt2)
    np.testing.assert_array_equal(ks, ks2)
    np.testing.assert_allclose(pk[sel], pk2[sel], rtol=1e-11, atol=1e-14 * pk2[sel].max())


@pytest.mark.parametrize("field,prec", [("phi", "fp32"), ("phi1", "fp32"), ("phidot", "fp32"),
                                        ("phidot2", "fp64")])
def test_powerspec_own_transform_fields(field, prec):
    """Complex-field spectra (two real transforms in one batch), phi1 and the
    velocities (the FP32 form's phidot read-back, the FP64 form's stored
    pairs) through the own transform, against the oracle on the read-back
    fields."""
    lat = _evolved(3, 128, 5, prec)
    g = lat.get_state(fields=("phi1", "phi2", "phidot1", "phidot2"))
    re, im = {"phi": ("phi1", "phi2"), "phi1": ("phi1", None), "phidot": ("phidot1", "phidot2"),
              "phidot2": ("phidot2", None)}[field]
    ks_r, pk_r, cnt_r = O.powerspec_log(g[re], 3, 128, 0.5, 64,
                                        field_imag=None if im is None else g[im])
    ks, pk, cnt = lat.powerspec(binning="log", nbins=64, field=field)
    np.testing.assert_array_equal(cnt, cnt_r)
    sel = cnt > 0
    np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max())


@pytest.mark.parametrize("N,prec", [(1024, "fp32"), (512, "fp64")])
def test_powerspec_own_transform_large(N, prec, monkeypatch):
    """The own transform at the benchmark's 1024^3 (and 512^3), where no
    oracle transform fits the test's time: against rocFFT on the same device
    field, log and linear bins."""
    lat = axs.Lattice(3, N, precision=prec)
    lat.init("kinv", seed=7)
    own = [lat.powerspec(binning="log", nbins=64), lat.powerspec(binning="linear")]
    lat.close()
    monkeypatch.setenv("AXS_PK_ROCFFT", "1")
    ref = axs.Lattice(3, N, precision=prec)
    ref.init("kinv", seed=7)
    rf = [ref.powerspec(binning="log", nbins=64), ref.powerspec(binning="linear")]
    ref.close()
    for (ks, pk, cnt), (ks2, pk2, cnt2) in zip(own, rf):
        np.testing.assert_array_equal(cnt, cnt2)
        np.testing.assert_array_equal(ks, ks2)
        sel = cnt > 0
        np.testing.assert_allclose(pk[sel], pk2[sel], rtol=1e-11, atol=1e-14 * pk2[sel].max())
