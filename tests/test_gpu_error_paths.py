"""Error paths of axs_step and the FP32 form's derived buffers.

A failure point (the environment variable AXS_STEP_FAIL_AFTER, read when a
context is created) makes axs_step report an error in place of a launch;
every kernel that does run exits normally.  After a direct loop fails
partway, the launched steps stand and every sampling call must describe
that new state; after a failed graph launch, the state and tau are those
before the call.  In the FP32 form a ker rebuilt from phi belongs to the
tau it was computed at, so sampling after axs_set_tau must not depend on
which calls came before."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import random_state

pytestmark = pytest.mark.gpu


def _sample(lat, nd):
    g = lat.get_state()
    n, c = lat.strings()
    return {"state": g, "n": n, "c": c, "counts": lat.string_counts(),
            "energy": lat.energy(), "background": lat.background(), "tau": lat.tau}


def _assert_same(a, b):
    for k in a["state"]:
        np.testing.assert_array_equal(a["state"][k], b["state"][k], err_msg=k)
    assert a["n"] == b["n"]
    np.testing.assert_array_equal(a["c"], b["c"])
    np.testing.assert_array_equal(a["counts"][0], b["counts"][0])
    np.testing.assert_array_equal(a["counts"][1], b["counts"][1])
    np.testing.assert_array_equal(a["energy"], b["energy"])
    np.testing.assert_array_equal(a["background"], b["background"])
    assert a["tau"] == b["tau"]


@pytest.mark.parametrize("nd,N,prec", [(3, 64, "fp32"), (3, 64, "fp64"), (2, 128, "fp32"),
                                       (2, 96, "fp64")])
def test_failed_direct_step_keeps_launched_steps(monkeypatch, nd, N, prec):
    """Two launches succeed, the third fails: tau and the state are those of
    the launched steps, and strings, counts, energy, background and the
    read-back equal a context that stepped that far without failing -- the
    caches filled before the call (sums, string search, read-back ker and
    phidot) are not served for the new state."""
    p1, p2, v1, v2, _, _ = random_state(nd, N, seed=31)
    monkeypatch.setenv("AXS_STEP_FAIL_AFTER", "2")
    bad = axs.Lattice(nd, N, precision=prec)
    monkeypatch.delenv("AXS_STEP_FAIL_AFTER")
    bad.set_state(p1, p2, v1, v2, tau=1.0)
    _sample(bad, nd)                     # fill every cache on the first state
    depth = bad.kernel_config()["depth"]
    with pytest.raises(axs.AxsError) as ei:
        bad.step(10 * depth)
    assert ei.value.code == axs.AXS_ERR_HIP
    done = int(round((bad.tau - 1.0) / 0.1))
    assert done == 2 * depth
    good = axs.Lattice(nd, N, precision=prec)
    good.set_state(p1, p2, v1, v2, tau=1.0)
    good.step(done)
    _assert_same(_sample(bad, nd), _sample(good, nd))
    # and against the oracle on the read-back state
    g = bad.get_state()
    n_ref, c_ref = O.cores(O.axion(g["phi1"], g["phi2"]), nd, N)
    n, c = bad.strings()
    assert n == n_ref
    np.testing.assert_array_equal(c, c_ref)
    ref = O.energy(O.params(nd, N), bad.tau, g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    np.testing.assert_allclose(bad.energy(), ref, rtol=1e-11)


@pytest.mark.parametrize("nd,N,prec", [(3, 64, "fp32"), (2, 128, "fp64")])
def test_failed_graph_launch_restores_state(monkeypatch, nd, N, prec):
    """With AXS_STEP_GRAPH the loop is captured and launched as one graph; a
    failed graph launch leaves tau, the state and every sampled value as
    they were before the call."""
    p1, p2, v1, v2, _, _ = random_state(nd, N, seed=32)
    monkeypatch.setenv("AXS_STEP_FAIL_AFTER", "0")
    lat = axs.Lattice(nd, N, precision=prec)
    monkeypatch.delenv("AXS_STEP_FAIL_AFTER")
    cfg = lat.kernel_config()
    lat.set_kernel(kernel="auto", rows=cfg["rows"], chunk=cfg["chunk"], depth=cfg["depth"],
                   flags=cfg["flags"] | axs.AXS_STEP_GRAPH)
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    lat.step(0)
    before = _sample(lat, nd)
    with pytest.raises(axs.AxsError):
        lat.step(4 * cfg["depth"])
    _assert_same(_sample(lat, nd), before)


def test_fp32_sampling_after_set_tau_is_history_free():
    """A read-back rebuilds ker at the current tau; after axs_set_tau the
    energy, background and read-back must be those of the new tau, equal to
    a context that never read back, and to the oracle on that read-back."""
    nd, N = 3, 32
    p1, p2, v1, v2, _, _ = random_state(nd, N, seed=33)
    out = []
    for readback_first in (True, False):
        lat = axs.Lattice(nd, N, precision="fp32")
        lat.set_state(p1, p2, v1, v2, tau=1.0)
        lat.step(3)
        if readback_first:
            lat.get_state()              # ker and phidot rebuilt at tau = 1.3
            lat.snapshot_begin()
            lat.snapshot_end()
        lat.tau = 2.5
        e, b = lat.energy(), lat.background()
        g = lat.get_state()
        lat.snapshot_begin()
        snap, tau_s = lat.snapshot_end()
        assert tau_s == 2.5
        for k in g:
            np.testing.assert_array_equal(snap[k], g[k])
        out.append((e, b, g))
    (e0, b0, g0), (e1, b1, g1) = out
    np.testing.assert_array_equal(e0, e1)
    np.testing.assert_array_equal(b0, b1)
    for k in g0:
        np.testing.assert_array_equal(g0[k], g1[k], err_msg=k)
    ref = O.energy(O.params(nd, N), 2.5, g0["phi1"], g0["phi2"], g0["phidot1"], g0["phidot2"])
    np.testing.assert_allclose(e0, ref, rtol=1e-11)
    # the rebuilt ker follows tau: it differs from the one at tau = 1.3
    lat = axs.Lattice(nd, N, precision="fp32")
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    lat.step(3)
    k13 = lat.get_state()["ker1"]
    lat.tau = 2.5
    assert not np.array_equal(lat.get_state()["ker1"], k13)
    np.testing.assert_array_equal(lat.get_state()["ker1"], g0["ker1"])


def test_fp32_uploaded_ker_is_kept_across_set_tau():
    """An uploaded ker (the reference's all_data ker arrays) is part of the
    state, not a function of tau: set_tau keeps it and the phidot read back."""
    nd, N = 3, 32
    p1, p2, v1, v2, k1, k2 = random_state(nd, N, seed=34, ker=True)
    lat = axs.Lattice(nd, N, precision="fp32")
    lat.set_state(p1, p2, v1, v2, k1, k2, tau=1.0)
    g0 = lat.get_state()
    lat.tau = 3.0
    g1 = lat.get_state()
    for k in g0:
        np.testing.assert_array_equal(g0[k], g1[k], err_msg=k)
    np.testing.assert_array_equal(g1["ker1"], k1.astype(np.float32).astype(np.float64))
