"""Asynchronous snapshots (axs_snapshot_begin / axs_snapshot_end): the copy
holds exactly the state at the time it was begun, stepping goes on during
the transfer without being perturbed, and the end may run on another
thread."""
import threading

import numpy as np
import pytest

import axs
from fields import random_state

pytestmark = pytest.mark.gpu

ALL = ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2")


@pytest.mark.parametrize("nd,N,precision", [(2, 128, "fp64"), (3, 64, "fp64"), (3, 64, "fp32"),
                                            (3, 96, "fp32")])
def test_snapshot_is_the_state_at_begin(nd, N, precision):
    state = random_state(nd, N, seed=61, ker=True)
    a = axs.Lattice(nd, N, precision=precision)
    a.set_state(*state, tau=1.0)
    b = axs.Lattice(nd, N, precision=precision)
    b.set_state(*state, tau=1.0)
    a.step(10)
    want = a.get_state()
    tau0 = a.tau
    a.snapshot_begin()
    a.step(25)                       # stepping continues while the copy runs
    got, tau = a.snapshot_end()
    assert tau == tau0
    for k in ALL:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    b.step(35)                       # the snapshot did not perturb the run
    ga, gb = a.get_state(), b.get_state()
    for k in ALL:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_partial_snapshot_and_end_on_another_thread():
    # the driver's use: the first N^2 values of phi (the reference's 3D
    # field snapshot), finished by a writer thread while the main thread
    # steps
    N = 64
    state = random_state(3, N, seed=62, ker=True)
    lat = axs.Lattice(3, N)
    lat.set_state(*state, tau=1.0)
    lat.step(4)
    want = lat.get_state(fields=("phi1", "phi2"))
    lat.snapshot_begin(fields=("phi1", "phi2"), count=N * N)
    box = {}
    th = threading.Thread(target=lambda: box.update(r=lat.snapshot_end()))
    th.start()
    lat.step(30)
    th.join()
    got, _ = box["r"]
    np.testing.assert_array_equal(got["phi1"], want["phi1"][:N * N])
    np.testing.assert_array_equal(got["phi2"], want["phi2"][:N * N])


def test_repeated_snapshots_reuse_buffers():
    N = 64
    lat = axs.Lattice(2, N)
    lat.set_state(*random_state(2, N, seed=63, ker=True), tau=1.0)
    for n in range(4):
        want = lat.get_state()
        lat.snapshot_begin(fields=("phi1", "phi2", "ker1", "ker2"), count=(n + 1) * 512)
        lat.step(3)
        got, _ = lat.snapshot_end()
        for k in ("phi1", "phi2", "ker1", "ker2"):
            np.testing.assert_array_equal(got[k], want[k][:(n + 1) * 512])


def test_snapshot_errors():
    N = 32
    lat = axs.Lattice(2, N)
    with pytest.raises(axs.AxsError, match="no state"):
        lat.snapshot_begin()
    lat.set_state(*random_state(2, N, seed=64, ker=True), tau=1.0)
    with pytest.raises(axs.AxsError, match="no snapshot in flight"):
        lat.snapshot_end()
    lat.snapshot_begin()
    with pytest.raises(axs.AxsError, match="already in flight"):
        lat.snapshot_begin()
    lat.snapshot_end()
    with pytest.raises(axs.AxsError, match="count"):
        lat.snapshot_begin(count=N * N + 1)
    with pytest.raises(axs.AxsError, match="fields"):
        lat.snapshot_begin(fields=())
