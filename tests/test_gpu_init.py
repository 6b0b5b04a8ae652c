"""Initial conditions generated in device memory (axs_init_member): the k^-1
field of the benchmark configurations and the thermal spectrum of
gaussian_thermal (init.cpp:33-200).  The device generator draws the same
Philox4x32-10 normals per Fourier mode as the host generator (pinned by the
published known-answer vectors in test_host.py) and inverse-transforms with
rocFFT, so its fields equal the host ones up to FFT and libm rounding; the
FP32 form holds the FP64 field rounded once."""
import numpy as np
import pytest

import axs

pytestmark = pytest.mark.gpu


def _rel_max(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("mode", ["kinv", "thermal"])
@pytest.mark.parametrize("nd,N", [(2, 128), (3, 32), (3, 48)])
def test_device_init_matches_host(mode, nd, N):
    lat = axs.Lattice(nd, N, precision="fp64", T_initial=4.0)
    lat.init(mode, seed=1998)
    g = lat.get_state()
    h = axs.host_init(mode, nd, N, space_step=0.5, lam=1.0, T_initial=4.0, seed=1998)
    for k, ref in zip(("phi1", "phi2", "phidot1", "phidot2"), h):
        if np.all(ref == 0):
            assert np.all(g[k] == 0), k
        else:
            assert _rel_max(g[k], ref) < 1e-12, (k, _rel_max(g[k], ref))
    assert np.all(g["ker1"] == 0) and np.all(g["ker2"] == 0)      # s_run.cpp:44-47
    if mode == "kinv":
        assert abs(np.mean(g["phi1"] ** 2 + g["phi2"] ** 2) - 1.0) < 1e-12


@pytest.mark.parametrize("mode", ["kinv", "thermal"])
def test_fp32_init_is_the_fp64_field_rounded_once(mode):
    nd, N = 3, 40
    a = axs.Lattice(nd, N, precision="fp64")
    b = axs.Lattice(nd, N, precision="fp32")
    a.init(mode, seed=7)
    b.init(mode, seed=7)
    ga, gb = a.get_state(fields=("phi1", "phi2")), b.get_state(fields=("phi1", "phi2"))
    for k in ga:
        np.testing.assert_array_equal(gb[k], ga[k].astype(np.float32).astype(np.float64))


def test_ensemble_members_seeded_independently():
    Ns = [64, 32, 64]
    lat = axs.Lattice(3, members=Ns)
    for r in range(len(Ns)):
        lat.init("kinv", seed=1998 + r, member=r)
    for r, N in enumerate(Ns):
        single = axs.Lattice(3, N)
        single.init("kinv", seed=1998 + r)
        np.testing.assert_array_equal(lat.get_state(member=r)["phi1"], single.get_state()["phi1"])
    assert not np.allclose(lat.get_state(member=0)["phi1"], lat.get_state(member=2)["phi1"])
