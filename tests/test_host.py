"""Host driver pieces that run without a GPU: parameter-file reader
(read_parameters.cpp:182-238), step count (s_run.cpp:82), xi (s_run.cpp:
111-131), host FFT and the host initial conditions (init.cpp:33-200)."""
import os
import subprocess

import numpy as np
import pytest

import axs
import oracle as O


def test_default_param_parses(root):
    # the reference's own parameter file
    p = axs.read_params(os.path.join(root, "tests", "golden", "default.param"))
    assert p["lambda"] == 1.0 and p["enable_PRS"] == 1 and p["T_initial"] == 4.0
    assert p["NDIMS"] == 2 and p["N"] == 128 and p["space_step"] == 1.0
    assert p["time_step"] == pytest.approx(float(np.float32(0.3)), abs=0)
    assert p["stencil_setting"] == 0 and p["seed"] == 1998 and p["thr"] == 1
    assert p["n_samples"] == 40 and p["n_snapshots"] == 5 and p["tau_initial"] == 1.0
    assert p["refinement_threshold"] == 20.0
    assert p["output_directory"].endswith("Axion-Strings/output_files")
    # optional tags take their defaults
    assert p["precision"] == "fp64" and p["init_mode"] == "thermal" and p["n_steps"] == 0
    assert p["pk_binning"] == "linear" and p["pk_nbins"] == 64


def _write(tmp_path, text):
    f = tmp_path / "run.param"
    f.write_text(text)
    return str(f)


def _base(root):
    return open(os.path.join(root, "tests", "golden", "default.param")).read()


def test_duplicate_and_unknown_tags_are_errors(root, tmp_path):
    base = _base(root)
    with pytest.raises(ValueError):
        axs.read_params(_write(tmp_path, base + "\nN 64\n"))          # duplicate
    with pytest.raises(ValueError):
        axs.read_params(_write(tmp_path, base + "\nfoo 1\n"))         # unknown
    missing = "\n".join(l for l in base.splitlines() if not l.startswith("thr "))
    with pytest.raises(ValueError):
        axs.read_params(_write(tmp_path, missing))                    # missing
    with pytest.raises(ValueError):
        axs.read_params(_write(tmp_path, base + "\nn_steps 5\nn_steps 6\n"))


def test_error_messages_match_reference(root, tmp_path):
    # read_parameters.cpp:222 / :233 messages, exit status failure
    base = _base(root)
    f = _write(tmp_path, base + "\nfoo 1\n")
    r = subprocess.run([axs.DRIVER_PATH, f], capture_output=True, text=True)
    assert r.returncode != 0
    assert f"Error in file {f}:   Tag 'foo' not allowed or multiple defined." in r.stdout
    missing = "\n".join(l for l in base.splitlines() if not l.startswith("seed "))
    f = _write(tmp_path, missing)
    r = subprocess.run([axs.DRIVER_PATH, f], capture_output=True, text=True)
    assert r.returncode != 0
    assert f"Error. I miss a value for tag 'seed' in parameter file '{f}'." in r.stdout
    r = subprocess.run([axs.DRIVER_PATH], capture_output=True, text=True)
    assert r.returncode != 0 and "Error: usage ./main <ParameterFile>" in r.stdout


def test_optional_tags(root, tmp_path):
    base = _base(root)
    p = axs.read_params(_write(tmp_path, base + "\nn_steps 500\nprecision fp32\ninit_mode kinv\n"
                               "pk_binning log\npk_nbins 32\nsample_energy 1\n"))
    assert (p["n_steps"], p["precision"], p["init_mode"]) == (500, "fp32", "kinv")
    assert (p["pk_binning"], p["pk_nbins"], p["sample_energy"]) == ("log", 32, 1)


def test_final_step():
    h = axs.load_host_library()
    # s_run.cpp:82: round(lct/dt) - round(dx/dt) + 1, lct = 0.5 N dx
    assert h.axs_host_final_step(256, 0.5, 0.1) == 636
    assert h.axs_host_final_step(1024, 0.5, 0.1) == 2556
    assert h.axs_host_final_step(128, 1.0, 0.3) == 211      # default.param


@pytest.mark.parametrize("nd,count,tau,N,dx", [(2, 100, 1.0, 256, 0.5), (2, 77, 3.7, 256, 0.5),
                                               (3, 300, 1.0, 256, 0.5), (3, 12345, 21.3, 64, 0.5)])
def test_xi_matches_oracle(nd, count, tau, N, dx):
    h = axs.load_host_library()
    want = O.xi2(count, tau, N, dx) if nd == 2 else O.xi3(count, tau, N, dx)
    assert h.axs_host_xi(nd, count, tau, N, dx) == want


@pytest.mark.parametrize("nd,N", [(1, 8), (1, 12), (1, 30), (1, 192), (1, 1024), (1, 49),
                                  (2, 24), (2, 64), (3, 16), (3, 12), (3, 20)])
def test_host_fft_matches_numpy(nd, N):
    rng = np.random.default_rng(N + nd)
    x = rng.standard_normal((N,) * nd) + 1j * rng.standard_normal((N,) * nd)
    got = axs.host_fft(x, nd, N, -1).reshape(x.shape)
    np.testing.assert_allclose(got, np.fft.fftn(x), rtol=0, atol=1e-10 * np.sqrt(x.size) * 10)
    back = axs.host_fft(got, nd, N, +1).reshape(x.shape) / x.size
    np.testing.assert_allclose(back, x, rtol=0, atol=1e-12)


@pytest.mark.parametrize("nd,N", [(2, 128), (3, 32)])
def test_host_init_kinv(nd, N):
    p1, p2, v1, v2 = axs.host_init("kinv", nd, N, seed=1998)
    assert abs(np.mean(p1 ** 2 + p2 ** 2) - 1.0) < 1e-12
    assert np.all(v1 == 0) and np.all(v2 == 0)
    # only modes 0 < |k| < pi/(4 dx) are populated: |n| < N/8 at dx = 0.5
    F = np.fft.fftn((p1 + 1j * p2).reshape((N,) * nd))
    f = np.fft.fftfreq(N, 1.0 / N)
    k2 = sum(x ** 2 for x in np.meshgrid(*[f] * nd, indexing="ij"))
    outside = np.abs(F[(k2 == 0) | (k2 >= (N / 8) ** 2)])
    assert outside.max() < 1e-9 * np.abs(F).max()
    # different seeds differ, same seed repeats
    q1, _, _, _ = axs.host_init("kinv", nd, N, seed=1998)
    r1, _, _, _ = axs.host_init("kinv", nd, N, seed=1999)
    assert np.array_equal(p1, q1) and not np.allclose(p1, r1)


def test_host_init_thermal_spectrum():
    # init.cpp:85-100: <|phi_k|^2> = 2 n_B / w and <|phidot_k|^2> = 2 n_B w per
    # mode (complex normal amplitudes), phi = IFFT / N^d; zero k = 0 mode
    nd, N, dx, lam, T0 = 2, 256, 0.5, 1.0, 4.0
    p1, p2, v1, v2 = axs.host_init("thermal", nd, N, space_step=dx, lam=lam, T_initial=T0)
    L = N ** nd
    F = np.fft.fftn((p1 + 1j * p2).reshape(N, N))        # = spectrum (IFFT/N^d inverted)
    V = np.fft.fftn((v1 + 1j * v2).reshape(N, N))
    f = np.fft.fftfreq(N, 1.0 / N)
    kk = np.sqrt(sum(x ** 2 for x in np.meshgrid(f, f, indexing="ij"))) * 2 * np.pi / (N * dx)
    meff2 = float(np.float32(lam * (T0 ** 2 / 3.0 - 1.0)))
    w = np.sqrt(kk ** 2 + meff2)
    with np.errstate(divide="ignore"):
        nb = 1.0 / (np.exp(w / T0) - 1.0)
    assert abs(F[0, 0]) < 1e-9 and abs(V[0, 0]) < 1e-9
    for lo, hi in ((0.2, 0.5), (1.0, 2.0), (4.0, 6.0)):
        sel = (kk >= lo) & (kk < hi)
        r_phi = np.mean(np.abs(F[sel]) ** 2 / (2 * nb[sel] / w[sel]))
        r_vel = np.mean(np.abs(V[sel]) ** 2 / (2 * nb[sel] * w[sel]))
        assert 0.85 < r_phi < 1.15 and 0.85 < r_vel < 1.15, (lo, r_phi, r_vel)


def test_philox4x32_10_known_answers():
    # the published Philox4x32-10 known-answer vectors (Random123's
    # kat_vectors: counter, key -> output): all-zero, all-ones, and the
    # digits-of-pi case
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert axs.host_philox(ctr, key) == out, [hex(x) for x in axs.host_philox(ctr, key)]
