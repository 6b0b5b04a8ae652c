"""End-to-end: the host driver axs_main (run_standard, s_run.cpp:26-232) on
a small lattice, every output file compared with a replay of the same run
on the CPU oracle; checkpoint / restart continues bit-exactly."""
import math
import os
import subprocess

import numpy as np
import pytest

import axs
import oracle as O

pytestmark = pytest.mark.gpu

PARAM = """% test run
lambda              {lam}
enable_PRS          1
T_initial           {T0}
NDIMS               {nd}
N                   {N}
space_step          0.5
time_step           0.1
stencil_setting     0
seed                1998
write_output_file   1
output_file_path    {out}/output.txt
save_snapshots      1
n_snapshots         3
output_directory    {out}
save_fields         1
save_strings        1
save_pk             1
sample_time_series  1
n_samples           5
ts_output_path      {out}/time-series.csv
sample_strings      1
sample_background   1
thr                 1
enable_amr          0
init_from_snapshot  0
tau_initial         1.0
refinement_threshold 20.0
n_steps             {steps}
init_mode           {init}
{extra}
"""


def _cadence(t, n, final):
    if t == 0:
        return True
    every = final // (n - 1)
    return every > 0 and t % every == 0


def _run_driver(tmp_path, nd, N, steps, init="thermal", T0=4.0, extra="", lam=1.0):
    out = tmp_path / "out"
    out.mkdir(exist_ok=True)
    pf = tmp_path / "run.param"
    pf.write_text(PARAM.format(nd=nd, N=N, steps=steps, out=out, init=init, T0=T0, extra=extra,
                               lam=lam))
    r = subprocess.run([axs.DRIVER_PATH, str(pf)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


@pytest.mark.parametrize("nd,N,steps,init,eom,lap", [
    (2, 128, 120, "thermal", "field_rescaled", "stencil"), (3, 32, 60, "kinv", "field_rescaled", "stencil"),
    (2, 64, 50, "thermal", "hamiltonian", "stencil"), (2, 64, 40, "kinv", "field_rescaled", "spectral")])
def test_driver_outputs_match_oracle(tmp_path, nd, N, steps, init, eom, lap):
    # the Hamiltonian form and the spectral Laplacian are the reference's
    # alternative step paths, selected here by the optional tags
    out = _run_driver(tmp_path, nd, N, steps, init=init, extra=f"eom {eom}\nlaplacian {lap}")
    spectral = lap == "spectral"
    eom_i, lap_i = int(eom == "hamiltonian"), int(spectral)
    L = N ** nd
    p1, p2, v1, v2 = axs.host_init(init, nd, N, space_step=0.5, lam=1.0, T_initial=4.0, seed=1998)
    st = O.State(p1, p2, v1, v2)
    prm = O.params(nd, N, T_initial=4.0)
    A = O.Matrix(nd, N)
    ts_rows, snaps = [], []
    for t in range(steps):
        if _cadence(t, 5, steps):
            ax = O.axion(st.phi1, st.phi2)
            n, c = O.cores(ax, nd, N)
            if nd == 2:
                xi = O.xi2(O.num_cores2(c), st.tau, N, 0.5)
            else:
                xi = O.xi3(n, st.tau, N, 0.5)
            bg = O.background(st.phi1, st.phi2, st.phidot1, st.phidot2) / L
            t_phys = np.float32(0.5 * st.tau * st.tau)
            ts_rows.append([float(t_phys), float(np.float32(st.tau)), float(xi), n] + list(bg))
        if _cadence(t, 3, steps):
            ks, pk, cnt = O.powerspec(st.phi2, nd, N, 0.5)
            n, c = O.cores(O.axion(st.phi1, st.phi2), nd, N)
            snaps.append((st.tau, st.phi1.copy(), st.phi2.copy(), c, ks, pk, cnt))
        st.run_form(prm, A, 1, eom=eom_i, laplacian=lap_i)

    # time series: header and values (printed with %f)
    lines = open(out / "time-series.csv").read().splitlines()
    assert lines[0] == ("time,scalefactor,xi,n_plaquettes,phi1_bar,phi2_bar,phidot1_bar,"
                        "phidot2_bar,axion_bar,saxion_bar")
    assert len(lines) - 1 == len(ts_rows)
    for line, want in zip(lines[1:], ts_rows):
        got = [float(x) for x in line.split(",")]
        assert int(got[3]) == want[3]                       # n_plaquettes, exact
        np.testing.assert_allclose(got, want, rtol=0, atol=1.5e-6)

    # snapshot timings
    lines = open(out / "snapshot-timings.csv").read().splitlines()
    assert lines[0] == "snapshot,tau,hubble_scale,string_tension,"
    for s, (tau, *_rest) in enumerate(snaps):
        fields = lines[1 + s].split(",")
        assert fields[0] == str(s) and float(fields[1]) == pytest.approx(tau, abs=1e-6)
        _check_timing_columns(fields, tau, 1.0)

    for s, (tau, f1, f2, c, ks, pk, cnt) in enumerate(snaps):
        raw1 = np.fromfile(out / f"snapshot{s}-phi1", dtype=np.float64)
        raw2 = np.fromfile(out / f"snapshot{s}-phi2", dtype=np.float64)
        nsave = L if nd == 2 else N * N                    # 3D: first N^2 (fileio.cpp:31-46)
        if spectral:        # FFT rounding: rocFFT against the oracle's transform
            assert np.abs(raw1 - f1[:nsave]).max() < 1e-11 * np.abs(f1).max()
            assert np.abs(raw2 - f2[:nsave]).max() < 1e-11 * np.abs(f2).max()
        else:
            np.testing.assert_array_equal(raw1, f1[:nsave])
            np.testing.assert_array_equal(raw2, f2[:nsave])
        text = open(out / f"snapshot{s}-strings.csv").read()
        want = "".join(", ".join(str(int(v)) for v in row) + "\n" for row in c)
        assert text == want
        rows = open(out / f"snapshot{s}-pk.csv").read().splitlines()
        assert rows[0] == "k,pk,n_modes"
        sel = np.nonzero(cnt)[0]
        assert len(rows) - 1 == len(sel)
        for row, b in zip(rows[1:], sel):
            k_, p_, n_ = row.split(",")
            assert int(n_) == cnt[b]
            assert float(k_) == pytest.approx(ks[b], rel=1e-5)
            # empty-spectrum bins hold only transform rounding (~1e-26)
            assert float(p_) == pytest.approx(pk[b], rel=1e-5, abs=1e-12 * pk[sel].max())
    log = open(out / "output.txt").read()
    assert "Running with" in log and "Time taken:" in log and "Writing snapshot 0:" in log


def _check_timing_columns(fields, tau, lam):
    # physics.cpp:51-53 and 65-67: H = 1/tau^2, so hubble_scale = 1/H = tau^2,
    # and string_tension = log(m_saxion / H) = log(sqrt(lambda) tau^2), 0 when
    # lambda = 0 (s_run.cpp:162-163); float values printed with %f
    assert float(fields[2]) == pytest.approx(tau * tau, rel=3e-7, abs=1e-6)
    want = math.log(math.sqrt(lam) * tau * tau) if lam != 0.0 else 0.0
    assert float(fields[3]) == pytest.approx(want, rel=3e-7, abs=1.5e-6)


@pytest.mark.parametrize("lam", [0.5, 2.0, 0.0])
def test_snapshot_timing_columns(tmp_path, lam):
    """snapshot-timings.csv: tau, 1/H and the string tension of every
    snapshot for couplings other than 1, and the zero column at lambda = 0"""
    nd, N, steps = 2, 64, 61
    out = _run_driver(tmp_path, nd, N, steps, init="kinv", lam=lam)
    lines = open(out / "snapshot-timings.csv").read().splitlines()
    assert lines[0] == "snapshot,tau,hubble_scale,string_tension,"
    taus, t = [], 1.0
    for step in range(steps):
        if _cadence(step, 3, steps):
            taus.append(t)
        t = t + float(np.float32(0.1))
    assert len(lines) - 1 == len(taus) == 3
    for s, tau in enumerate(taus):
        fields = lines[1 + s].split(",")
        assert fields[0] == str(s) and float(fields[1]) == pytest.approx(tau, abs=1e-6)
        assert fields[4] == ""
        _check_timing_columns(fields, tau, lam)


def test_checkpoint_restart_is_bit_exact(tmp_path):
    nd, N = 2, 64
    a = tmp_path / "a"
    a.mkdir()
    ck1 = tmp_path / "ck1.bin"
    _run_driver(a, nd, N, 40, extra=f"checkpoint_path {ck1}")
    b = tmp_path / "b"
    b.mkdir()
    ck2 = tmp_path / "ck2.bin"
    _run_driver(b, nd, N, 90, extra=f"restart_path {ck1}\ncheckpoint_path {ck2}")
    c = tmp_path / "c"
    c.mkdir()
    ck3 = tmp_path / "ck3.bin"
    _run_driver(c, nd, N, 90, extra=f"checkpoint_path {ck3}")
    assert open(ck2, "rb").read() == open(ck3, "rb").read()
