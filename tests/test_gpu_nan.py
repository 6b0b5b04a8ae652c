"""The reference's debug scan (s_run.cpp:4-12): the first site whose phi1 or
phi2 is NaN, through the C-ABI (axs_member_nan_check) and through the
driver's debug_nan tag, which stops the run with the reference's message."""
import struct

import numpy as np
import pytest

import axs
from fields import random_state
from test_gpu_driver import PARAM

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nd,N,prec", [(3, 64, "fp64"), (3, 64, "fp32"), (2, 256, "fp32"),
                                       (3, 48, "fp64")])
def test_nan_check_finds_the_first(nd, N, prec):
    lat = axs.Lattice(nd, N, precision=prec)
    p1, p2, v1, v2, _, _ = random_state(nd, N, seed=3)
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    assert lat.nan_check() == -1
    L = N ** nd
    hits = [L - 1, 12345 % L, 777 % L]
    q1, q2 = p1.copy(), p2.copy()
    q1[hits[0]] = np.nan
    q2[hits[1]] = np.nan
    q1[hits[2]] = np.nan
    lat.set_state(q1, q2, v1, v2, tau=1.0)
    assert lat.nan_check() == min(hits)
    q2[0] = np.nan                                   # the very first site
    lat.set_state(q1, q2, v1, v2, tau=1.0)
    assert lat.nan_check() == 0
    lat.set_state(p1, p2, v1, v2, tau=1.0)
    lat.step(3)
    assert lat.nan_check() == -1


def test_nan_check_ensemble_members():
    Ns = [64, 96, 64]
    lat = axs.Lattice(2, members=Ns, precision="fp32")
    for r, N in enumerate(Ns):
        p1, p2, v1, v2, _, _ = random_state(2, N, seed=10 + r)
        if r == 1:
            p1[N * 5 + 7] = np.nan
        lat.set_state(p1, p2, v1, v2, member=r)
    assert [lat.nan_check(member=r) for r in range(3)] == [-1, 96 * 5 + 7, -1]


def _run(tmp_path, name, steps, extra):
    import subprocess
    d = tmp_path / name
    out = d / "out"
    out.mkdir(parents=True)
    pf = d / "run.param"
    pf.write_text(PARAM.format(nd=2, N=64, steps=steps, out=out, init="thermal", T0=4.0,
                               extra=extra, lam=1.0))
    r = subprocess.run([axs.DRIVER_PATH, str(pf)], capture_output=True, text=True, timeout=600)
    return r, out


def test_driver_debug_nan(tmp_path):
    """debug_nan 1 changes nothing on a clean run; restarted from a
    checkpoint with a NaN it stops before the first step with the
    reference's message and a failure status."""
    ck = tmp_path / "ck.bin"
    r, out_a = _run(tmp_path, "a", 40, f"checkpoint_path {ck}\ndebug_nan 1")
    assert r.returncode == 0, r.stdout + r.stderr
    r, out_b = _run(tmp_path, "b", 40, "")
    assert r.returncode == 0, r.stdout + r.stderr
    assert (out_a / "time-series.csv").read_text() == (out_b / "time-series.csv").read_text()
    # header: magic[8], int32 ndims, int32 N, int64 step, double tau; then
    # phi1, phi2, phidot1, phidot2, ker1, ker2 as doubles
    raw = bytearray(ck.read_bytes())
    hdr = struct.calcsize("<8siiqd")
    step = struct.unpack_from("<8siiqd", raw)[3]
    L = 64 * 64
    struct.pack_into("<d", raw, hdr + 8 * (L + 100), float("nan"))      # phi2 of site 100
    bad = tmp_path / "bad.bin"
    bad.write_bytes(bytes(raw))
    r, out_c = _run(tmp_path, "c", 90, f"restart_path {bad}\ndebug_nan 1")
    assert r.returncode != 0
    text = (out_c / "output.txt").read_text()
    assert "Error: NaN encountered in solution vector.\n" in text
    assert f" tstep = {step}\n" in text
