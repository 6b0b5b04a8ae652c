"""Randomised (seeded, deterministic) parity of the sampling calls against the
oracle: random dimension, storage form, ensemble of uneven sides (including
sides that are not multiples of the kernels' 64-site tiles), field (smooth
k^-1, random phases, or the two mixed) and step count.  For every member:
the flagged-site count and, in 2D, the core count from axs_string_counts
(string_finding.cpp:3-86, s_run.cpp:99-109); the coordinate list of one
member; the spectra of all members from one axs_powerspec_all call
(pk_estimation.cpp:37-214); the energy and background sums of one member.
Strings are compared exactly, the reductions to FP64 rounding."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import kinv_field

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(seed)
    nd = int(rng.choice([2, 3]))
    precision = str(rng.choice(["fp32", "fp64"]))
    hi = 300 if nd == 2 else 100
    nmem = int(rng.integers(1, 5))
    Ns = [int(rng.integers(4, hi // 2 + 1)) * 2 for _ in range(nmem)]
    kinds = [str(rng.choice(["kinv", "random", "mixed"])) for _ in range(nmem)]
    nsteps = int(rng.integers(0, 6))
    nbins = int(rng.integers(5, 40))
    return rng, nd, precision, Ns, kinds, nsteps, nbins


def _field(rng, nd, N, kind, seed):
    if kind == "kinv" and N >= 16:
        return kinv_field(nd, N, seed=seed)
    p1, p2 = rng.standard_normal(N ** nd), rng.standard_normal(N ** nd)
    if kind == "mixed" and N >= 16:
        q1, q2 = kinv_field(nd, N, seed=seed)
        p1, p2 = q1 + 0.05 * p1, q2 + 0.05 * p2
    return p1, p2


@pytest.mark.parametrize("seed", range(16))
def test_random_sampling_matches_oracle(seed):
    rng, nd, precision, Ns, kinds, nsteps, nbins = _case(seed)
    tag = f"seed {seed}: nd {nd} {precision} sides {Ns} fields {kinds} steps {nsteps}"
    lat = axs.Lattice(nd, members=Ns, precision=precision)
    for r, (N, kind) in enumerate(zip(Ns, kinds)):
        p1, p2 = _field(rng, nd, N, kind, 500 + seed * 10 + r)
        v = 0.1 * rng.standard_normal((2, N ** nd))
        lat.set_state(p1, p2, v[0], v[1], member=r)
    lat.tau = 1.0
    lat.step(nsteps)
    states = [lat.get_state(member=r) for r in range(len(Ns))]

    npl, ncores = lat.string_counts()
    refs = []
    for r, (N, g) in enumerate(zip(Ns, states)):
        n_ref, c_ref = O.cores(O.axion(g["phi1"], g["phi2"]), nd, N, thr=1)
        refs.append(c_ref)
        assert npl[r] == n_ref, f"{tag}: member {r} count"
        assert ncores[r] == (O.num_cores2(c_ref) if nd == 2 else n_ref), f"{tag}: member {r} cores"
    r = int(rng.integers(0, len(Ns)))
    n, c = lat.strings(member=r)
    assert n == len(refs[r]), f"{tag}: member {r} list length"
    np.testing.assert_array_equal(c, refs[r], err_msg=f"{tag}: member {r} list")

    spectra = lat.powerspec_all(binning="log", nbins=nbins)
    for r, (N, g) in enumerate(zip(Ns, states)):
        ks_r, pk_r, cnt_r = O.powerspec_log(g["phi2"], nd, N, 0.5, nbins)
        ks, pk, cnt = spectra[r]
        np.testing.assert_array_equal(cnt, cnt_r, err_msg=f"{tag}: member {r} mode counts")
        sel = cnt > 0
        np.testing.assert_allclose(pk[sel], pk_r[sel], rtol=1e-9, atol=1e-12 * pk_r[sel].max(),
                                   err_msg=f"{tag}: member {r} spectrum")

    g = states[0]
    e = lat.energy()
    e_ref = O.energy(O.params(nd, Ns[0]), lat.tau, g["phi1"], g["phi2"], g["phidot1"],
                     g["phidot2"])
    np.testing.assert_allclose(e, e_ref, rtol=1e-11 if precision == "fp64" else 1e-5,
                               atol=1e-9 * np.abs(e_ref).sum(), err_msg=f"{tag}: energy")
    b = lat.background()
    b_ref = O.background(g["phi1"], g["phi2"], g["phidot1"], g["phidot2"])
    scale = np.array([np.abs(g[k]).sum() for k in ("phi1", "phi2", "phidot1", "phidot2")] +
                     [b_ref[4], b_ref[5]])
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert np.all(np.abs(b - b_ref) <= tol * scale + 1e-300), f"{tag}: background"
