"""GPU parity of the string finder away from the default threshold and on
hand-built 3D plaquette geometries: the finder's flags are compared with the
oracle's cores2/cores3 (string_finding.cpp:3-86) at thr = 0, 5, 30 (accept =
0.5f * (1.0f - thr/100.0f), string_finding.cpp:7-8, 39-40) and on lattices
where each of the 3D plaquettes a, b and c alone decides a site
(string_finding.cpp:49-62)."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import kinv_field

pytestmark = pytest.mark.gpu

THRS = [0, 5, 30]


def _compare(lat, nd, N, thr, member=0):
    g = lat.get_state(member=member, fields=("phi1", "phi2"))
    n_ref, c_ref = O.cores(O.axion(g["phi1"], g["phi2"]), nd, N, thr=thr)
    n, c = lat.strings(member=member)
    assert n == n_ref, (nd, N, thr, n, n_ref)
    np.testing.assert_array_equal(c, c_ref)
    return n_ref, c_ref


@pytest.mark.parametrize("thr", THRS)
@pytest.mark.parametrize("nd,N,prec", [(2, 256, "fp64"), (2, 130, "fp32"), (3, 64, "fp64"),
                                       (3, 66, "fp32"), (3, 100, "fp64")])
def test_strings_threshold_random_phases(thr, nd, N, prec):
    """dense random phases: the set, its order and the counts at thr != 1"""
    rng = np.random.default_rng(1000 * thr + N)
    p1, p2 = rng.standard_normal(N ** nd), rng.standard_normal(N ** nd)
    z = np.zeros(N ** nd)
    lat = axs.Lattice(nd, N, precision=prec, thr=thr)
    lat.set_state(p1, p2, z, z, tau=1.0)
    n_ref, c_ref = _compare(lat, nd, N, thr)
    assert n_ref > 0
    npl, ncores = lat.string_counts()
    assert npl[0] == n_ref
    assert ncores[0] == (O.num_cores2(c_ref) if nd == 2 else n_ref)


@pytest.mark.parametrize("thr", THRS)
@pytest.mark.parametrize("nd,N,nsteps,prec", [(2, 256, 120, "fp64"), (3, 96, 60, "fp64"),
                                              (3, 64, 40, "fp32")])
def test_strings_threshold_evolved(thr, nd, N, nsteps, prec):
    """an evolved k^-1 field (smooth, sparse strings) at thr != 1"""
    p1, p2 = kinv_field(nd, N, seed=1998)
    z = np.zeros(N ** nd)
    lat = axs.Lattice(nd, N, precision=prec, thr=thr)
    lat.set_state(p1, p2, z, z, tau=1.0)
    lat.step(nsteps)
    n_ref, _ = _compare(lat, nd, N, thr)
    assert n_ref > 0


FLOAT_ACCEPT = {0: 0.5, 5: 0.4749999940395355, 30: 0.3499999940395355}


@pytest.mark.parametrize("thr", THRS)
@pytest.mark.parametrize("N,prec", [(64, "fp64"), (66, "fp32")])
def test_strings_3d_threshold_straddle(thr, N, prec):
    """The phase turns by the float threshold of thr per site along k, with
    1e-9-turn noise: plaquettes b and c of every site sum to the threshold
    within 1e-8, so the float screen decides none of them and every site
    goes through the FP64 fallback; the flags must be the oracle's."""
    rng = np.random.default_rng(17 + thr)
    k = np.arange(N ** 3) % N
    turns = FLOAT_ACCEPT[thr] * k + 1e-9 * rng.standard_normal(N ** 3)
    p1, p2 = np.cos(2 * np.pi * turns), np.sin(2 * np.pi * turns)
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N, precision=prec, thr=thr)
    lat.set_state(p1, p2, z, z, tau=1.0)
    n_ref, _ = _compare(lat, 3, N, thr)
    assert 0.05 * N ** 3 < n_ref < 0.99 * N ** 3


@pytest.mark.parametrize("thr", [5, 30])
@pytest.mark.parametrize("nd,Ns,prec", [(2, [64, 130, 256], "fp64"), (2, [96, 64], "fp32"),
                                        (3, [64, 34, 96], "fp64"), (3, [66, 64], "fp32")])
def test_string_counts_ensemble_threshold(thr, nd, Ns, prec):
    """ensemble-wide counts at thr != 1 against the oracle per member"""
    lat = axs.Lattice(nd, members=Ns, precision=prec, thr=thr)
    fields = []
    for r, N in enumerate(Ns):
        rng = np.random.default_rng(500 + r)
        p1, p2 = rng.standard_normal(N ** nd), rng.standard_normal(N ** nd)
        z = np.zeros(N ** nd)
        lat.set_state(p1, p2, z, z, member=r)
        g = lat.get_state(member=r, fields=("phi1", "phi2"))
        fields.append((g["phi1"], g["phi2"]))
    npl, ncores = lat.string_counts()
    for r, (N, (p1, p2)) in enumerate(zip(Ns, fields)):
        n_ref, c_ref = O.cores(O.axion(p1, p2), nd, N, thr=thr)
        assert npl[r] == n_ref, (r, N)
        assert ncores[r] == (O.num_cores2(c_ref) if nd == 2 else n_ref), (r, N)


def _turn_field(turns):
    t = np.asarray(turns, np.float64).reshape(-1)
    return np.sin(2 * np.pi * t), np.cos(2 * np.pi * t)


# (i, j, k) positions on tile edges of the one-pass finder (30-row, 64-column
# tiles along j and k, chunks along i) and the lattice wrap
CORNERS = [(0, 0, 0), (5, 29, 63), (31, 30, 64), (63, 59, 1), (40, 89, 95), (95, 95, 95)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("thr,delta,n_per", [(1, 0.3, 4), (1, 0.2, 0), (30, 0.2, 4),
                                             (30, 0.4, 6), (5, 0.24, 4)])
def test_strings_3d_single_raised_corners(prec, thr, delta, n_per):
    """Isolated sites turned by delta on a flat field: every plaquette that
    holds one as corner 2 or 3 sums to 2 min(delta, 1 - delta), as corner 1
    or 4 to min(delta, 1 - delta).  The flagged set is 4 (or, where one edge
    passes, 6) sites per raised corner at fixed offsets -- the oracle's."""
    N = 96
    t = np.zeros((N, N, N))
    for P in CORNERS:
        t[P] = delta
    p1, p2 = _turn_field(t)
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N, precision=prec, thr=thr)
    lat.set_state(p1, p2, z, z, tau=1.0)
    n_ref, _ = _compare(lat, 3, N, thr)
    assert n_ref == n_per * len(CORNERS)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_strings_3d_plaquette_b_alone(prec):
    """Pairs s+(1,0,0) at +0.2 turns and s+(0,1,1) at -0.2: only plaquette b
    of s (corners s, s+(1,0,0), s+(0,1,1), s+(0,0,1)) holds both, so each s
    is flagged by b alone and nothing else is."""
    N = 96
    t = np.zeros((N, N, N))
    sites = [(2, 2, 2), (10, 29, 63), (20, 30, 0), (95, 95, 95), (50, 59, 127 % N)]
    for (i, j, k) in sites:
        t[(i + 1) % N, j, k] = 0.2
        t[i, (j + 1) % N, (k + 1) % N] = -0.2
    p1, p2 = _turn_field(t)
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N, precision=prec)
    lat.set_state(p1, p2, z, z, tau=1.0)
    n_ref, c_ref = _compare(lat, 3, N, 1)
    assert {tuple(int(v) for v in x) for x in c_ref} == set(sites)
