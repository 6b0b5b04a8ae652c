"""Pins the oracle's string finder (cores2/cores3, string_finding.cpp:3-86),
core count (s_run.cpp:101-109) and xi (s_run.cpp:111-131)."""
import numpy as np
import pytest

import oracle as O
from fields import kinv_field, vortex2d


def _axion_from_angles(theta):
    # phi1 = sin(theta), phi2 = cos(theta) gives atan2(phi1, phi2) = theta
    return O.axion(np.sin(theta), np.cos(theta))


def test_uniform_field_has_no_strings():
    for nd, N in ((2, 16), (3, 8)):
        L = N ** nd
        n, _ = O.cores(O.axion(np.full(L, 0.3), np.full(L, 0.8)), nd, N)
        assert n == 0


def test_hand_computed_plaquettes():
    # 2D N=4, angles chosen by hand at the corners of plaquette (1,1):
    # (1,1)=-pi/2, (2,1)=0, (2,2)=pi/2, (1,2)=pi (rest 0).  Normalised
    # n = (theta + pi)/(2 pi): 0.25, 0.5, 0.75, 1.0 -> edges 0.25, 0.25,
    # min(0.25, 0.75) = 0.25 -> sum 0.75 > 0.495: flagged.
    N = 4
    th = np.zeros((N, N))              # [j, i]
    th[1, 1], th[1, 2], th[2, 2], th[2, 1] = -np.pi / 2, 0.0, np.pi / 2, np.pi * 0.999999
    n, c = O.cores(_axion_from_angles(th.reshape(-1)), 2, N)
    flagged = {tuple(x) for x in c}
    assert (1, 1) in flagged
    # plaquette (2,2) has corners (2,2)=pi/2 -> 0.75, (3,2)=0 -> 0.5,
    # (3,3)=0 -> 0.5, (2,3)=0 -> 0.5: edges 0.25, 0, 0 -> 0.25: not flagged
    assert (2, 2) not in flagged
    # three edges only (Q7): a winding plaquette whose large jump is on the
    # omitted fourth edge (i,j+1)->(i,j) still passes through its other edges
    th2 = np.zeros((N, N))
    th2[0, 1], th2[1, 1] = 0.4 * np.pi, 0.8 * np.pi   # (1,0), (1,1)
    th2[1, 0] = -0.8 * np.pi                          # (0,1)
    n2, c2 = O.cores(_axion_from_angles(th2.reshape(-1)), 2, N)
    # plaquette (0,0): n = .5, .7, .9, .1 -> edges .2, .2, min(.8,.2)=.2 -> .6
    assert (0, 0) in {tuple(x) for x in c2}


def test_threshold_is_float_computed():
    # Q6: accept = 0.5f * (1.0f - thr/100.0f) in float = 0.49500000476837158
    # A plaquette whose three-edge sum lies between 0.495 and that float
    # value is NOT flagged.
    N = 4
    th = np.zeros((N, N))
    # corners of (0,0): n1=.5 at (0,0); (1,0): .5+a; (1,1): .5+2a; (0,1): .5+3a
    a = 0.4950000010 / 3
    th[0, 1], th[1, 1], th[1, 0] = 2 * np.pi * a, 2 * np.pi * 2 * a, 2 * np.pi * 3 * a
    n, c = O.cores(_axion_from_angles(th.reshape(-1)), 2, N, thr=1)
    assert (0, 0) not in {tuple(x) for x in c}
    a = 0.4950000060 / 3
    th[0, 1], th[1, 1], th[1, 0] = 2 * np.pi * a, 2 * np.pi * 2 * a, 2 * np.pi * 3 * a
    n, c = O.cores(_axion_from_angles(th.reshape(-1)), 2, N, thr=1)
    assert (0, 0) in {tuple(x) for x in c}


def _winding(theta2d):
    """brute-force winding number of every plaquette (i,j) of a [j,i] array"""
    t = theta2d / (2 * np.pi)

    def wrap(d):
        return d - np.round(d)
    a = t
    b = np.roll(t, -1, axis=1)                       # (i+1, j)
    c = np.roll(np.roll(t, -1, axis=1), -1, axis=0)  # (i+1, j+1)
    d = np.roll(t, -1, axis=0)                       # (i, j+1)
    return np.rint(wrap(b - a) + wrap(c - b) + wrap(d - c) + wrap(a - d)).astype(int)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_every_winding_plaquette_is_flagged(seed):
    # For winding +-1 the four wrapped edge jumps sum to +-1, so the three
    # counted ones sum to >= 0.5 > accept: every pierced plaquette is flagged.
    N = 64
    p1, p2 = kinv_field(2, N, seed=seed)
    ax = O.axion(p1, p2)
    w = _winding(ax.reshape(N, N))
    n, c = O.cores(ax, 2, N)
    flagged = np.zeros((N, N), bool)
    flagged[c[:, 1], c[:, 0]] = True
    assert np.count_nonzero(w) > 0
    assert np.all(flagged[w != 0])


def test_isolated_vortex_pair():
    # smooth vortex-antivortex pair on a 32^2 lattice: the plaquettes
    # containing the two cores are flagged, distant smooth regions are not
    # (the product field is not periodic, so the wrap column i = N-1 carries
    # a phase jump between the two cores and is left out of the check)
    N = 32
    p1, p2 = vortex2d(N, [(8.5, 8.5), (24.5, 20.5)], [1, -1])
    ax = O.axion(p1, p2)
    n, c = O.cores(ax, 2, N)
    flagged = {tuple(x) for x in c}
    assert (8, 8) in flagged and (24, 20) in flagged
    w = _winding(ax.reshape(N, N))
    assert w[8, 8] != 0 and w[20, 24] != 0
    assert all(abs(i - 8) <= 3 and abs(j - 8) <= 3 or abs(i - 24) <= 3 and abs(j - 20) <= 3
               for i, j in flagged if i < N - 1 and j < N - 1)


def test_list_is_in_ascending_site_order():
    N = 48
    p1, p2 = kinv_field(2, N, seed=5)
    n, c = O.cores(O.axion(p1, p2), 2, N)
    m = c[:, 0] + N * c[:, 1]
    assert n > 10 and np.all(np.diff(m) > 0)
    p1, p2 = kinv_field(3, 16, seed=5)
    n, c = O.cores(O.axion(p1, p2), 3, 16)
    m = (c[:, 0] * 16 + c[:, 1]) * 16 + c[:, 2]
    assert n > 10 and np.all(np.diff(m) > 0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_straight_string_3d(axis):
    # A vortex line along one axis: every slice across it flags the same
    # sites (translation invariance), and at least the core plaquette.
    N = 12
    j2, i2 = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    z = (i2 - 5.5) + 1j * (j2 - 6.5)          # vortex in the transverse plane
    th2 = np.angle(z)                          # [b, a] transverse coords
    th = np.zeros((N, N, N))
    for s in range(N):
        if axis == 0:
            th[s, :, :] = th2            # varies in (j,k); line along i
        elif axis == 1:
            th[:, s, :] = th2            # varies in (i,k); line along j
        else:
            th[:, :, s] = th2            # varies in (i,j); line along k
    n, c = O.cores(_axion_from_angles(th.reshape(-1)), 3, N)
    assert n > 0
    sets = {}
    for x in c:
        key = x[axis]
        rest = tuple(np.delete(x, axis))
        sets.setdefault(key, set()).add(rest)
    assert len(sets) == N
    first = sets[0]
    assert all(v == first for v in sets.values())


def test_num_cores_known_lists():
    # s_run.cpp:101-109: subtract consecutive pairs differing by (1,0) or (0,1)
    c = np.array([[0, 0], [1, 0], [3, 0], [3, 1], [5, 7]], np.int32)
    assert O.num_cores2(c) == 5 - 2
    c = np.array([[2, 2]], np.int32)
    assert O.num_cores2(c) == 1
    c = np.array([[0, 0], [0, 1], [1, 1], [2, 1], [2, 3]], np.int32)
    assert O.num_cores2(c) == 5 - 3


def test_xi_known_values():
    # 2D: N=256, dx=0.5, tau=1 -> t = 0.5, a = 1, area = 128^2 = 16384
    assert O.xi2(100, 1.0, 256, 0.5) == np.float32(100 * 0.25 / 16384)
    # 3D: n=300 -> length 200 (float), a=1 -> xi = 200 * 0.25 / 128^3
    assert O.xi3(300, 1.0, 256, 0.5) == np.float32(np.float32(200.0) * 0.25 / 128 ** 3)
    # tau = 2: t = 2, a = 2; 2D area = (256*0.5*2)^2 = 65536 -> xi = n*4/65536
    assert O.xi2(64, 2.0, 256, 0.5) == np.float32(64 * 4.0 / 65536)


# ---------------------------------------------------------------------------
# 3D plaquette geometry (string_finding.cpp:49-62) and the threshold
# percentage (string_finding.cpp:7-8, 39-40), pinned by flag sets derived by
# hand.  Corner offsets of site s = (i, j, k), in the order the edges run:
#   a: s, s+(1,0,0), s+(1,1,0), s+(0,1,0)
#   b: s, s+(1,0,0), s+(0,1,1), s+(0,0,1)     (not planar: Q8)
#   c: s, s+(0,1,0), s+(0,1,1), s+(0,0,1)
# Three edges per plaquette: corner 1 and corner 4 each touch one counted
# edge, corners 2 and 3 two.

def _flag_set_3d(turns, N, thr=1):
    """flagged sites of a field given as angles in turns, [i, j, k]"""
    n, c = O.cores(2 * np.pi * np.asarray(turns, np.float64).reshape(-1), 3, N, thr=thr)
    return {tuple(int(v) for v in x) for x in c}


def _shift(p, d, N):
    return tuple((a - b) % N for a, b in zip(p, d))


@pytest.mark.parametrize("delta", [0.3, 0.6, -0.3])
@pytest.mark.parametrize("P", [(2, 3, 1), (0, 0, 0), (5, 5, 5)])
def test_single_raised_corner_3d(delta, P):
    # One site turned by delta (circular distance 0.3 or 0.4 to the rest),
    # the rest at angle 0.  A plaquette with P as corner 2 or 3 sums to
    # 2 * 0.3 (0.4) > 0.495; with P as corner 1 or 4 to 0.3 (0.4) < 0.495.
    # P is corner 2/3 of: a at s = P-(1,0,0), P-(1,1,0); b at P-(1,0,0),
    # P-(0,1,1); c at P-(0,1,0), P-(0,1,1).
    N = 6
    t = np.zeros((N, N, N))
    t[P] = delta
    want = {_shift(P, d, N) for d in ((1, 0, 0), (1, 1, 0), (0, 1, 1), (0, 1, 0))}
    assert _flag_set_3d(t, N) == want


def test_plaquette_a_alone_decides():
    # P = s+(1,1,0) is corner 3 of a at s (sum 0.6) and no corner of b or c
    # at s, so s is flagged by a alone; s+(0,1,0) = P-(1,0,0) by a and b.
    N = 6
    s = (2, 2, 2)
    t = np.zeros((N, N, N))
    t[3, 3, 2] = 0.3
    got = _flag_set_3d(t, N)
    assert s in got
    assert got == {(2, 3, 2), (2, 2, 2), (3, 2, 1), (3, 2, 2)}


def test_plaquette_c_alone_decides():
    # P = s+(0,1,0) is corner 2 of c at s (0.6), corner 4 of a at s (0.3),
    # no corner of b at s: s is flagged by c alone.
    N = 6
    t = np.zeros((N, N, N))
    t[2, 3, 2] = 0.3
    got = _flag_set_3d(t, N)
    assert (2, 2, 2) in got
    assert got == {(1, 3, 2), (1, 2, 2), (2, 2, 1), (2, 2, 2)}


@pytest.mark.parametrize("s", [(2, 2, 2), (5, 0, 5)])
def test_plaquette_b_alone_decides(s):
    # s+(1,0,0) at +0.2 turns and s+(0,1,1) at -0.2, everything else at 0.
    # At s: a = 0.2 + 0.2 + 0 = 0.4, c = 0 + 0.2 + 0.2 = 0.4, but
    # b = 0.2 + min(0.4, 0.6) + 0.2 = 0.8 > 0.495.  Any other plaquette
    # holds at most one of the two sites (only b pairs corners that differ
    # by (-1, 1, 1)) and sums to at most 0.4, so s is the only flagged site.
    # A planar b through (i+1, j, k+1) would flag nothing here.
    N = 6
    t = np.zeros((N, N, N))
    t[(s[0] + 1) % N, s[1], s[2]] = 0.2
    t[s[0], (s[1] + 1) % N, (s[2] + 1) % N] = -0.2
    assert _flag_set_3d(t, N) == {s}


def test_straight_string_along_j_exact_set():
    # Angles in quarter turns q(i, k), the same for every j: one winding
    # cell (corners (1,1), (2,1), (2,2), (1,2) at 0, 1, 2, 3 quarters).
    #       k: 0 1 2 3
    #   i=0:   0 0 0 0
    #   i=1:   0 0 3 0
    #   i=2:   0 1 2 0
    #   i=3:   0 0 0 0
    # Along j the plaquettes of site (i, j, k) reduce to
    #   a = 2 d(q(i,k), q(i+1,k)),
    #   b = d(q(i,k), q(i+1,k)) + d(q(i+1,k), q(i,k+1)),
    #   c = d(q(i,k), q(i,k+1)),
    # with d in quarters 0, .25, .5, .25 for |dq| = 0, 1, 2, 3.  a flags every
    # (i, k) whose i-neighbour differs: (0,2), (1,1), (1,2), (2,1), (2,2); b
    # and c add nothing (b is at most .25 where a is 0, c never reaches .5).
    # A planar b would also flag (1, 0): .25 + .25 through (2,0)-(2,1)-(1,1).
    N = 4
    q = np.array([[0, 0, 0, 0], [0, 0, 3, 0], [0, 1, 2, 0], [0, 0, 0, 0]], np.float64)
    t = np.repeat((q / 4.0)[:, None, :], N, axis=1)        # [i, j, k]
    cells = {(0, 2), (1, 1), (1, 2), (2, 1), (2, 2)}
    want = {(i, j, k) for (i, k) in cells for j in range(N)}
    assert _flag_set_3d(t, N) == want


# float thresholds 0.5f * (1.0f - thr/100.0f): thr 0 -> 0.5 exactly,
# thr 1 -> 0.4950000047683716 (above 0.495), thr 5 -> 0.4749999940395355
# (below 0.475), thr 30 -> 0.3499999940395355 (below 0.35)
FLOAT_ACCEPT = {0: 0.5, 1: 0.4950000047683716, 5: 0.4749999940395355, 30: 0.3499999940395355}


@pytest.mark.parametrize("thr", [0, 1, 5, 30])
def test_threshold_percent_2d(thr):
    # three equal edges of s/3 turns around plaquette (0,0): flagged iff the
    # sum exceeds the float threshold; sums a few 1e-9 either side of it,
    # and of the decimal value when the two differ by more than that
    N = 4
    acc = FLOAT_ACCEPT[thr]

    def flagged(s):
        a = s / 3
        th = np.zeros((N, N))
        th[0, 1], th[1, 1], th[1, 0] = 2 * np.pi * a, 2 * np.pi * 2 * a, 2 * np.pi * 3 * a
        n, c = O.cores(th.reshape(-1), 2, N, thr=thr)
        return (0, 0) in {tuple(x) for x in c}

    assert flagged(acc + 2e-9)
    assert not flagged(acc - 2e-9)
    dec = 0.5 * (1 - thr / 100)
    if thr == 1:
        assert not flagged(dec + 2e-9)      # the float threshold is above 0.495
    elif thr in (5, 30):
        assert flagged(dec - 2e-9)          # ... and below 0.475, 0.35


@pytest.mark.parametrize("thr,delta,n_want", [
    (1, 0.2, 0), (30, 0.2, 4),              # 2 * 0.2 = 0.4 against 0.495 / 0.35
    (1, 0.4, 4), (30, 0.4, 6),              # one edge of 0.4 passes 0.35 too
    (5, 0.24, 4), (1, 0.24, 0),             # 0.48 against 0.475 / 0.495
    (0, 0.2499, 0), (0, 0.2501, 4),         # 0.4998 / 0.5002 against 0.5
    (100, 0.01, 6), (100, 0.0, 0),          # thr 100: accept 0, any edge passes; flat: 0 > 0 fails
])
def test_threshold_percent_3d(thr, delta, n_want):
    # single raised corner: corners 2/3 give 2 * d, corners 1/4 give d
    # (d = min(delta, 1 - delta)); at thr = 30 one edge of 0.4 flags the two
    # sites with P as corner 1 or 4 of some plaquette as well: P itself and
    # P-(0,0,1) (a's corner 4 at P-(0,1,0) is already flagged).
    N = 6
    P = (3, 2, 4)
    t = np.zeros((N, N, N))
    t[P] = delta
    got = _flag_set_3d(t, N, thr=thr)
    assert len(got) == n_want
    core = {_shift(P, d, N) for d in ((1, 0, 0), (1, 1, 0), (0, 1, 1), (0, 1, 0))}
    if n_want == 4:
        assert got == core
    if n_want == 6:
        assert got == core | {P, _shift(P, (0, 0, 1), N)}
