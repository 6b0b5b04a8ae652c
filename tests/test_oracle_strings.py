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
