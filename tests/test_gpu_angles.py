"""The device's angles: the axion field (atan2(phi1, phi2), s_run.cpp:92-94)
and the string finder's decisions where a plaquette sum lies within a few
ulp of the threshold (string_finding.cpp:7-8, 22-27, 64-80).

The oracle's atan2 is the correctly rounded one (orc_atan2_cr, pinned to
mpmath in tests/test_atan2_cr.py).  The device's axion field must equal it
bit for bit on every site, in both storage forms, and so must the host build
of the device's double-double atan2.  Lattices are then built so that
thousands of plaquette sums lie within 3 ulp of the threshold, above and
below: every flag there depends on the last bit of the corner angles, and
the device's set must be the oracle's.  How often glibc's and numpy's atan2
would decide those flags differently is printed beside."""
import ctypes
import math

import numpy as np
import pytest

import axs
import oracle as O

pytestmark = pytest.mark.gpu

TWOPI = 2 * math.pi
FLOAT_ACCEPT = {1: 0.4950000047683716, 5: 0.4749999940395355}


def _host_dd(y, x):
    h = axs.load_host_library()
    f = h.axs_host_atan2_dd
    f.argtypes = [ctypes.c_int64] + [ctypes.POINTER(ctypes.c_double)] * 3
    f.restype = None
    y = np.ascontiguousarray(y, np.float64).reshape(-1)
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    out = np.empty_like(y)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    f(y.size, P(y), P(x), P(out))
    return out


def _same_bits(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_axion_field_is_correctly_rounded(prec):
    """3D 216^3 (1e7 sites) of random components over a wide range of
    magnitudes, with zeros of both signs, subnormals, infinities and NaN
    (FP64): the device's axion equals the oracle's correctly rounded atan2
    and the host build of the same double-double code, bit for bit."""
    N = 216
    L = N ** 3
    rng = np.random.default_rng(11)
    p1 = rng.standard_normal(L) * np.ldexp(1.0, rng.integers(-20, 20, L))
    p2 = rng.standard_normal(L) * np.ldexp(1.0, rng.integers(-20, 20, L))
    pick = rng.choice(L, 6000, replace=False)
    p1[pick[:1000]] = 0.0
    p2[pick[1000:2000]] = -0.0
    p1[pick[2000:2500]], p2[pick[2000:2500]] = -0.0, -0.0
    if prec == "fp64":
        p1[pick[2500:3000]] = 4e-320 * rng.standard_normal(500)
        p2[pick[3000:3500]] = np.inf * rng.choice([-1, 1], 500)
        p1[pick[3500:4000]] = -np.inf
        p1[pick[4000:4500]] = np.nan
        p1[pick[4500:5000]] = 1e300 * rng.standard_normal(500)
    z = np.zeros(L)
    lat = axs.Lattice(3, N, precision=prec)
    lat.set_state(p1, p2, z, z, tau=1.0)
    g = lat.get_state(fields=("phi1", "phi2"))      # the stored (possibly float) field
    got = lat.axion()
    want = O.atan2_cr(g["phi1"], g["phi2"])
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert _same_bits(got[~nan], want[~nan])
    hd = _host_dd(g["phi1"], g["phi2"])
    assert np.array_equal(np.isnan(hd), nan)
    assert _same_bits(got[~nan], hd[~nan])
    # how often other atan2 implementations differ on these arguments
    fin = np.isfinite(g["phi1"]) & np.isfinite(g["phi2"])
    npy = np.arctan2(g["phi1"][fin], g["phi2"][fin])
    sub = np.flatnonzero(fin)[:1_000_000]
    libm = np.array([math.atan2(a, b) for a, b in zip(g["phi1"][sub], g["phi2"][sub])])
    n_np = int(np.count_nonzero(npy != got[fin]))
    n_libm = int(np.count_nonzero(libm != got[sub]))
    print(f"{prec}: numpy.arctan2 differs on {n_np} of {fin.sum()} sites, "
          f"glibc atan2 on {n_libm} of {sub.size}")
    assert n_libm < 0.005 * sub.size


def _turns(p1, p2):
    return (O.atan2_cr(p1, p2).reshape(np.shape(p1)) + math.pi) / TWOPI


def _edge(a, b):
    d = np.abs(b - a)
    e = 1 - d
    return np.where(e < d, e, d)


def _tune(rng, a1, a2, base, prev, acc, shape):
    """step a1 by whole ulps until base + edge(prev, angle(a1, a2)) lies
    within 3 ulp of acc (either side); returns the tuned a1 and the offsets
    in ulps (NaN where no step landed)"""
    ulp = np.spacing(acc)
    best = np.full(shape, np.nan)
    pick = np.zeros(shape)
    for k in range(-200, 201):
        b1 = a1 + k * np.spacing(np.abs(a1))
        s = base + _edge(prev, _turns(b1, a2))
        off = (s - acc) / ulp
        take = (np.abs(off) <= 3) & (np.isnan(best) | (rng.uniform(0, 1, shape) < 0.3))
        best = np.where(take, off, best)
        pick = np.where(take, k, pick)
    return a1 + pick * np.spacing(np.abs(a1)), best


def _margin_field_2d(N, seed, acc):
    """2D [j, i] field: the plaquette of every site with even i, j has
    corners 1-3 at random angles 0.05-0.2 turns apart and corner 4 tuned so
    that the three-edge sum is within 3 ulp of acc"""
    rng = np.random.default_rng(seed)
    p1, p2 = rng.standard_normal((N, N)), rng.standard_normal((N, N))
    J, I = np.meshgrid(np.arange(0, N, 2), np.arange(0, N, 2), indexing="ij")
    shape = J.shape
    r = rng.uniform(0.5, 2.0, (4,) + shape)
    t0 = rng.uniform(0, 1, shape)
    d1 = rng.uniform(0.05, 0.2, shape) * rng.choice([-1, 1], shape)
    d2 = rng.uniform(0.05, 0.2, shape) * rng.choice([-1, 1], shape)
    corners = [(J, I), (J, I + 1), (J + 1, I + 1), (J + 1, I)]
    for c, t in enumerate((t0, t0 + d1, t0 + d1 + d2)):
        th = TWOPI * t - math.pi
        p1[corners[c]], p2[corners[c]] = r[c] * np.sin(th), r[c] * np.cos(th)
    n = [_turns(p1[corners[c]], p2[corners[c]]) for c in range(3)]
    s12 = _edge(n[0], n[1]) + _edge(n[1], n[2])
    th = TWOPI * (n[2] + (acc - s12) * rng.choice([-1, 1], shape)) - math.pi
    a1, a2 = r[3] * np.sin(th), r[3] * np.cos(th)
    a1, off = _tune(rng, a1, a2, s12, n[2], acc, shape)
    p1[corners[3]], p2[corners[3]] = a1, a2
    return p1.reshape(-1), p2.reshape(-1), off


def _margin_field_3d(N, seed, acc):
    """3D [i, j, k] field: at every site s with even i, j, k plaquette a
    (s, s+i, s+i+j, s+j) sums to within 3 ulp of acc, while s+k and s+j+k
    take s's value so that b = 2 d(s, s+i) and c = 2 d(s, s+j) stay below
    0.37: s's flag is decided by a alone"""
    rng = np.random.default_rng(seed)
    p1, p2 = rng.standard_normal((N, N, N)), rng.standard_normal((N, N, N))
    I, J, K = np.meshgrid(*(np.arange(0, N, 2),) * 3, indexing="ij")
    shape = I.shape
    r = rng.uniform(0.5, 2.0, (3,) + shape)
    t0 = rng.uniform(0, 1, shape)
    sgn = rng.choice([-1, 1], shape)
    d = rng.uniform(0.13, 0.17, shape) * sgn
    s_, si, sij, sj = (I, J, K), (I + 1, J, K), (I + 1, J + 1, K), (I, J + 1, K)
    for idx, (c, t) in enumerate(zip((s_, si, sij), (t0, t0 + d, t0 + 2 * d))):
        th = TWOPI * t - math.pi
        p1[c], p2[c] = r[idx] * np.sin(th), r[idx] * np.cos(th)
    n0, ni, nij = _turns(p1[s_], p2[s_]), _turns(p1[si], p2[si]), _turns(p1[sij], p2[sij])
    s12 = _edge(n0, ni) + _edge(ni, nij)
    # corner s+j back towards s: d3 = acc - s12 in the direction of s
    th = TWOPI * (nij - sgn * (acc - s12)) - math.pi
    a1, a2 = np.sin(th), np.cos(th)
    a1, off = _tune(rng, a1, a2, s12, nij, acc, shape)
    p1[sj], p2[sj] = a1, a2
    for c in ((I, J, K + 1), (I, J + 1, K + 1)):
        p1[c], p2[c] = p1[s_], p2[s_]
    return p1.reshape(-1), p2.reshape(-1), off


def _flag_diffs(p1, p2, nd, N, thr, c_ref):
    """flags that glibc's and numpy's atan2 would decide differently"""
    ref = {tuple(x) for x in c_ref}
    libm = np.array([math.atan2(a, b) for a, b in zip(p1, p2)])
    out = []
    for ax in (libm, np.arctan2(p1, p2)):
        _, c = O.cores(ax, nd, N, thr=thr)
        out.append(len(ref ^ {tuple(x) for x in c}))
    return out


@pytest.mark.parametrize("thr", [1, 5])
@pytest.mark.parametrize("members", [None, [256, 128, 256]])
def test_strings_2d_sums_at_the_threshold(thr, members):
    acc = FLOAT_ACCEPT[thr]
    Ns = members or [256]
    lat = axs.Lattice(2, Ns[0] if members is None else 0, thr=thr, members=members)
    fields = []
    for r, N in enumerate(Ns):
        p1, p2, off = _margin_field_2d(N, 100 * thr + r, acc)
        assert np.count_nonzero(~np.isnan(off)) > 0.95 * off.size
        z = np.zeros(N * N)
        lat.set_state(p1, p2, z, z, member=r)
        fields.append((p1, p2))
    npl, ncores = lat.string_counts()
    for r, (N, (p1, p2)) in enumerate(zip(Ns, fields)):
        n_ref, c_ref = O.cores(O.axion(p1, p2), 2, N, thr=thr)
        n, c = lat.strings(member=r)
        assert n == n_ref and npl[r] == n_ref
        np.testing.assert_array_equal(c, c_ref)
        assert ncores[r] == O.num_cores2(c_ref)
        if r == 0:
            d_libm, d_np = _flag_diffs(p1, p2, 2, N, thr, c_ref)
            print(f"2D N={N} thr={thr}: {n_ref} flagged; glibc angles would flip {d_libm}, "
                  f"numpy angles {d_np}")


@pytest.mark.parametrize("thr", [1, 5])
@pytest.mark.parametrize("N", [64, 96])
def test_strings_3d_sums_at_the_threshold(thr, N):
    """every engineered site goes through the one-pass finder's FP64
    fallback (its float screen cannot decide a sum within ulps of the
    threshold) and must come out as the oracle decides"""
    acc = FLOAT_ACCEPT[thr]
    p1, p2, off = _margin_field_3d(N, 7 * thr + N, acc)
    assert np.count_nonzero(~np.isnan(off)) > 0.95 * off.size
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N, thr=thr)
    lat.set_state(p1, p2, z, z, tau=1.0)
    n_ref, c_ref = O.cores(O.axion(p1, p2), 3, N, thr=thr)
    n, c = lat.strings()
    assert n == n_ref
    np.testing.assert_array_equal(c, c_ref)
    # the engineered sites: flagged exactly where the correctly rounded sum
    # of plaquette a is above the threshold
    I, J, K = np.meshgrid(*(np.arange(0, N, 2),) * 3, indexing="ij")
    flagged = {tuple(x) for x in c_ref}
    eng = [(i, j, k) for i, j, k in zip(I.ravel(), J.ravel(), K.ravel())]
    above = (off.reshape(-1) > 0)
    got = np.array([s in flagged for s in eng])
    ok = ~np.isnan(off.reshape(-1))
    np.testing.assert_array_equal(got[ok], above[ok])
    d_libm, d_np = _flag_diffs(p1, p2, 3, N, thr, c_ref)
    print(f"3D N={N} thr={thr}: {n_ref} flagged ({int(above.sum())} engineered above); "
          f"glibc angles would flip {d_libm}, numpy angles {d_np}")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_background_angle_each_value(prec):
    """The |atan2(phi1, phi2)| term of the background sums (s_run.cpp:145),
    value by value: each member of an ensemble has one site at (a, b) and
    every other site at (0, 1), whose angle is exactly 0, so the member's
    angle sum is the device's |atan2(a, b)| itself.  Each must lie within
    two ulps (and 4.5e-16) of the correctly rounded |atan2|; pairs cover
    every octant, ratios at the arctangent table's midpoints and both
    slice layouts (sides 64 and 16)."""
    rng = np.random.default_rng(2024)
    M = 384
    Ns = [64] * (M // 2) + [16] * (M // 2)
    a = rng.standard_normal(M) * np.exp(rng.uniform(-5, 5, M))
    b = rng.standard_normal(M) * np.exp(rng.uniform(-5, 5, M))
    q = M // 3
    b[:q] = rng.choice([-1, 1], q) * rng.uniform(0.1, 10, q)
    a[:q] = b[:q] * (rng.integers(0, 256, q) + 0.5) / 256 * rng.choice([-1, 1], q)
    if prec == "fp32":
        a, b = a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
    lat = axs.Lattice(2, members=Ns, precision=prec)
    for r, N in enumerate(Ns):
        p1, p2 = np.zeros(N * N), np.ones(N * N)
        p1[0], p2[0] = a[r], b[r]
        z = np.zeros(N * N)
        lat.set_state(p1, p2, z, z, member=r)
    got = np.array([lat.background(member=r)[4] for r in range(M)])
    want = np.abs(O.atan2_cr(a, b))
    err = np.abs(got - want)
    assert err.max() <= 4.5e-16, (err.max(), a[err.argmax()], b[err.argmax()])
    assert np.all(err <= 2 * np.spacing(want))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_background_norm_each_value(prec):
    """The |phi| term of the background sums (s_run.cpp:146), value by
    value: each member of an ensemble has one site at (a, b) and every other
    site at (0, 0), so the member's |phi| sum is the device's
    sqrt(a^2 + b^2) itself.  Each must lie within one ulp of the root of the
    exact sum of squares (long double, then rounded); zeros, tiny and huge
    magnitudes and both slice layouts (sides 64 and 16) included."""
    rng = np.random.default_rng(2025)
    M = 384
    Ns = [64] * (M // 2) + [16] * (M // 2)
    a = rng.standard_normal(M) * np.exp(rng.uniform(-30, 30, M))
    b = rng.standard_normal(M) * np.exp(rng.uniform(-30, 30, M))
    a[:8], b[:8] = 0.0, 0.0
    b[8:16] = 0.0
    a[16:24] = b[16:24] * (1 + rng.uniform(-1e-6, 1e-6, 8))
    if prec == "fp32":
        a, b = a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
    lat = axs.Lattice(2, members=Ns, precision=prec)
    for r, N in enumerate(Ns):
        p1, p2 = np.zeros(N * N), np.zeros(N * N)
        p1[0], p2[0] = a[r], b[r]
        z = np.zeros(N * N)
        lat.set_state(p1, p2, z, z, member=r)
    got = np.array([lat.background(member=r)[5] for r in range(M)])
    ld = np.longdouble
    want = np.sqrt(ld(a) * ld(a) + ld(b) * ld(b)).astype(np.float64)
    assert np.all(got[:8] == 0.0)
    err = np.abs(got - want)
    assert np.all(err <= np.spacing(want)), (err / np.spacing(want)).max()
