"""The correctly rounded atan2 behind the axion field (atan2(phi1, phi2),
s_run.cpp:92-94) and the string finder's angles.

Two implementations: the oracle's (orc_atan2_cr: x87 extended-precision
atan2l rounded to double, libquadmath's atan2q near rounding midpoints) and
the library's double-double evaluation (csrc/axs_atan2.hpp), whose host
build (libaxs_host.so) has the same bits as its device build.  Both are
pinned to mpmath's atan2 at 200 bits rounded to nearest, on random, wide-
range and edge arguments, and to C99 Annex F on special values; they agree
with each other bit for bit on 10^7 random arguments.  glibc's atan2 is
characterised beside them: it is within one ulp, but not correctly rounded
on about 0.1% of random arguments -- the reason the oracle fixes the
correctly rounded value."""
import ctypes
import math

import numpy as np
import pytest

import axs
import oracle as O

mp = pytest.importorskip("mpmath")
from mpmath import libmp  # noqa: E402


def _cr(y, x):
    with mp.workprec(200):
        v = mp.atan2(mp.mpf(float(y)), mp.mpf(float(x)))
        return libmp.to_float(v._mpf_, rnd="n")


def _dd(y, x):
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


def _args(rng, n):
    """random, wide-range, near-axis, near-diagonal and tiny-ratio pairs"""
    parts = [
        (rng.standard_normal(n), rng.standard_normal(n)),
        (np.ldexp(rng.uniform(0.5, 1, n) * rng.choice([-1, 1], n), rng.integers(-1000, 1000, n)),
         np.ldexp(rng.uniform(0.5, 1, n) * rng.choice([-1, 1], n), rng.integers(-1000, 1000, n))),
    ]
    x = rng.standard_normal(n)
    parts.append((x * (1 + rng.uniform(-1e-3, 1e-3, n)), x))                 # |y| ~ |x|
    parts.append((x * np.ldexp(1.0, rng.integers(-60, -20, n)), x))          # near the x axis
    parts.append((x, x * np.ldexp(1.0, rng.integers(-60, -20, n))))          # near the y axis
    parts.append((x, -x * (1 + rng.uniform(-1e-6, 1e-6, n))))                # near +-3pi/4
    y = np.concatenate([p[0] for p in parts])
    x = np.concatenate([p[1] for p in parts])
    return y, x


@pytest.mark.parametrize("seed", [1, 2])
def test_both_match_mpmath(seed):
    rng = np.random.default_rng(seed)
    y, x = _args(rng, 3000)
    want = np.array([_cr(a, b) for a, b in zip(y, x)])
    np.testing.assert_array_equal(O.atan2_cr(y, x), want)
    np.testing.assert_array_equal(_dd(y, x), want)


SPECIAL = [0.0, -0.0, 1.0, -1.0, 0.5, -3.0, 5e-324, -5e-324, 2.2e-308, 1e300, -1e300,
           1.7976931348623157e308, math.inf, -math.inf, math.nan]


def test_special_values_c99_annex_f():
    ys, xs = np.meshgrid(SPECIAL, SPECIAL, indexing="ij")
    y, x = ys.ravel(), xs.ravel()
    libm = np.array([math.atan2(a, b) for a, b in zip(y, x)])
    for got in (O.atan2_cr(y, x), _dd(y, x)):
        nan = np.isnan(libm)
        assert np.array_equal(np.isnan(got), nan)
        np.testing.assert_array_equal(got[~nan], libm[~nan])
        # signed zeros: atan2(+-0, x >= +0) = +-0
        np.testing.assert_array_equal(np.signbit(got[~nan]), np.signbit(libm[~nan]))
    # finite non-zero pairs among them: correctly rounded as well
    fin = np.isfinite(y) & np.isfinite(x) & (y != 0) & (x != 0)
    want = np.array([_cr(a, b) for a, b in zip(y[fin], x[fin])])
    np.testing.assert_array_equal(_dd(y[fin], x[fin]), want)


def test_subnormal_and_extreme_ratios():
    rng = np.random.default_rng(9)
    n = 2000
    y = np.ldexp(rng.uniform(0.5, 1, n), rng.integers(-1074, -1000, n))     # subnormal or tiny
    x = np.ldexp(rng.uniform(0.5, 1, n), rng.integers(-1074, 1024, n))
    keep = (y != 0) & np.isfinite(x)
    y, x = y[keep], x[keep]
    a, b = O.atan2_cr(y, x), _dd(y, x)
    np.testing.assert_array_equal(a, b)
    # normal results: against mpmath (its conversion of subnormal results
    # rounds twice, so those are compared between the two implementations)
    norm = np.abs(a) >= 2.2250738585072014e-308
    want = np.array([_cr(p, q) for p, q in zip(y[norm], x[norm])])
    np.testing.assert_array_equal(a[norm], want)
    # and the swapped octants
    np.testing.assert_array_equal(O.atan2_cr(x, y), _dd(x, y))
    np.testing.assert_array_equal(O.atan2_cr(-y, -x), _dd(-y, -x))


def test_oracle_and_library_agree_on_1e7_random():
    rng = np.random.default_rng(2026)
    n = 10_000_000
    y, x = rng.standard_normal(n), rng.standard_normal(n)
    np.testing.assert_array_equal(O.atan2_cr(y, x), _dd(y, x))


def test_glibc_atan2_is_not_correctly_rounded():
    """glibc's atan2 (what math.atan2 calls) is within one ulp, and on about
    0.1% of random arguments one ulp off the correctly rounded value"""
    rng = np.random.default_rng(7)
    n = 200_000
    y, x = rng.standard_normal(n), rng.standard_normal(n)
    libm = np.array([math.atan2(a, b) for a, b in zip(y, x)])
    cr = O.atan2_cr(y, x)
    off = np.flatnonzero(libm != cr)
    print(f"glibc atan2: {off.size} of {n} arguments not correctly rounded")
    assert off.size < 0.005 * n
    assert np.all(np.abs(libm[off] - cr[off]) <= np.spacing(np.abs(cr[off])))
    for i in off[:20]:
        assert _cr(y[i], x[i]) == cr[i]


def _fast(y, x):
    h = axs.load_host_library()
    f = h.axs_host_abs_atan2_fast
    f.argtypes = [ctypes.c_int64] + [ctypes.POINTER(ctypes.c_double)] * 3
    f.restype = None
    y = np.ascontiguousarray(y, np.float64).reshape(-1)
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    out = np.empty_like(y)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    f(y.size, P(y), P(x), P(out))
    return out


@pytest.mark.parametrize("seed", [3, 4])
def test_abs_atan2_fast_bound(seed):
    """The |atan2| of the background sums (table reduction in plain double):
    within 4.5e-16 and two ulps of the correctly rounded |atan2| on random,
    wide-range, near-axis, near-diagonal and near-table-midpoint arguments,
    and on FP32 values (the FP32 form's phi); most results are exact."""
    rng = np.random.default_rng(seed)
    y, x = _args(rng, 100000)
    n = 100000
    xm = rng.standard_normal(n)
    mid = (np.arange(n) % 256 + 0.5) / 256 * (1 + rng.uniform(-1e-12, 1e-12, n))
    y = np.concatenate([y, xm * mid, rng.standard_normal(n).astype(np.float32)])
    x = np.concatenate([x, xm, rng.standard_normal(n).astype(np.float32)])
    got = _fast(y, x)
    want = np.abs(O.atan2_cr(y, x))
    err = np.abs(got - want)
    assert err.max() <= 4.5e-16
    assert np.all(err <= 2 * np.spacing(want))
    assert (err == 0).mean() > 0.6


def test_abs_atan2_f32in_matches_fast():
    """The float-argument shortcut of the FP32 background sums gives
    abs_atan2_fast's bits: random, wide-range and near-table-midpoint float
    pairs, and every pair of float special values (zeros, infinities, NaN,
    the smallest subnormal, the largest finite)."""
    h = axs.load_host_library()
    f = h.axs_host_abs_atan2_f32in
    f.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                  ctypes.POINTER(ctypes.c_double)]
    f.restype = None
    rng = np.random.default_rng(8)
    n = 200000
    y = (rng.standard_normal(n) * np.exp(rng.uniform(-80, 80, n))).astype(np.float32)
    x = (rng.standard_normal(n) * np.exp(rng.uniform(-80, 80, n))).astype(np.float32)
    xm = rng.standard_normal(n).astype(np.float32)
    mid = ((np.arange(n) % 256 + 0.5) / 256).astype(np.float32)
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4028235e38, 1.0, -1.0],
                  np.float32)
    ys, xs = np.meshgrid(sp, sp, indexing="ij")
    y = np.ascontiguousarray(np.concatenate([y, xm * mid, ys.ravel()]), np.float32)
    x = np.ascontiguousarray(np.concatenate([x, xm, xs.ravel()]), np.float32)
    out = np.empty(y.size)
    f(y.size, y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
      x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    want = _fast(y.astype(np.float64), x.astype(np.float64))
    np.testing.assert_array_equal(out, want)


def test_abs_atan2_fast_special_values():
    """C99 Annex F values exactly where a component is infinite or NaN or
    both are zero (NaN for NaN); the other pairs of the grid (one zero,
    subnormal, tiny or huge components -- scaled into range first -- on the
    axes and diagonals) within the bound."""
    ys, xs = np.meshgrid(SPECIAL, SPECIAL, indexing="ij")
    y, x = ys.ravel(), xs.ravel()
    got = _fast(y, x)
    want = np.abs(_dd(y, x))
    special = ~np.isfinite(y) | ~np.isfinite(x) | ((y == 0) & (x == 0))
    assert special.sum() > 0 and (~special).sum() > 0
    np.testing.assert_array_equal(got[special], want[special])
    err = np.abs(got[~special] - want[~special])
    assert np.all(err <= 2 * np.spacing(want[~special]))


def test_abs_atan2_fast_extreme_magnitudes():
    """Pairs whose larger magnitude lies outside [2^-1000, 2^1000] (scaled
    by a common power of two) against the correctly rounded value."""
    rng = np.random.default_rng(11)
    n = 20000
    e = rng.choice(np.r_[np.arange(-1074, -1000), np.arange(1000, 1024)], n)
    with np.errstate(over="ignore"):
        y = np.ldexp(rng.uniform(0.5, 1, n) * rng.choice([-1, 1], n), e)
        x = np.ldexp(rng.uniform(0.5, 1, n) * rng.choice([-1, 1], n), e - rng.integers(-40, 40, n))
    ok = np.isfinite(x) & np.isfinite(y)
    y, x = y[ok], x[ok]
    got = _fast(y, x)
    want = np.abs(O.atan2_cr(y, x))
    assert np.all(np.abs(got - want) <= np.maximum(2 * np.spacing(want), 1e-300))
