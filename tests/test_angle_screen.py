"""The FP32 angle of the 3D string finder's screen (angle_turns_f in
axion-strings_amd/csrc/axs_strings.hip): its coefficients are read from the
source and its float evaluation is emulated on the host (products exact,
each fused multiply-add rounded once, the hardware reciprocal taken one ulp
off either way), then compared with the exact scaled angle
(atan2(x, y) + pi) / (2 pi) of the float inputs (string_finding.cpp:14-16,
s_run.cpp:94) over random magnitudes and near-axis / near-diagonal inputs.
The screen's correctness rests on this bound: the documented 1.2e-7 turns
per angle keeps a plaquette sum within 2e-6 of its FP64 value, far inside
the screen's 2^-14 margin."""
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(__file__), "..", "axion-strings_amd", "csrc",
                   "axs_strings.hip")
f32 = np.float32


def _coefficients():
    text = open(SRC).read()
    m = re.search(r"kAtanC\[7\]\s*=\s*\{([^}]*)\}", text)
    assert m, "kAtanC not found"
    return [f32(float(v.strip().rstrip("f"))) for v in m.group(1).split(",")]


def _fma(a, b, c):
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)


def _turns(x, y, c, ulp):
    ax, ay = np.abs(x), np.abs(y)
    mx, mn = np.maximum(ax, ay), np.minimum(ax, ay)
    r = (1.0 / mx.astype(np.float64)).astype(np.float32)
    if ulp:
        r = np.nextafter(r, f32(np.inf) if ulp > 0 else f32(0))
    t = (mn * r).astype(np.float32)
    s = (t * t).astype(np.float32)
    p = np.full_like(t, c[6])
    for k in (5, 4, 3, 2, 1, 0):
        p = _fma(p, s, c[k])
    p = (p * t).astype(np.float32)
    p = np.where(ax > ay, (f32(np.pi / 2) - p).astype(np.float32), p)
    p = np.where(y < 0, (f32(np.pi) - p).astype(np.float32), p)
    p = np.where(x < 0, -p, p)
    return _fma(p, f32(1 / (2 * np.pi)), f32(0.5))


def test_screen_angle_error_bound():
    c = _coefficients()
    rng = np.random.default_rng(2024)
    worst = 0.0
    for trial in range(8):
        n = 400_000
        mag = 10.0 ** rng.uniform(-29, 29, n)
        ang = rng.uniform(-np.pi, np.pi, n)
        x = (mag * np.sin(ang)).astype(np.float32)
        y = (mag * np.cos(ang)).astype(np.float32)
        if trial % 4 == 1:
            y = (x * (1 + rng.uniform(-1e-6, 1e-6, n))).astype(np.float32)
        elif trial % 4 == 2:
            x = (y * rng.uniform(-1e-7, 1e-7, n)).astype(np.float32)
        elif trial % 4 == 3:
            y = (x * rng.uniform(-1e-7, 1e-7, n)).astype(np.float32)
        # the screen's trusted inputs (others take the FP64 path)
        ok = (np.maximum(np.abs(x), np.abs(y)) >= 1e-30) & (np.abs(x) <= 1e30) & \
             (np.abs(y) <= 1e30)
        x, y = x[ok], y[ok]
        exact = (np.arctan2(x.astype(np.float64), y.astype(np.float64)) + np.pi) / (2 * np.pi)
        for ulp in (0, 1, -1):
            d = np.abs(_turns(x, y, c, ulp).astype(np.float64) - exact)
            worst = max(worst, float(np.minimum(d, 1 - d).max()))
    assert worst < 1.2e-7, worst


def test_screen_angle_quadrants_and_axes():
    # the octant folding: exact axis and diagonal directions land on the
    # closed-form angles (atan2(x, y) is the direction of (y, x))
    c = _coefficients()
    x = np.array([0, 1, 1, 1, 0, -1, -1, -1, 1e-3, -1e-3], np.float32)
    y = np.array([1, 1, 0, -1, -1, -1, 0, 1, -1, -1], np.float32)
    want = (np.arctan2(x.astype(np.float64), y.astype(np.float64)) + np.pi) / (2 * np.pi)
    got = _turns(x, y, c, 0).astype(np.float64)
    d = np.abs(got - want)
    assert np.all(np.minimum(d, 1 - d) < 1.2e-7), (got, want)
