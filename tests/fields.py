"""Synthetic lattice fields for the tests (numpy; test data only).

Layout: 2D arrays are indexed [j, i] (m = i + N*j), 3D arrays [i, j, k]
(m = (i*N + j)*N + k); .reshape(-1) gives the reference's linear order.
"""
from __future__ import annotations

import numpy as np


def kgrid(ndims: int, N: int):
    """integer wave numbers of an unshifted FFT, broadcastable per axis"""
    f = np.fft.fftfreq(N, 1.0 / N)
    if ndims == 2:
        return [f[:, None], f[None, :]]
    return [f[:, None, None], f[None, :, None], f[None, None, :]]


def kinv_field(ndims: int, N: int, dx: float = 0.5, seed: int = 1998):
    """Complex Gaussian random field with Fourier amplitudes ~ 1/|k| for
    0 < |k| < pi/(4 dx), normalised to <|phi|^2> = 1; returns phi1, phi2 as
    flat float64 arrays."""
    rng = np.random.default_rng(seed)
    shape = (N,) * ndims
    g = [x * (2 * np.pi / (N * dx)) for x in kgrid(ndims, N)]
    k2 = sum(x * x for x in g)
    kmag = np.sqrt(k2)
    kcut = np.pi / (4 * dx)
    amp = np.zeros(shape)
    sel = (kmag > 0) & (kmag < kcut)
    amp[sel] = 1.0 / kmag[sel]
    z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * amp
    if not np.any(sel):
        raise ValueError(f"no lattice mode below k_cut for N={N}, dx={dx}")
    phi = np.fft.ifftn(z)
    norm = np.sqrt(np.mean(np.abs(phi) ** 2))
    phi /= norm
    return np.ascontiguousarray(phi.real).reshape(-1), np.ascontiguousarray(phi.imag).reshape(-1)


def random_state(ndims: int, N: int, seed: int = 7, vel_scale: float = 0.3, ker: bool = False):
    """kinv phi plus random velocities (and optionally random ker arrays)"""
    p1, p2 = kinv_field(ndims, N, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    L = N ** ndims
    v1 = vel_scale * rng.standard_normal(L)
    v2 = vel_scale * rng.standard_normal(L)
    if ker:
        return p1, p2, v1, v2, rng.standard_normal(L), rng.standard_normal(L)
    return p1, p2, v1, v2, None, None


def vortex2d(N: int, centres, windings, width: float = 0.0):
    """phi = prod_r ((x - x_r) + i (y - y_r))^{w_r} / |.|, x = i, y = j;
    returns phi1, phi2 flat (reference layout m = i + N*j)."""
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    z = np.ones((N, N), complex)
    for (cx, cy), w in zip(centres, windings):
        d = (i - cx) + 1j * (j - cy)
        z *= (d / np.abs(d)) ** w if w > 0 else (np.conj(d) / np.abs(d)) ** (-w)
    return np.ascontiguousarray(z.real).reshape(-1), np.ascontiguousarray(z.imag).reshape(-1)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
