"""Pins the oracle's stencil matrix and SpMV (restating
build_coefficient_matrix, s_integrate.cpp:5-149, and the MKL SpMV of
s_integrate.cpp:322-324) against the reference's own test data and closed
forms."""
import json
import os

import numpy as np
import pytest

import oracle as O


def test_matrix_matches_reference_fixture(root):
    # reference src/tests/fdm_matrix.cpp:47-64: exact 16x16 operator, 2D N=4
    fx = json.load(open(os.path.join(root, "tests", "golden", "fdm_matrix_N4.json")))
    want = np.array(fx["matrix"], dtype=np.float64)
    got = O.Matrix(2, 4).dense()
    np.testing.assert_array_equal(got, want)


def test_matrix_nnz():
    for nd, N, st, per in ((2, 8, 0, 5), (3, 6, 0, 7), (2, 16, 3, 17), (3, 8, 2, 19)):
        assert O.Matrix(nd, N, st).nnz == per * N ** nd


@pytest.mark.parametrize("ndims,N", [(2, 128), (3, 128)])
def test_zero_field(ndims, N):
    # tests/laplacian.cpp:15-24 / fdm_matrix.cpp:119-127: zero -> 0
    A = O.Matrix(ndims, N)
    y = A.spmv(np.zeros(N ** ndims))
    assert np.all(np.abs(y) <= 1e-30)


def test_plane_wave_2d():
    # fdm_matrix.cpp:139-173 (and laplacian.cpp:26-45): N=256, dx=1,
    # phi = sin(pi i / L) + cos(pi j / L), L = N/2, within 1e-5
    N = 256
    L = N / 2.0
    ii = np.arange(N)
    phi = (np.sin(np.float32(np.pi / L) * ii.astype(np.float32)).astype(np.float64)[None, :]
           + np.cos(np.float32(np.pi / L) * ii.astype(np.float32)).astype(np.float64)[:, None])
    lap = O.Matrix(2, N).spmv(phi.reshape(-1), alpha=1.0).reshape(N, N)
    exact = -(np.pi / L) ** 2 * phi
    assert np.abs(lap - exact).max() <= 1e-5
    # inline laplacian2D of s_internal.hpp:26-28 agrees at sampled sites
    flat = np.ascontiguousarray(phi.reshape(-1))
    for (i, j) in ((0, 0), (5, 255), (128, 17), (255, 255)):
        v = O.lib().orc_laplacian2D(flat.ctypes.data_as(O._DP), i, j, 1.0, N)
        assert abs(v - lap[j, i]) <= 1e-12


def test_plane_wave_3d():
    # fdm_matrix.cpp:205-243: phi = sin(pi i/L) + cos(pi j/L) + sin(pi k/L)
    # (run at N=128; the discretisation error (pi/L)^4/12 * 3 = 1.4e-6 < 1e-5)
    N = 128
    L = N / 2.0
    a = np.arange(N)
    s = np.sin(np.pi / L * a)
    c = np.cos(np.pi / L * a)
    phi = s[:, None, None] + c[None, :, None] + s[None, None, :]
    lap = O.Matrix(3, N).spmv(phi.reshape(-1), alpha=1.0).reshape(N, N, N)
    exact = -(np.pi / L) ** 2 * phi
    assert np.abs(lap - exact).max() <= 1e-5
    flat = np.ascontiguousarray(phi.reshape(-1))
    for (i, j, k) in ((0, 0, 0), (127, 3, 64), (1, 127, 127)):
        v = O.lib().orc_laplacian3D(flat.ctypes.data_as(O._DP), i, j, k, 1.0, N)
        assert abs(v - lap[i, j, k]) <= 1e-12


@pytest.mark.parametrize("ndims,N,modes", [(2, 32, (3, 5)), (3, 16, (1, 4, 7)), (2, 12, (4, 6))])
def test_discrete_eigenvalue(ndims, N, modes):
    # cos(2 pi (n . x)/N) is an eigenvector of the periodic 5/7-point operator
    # with eigenvalue -(4/dx^2) sum_d sin^2(pi n_d / N); alpha = 1/dx^2 = 4
    dx = 0.5
    grids = np.meshgrid(*[np.arange(N)] * ndims, indexing="ij")
    if ndims == 2:   # array [j, i]: first mode along i (fast axis)
        theta = 2 * np.pi * (modes[0] * grids[1] + modes[1] * grids[0]) / N
    else:
        theta = 2 * np.pi * sum(m * g for m, g in zip(modes, grids)) / N
    phi = np.cos(theta)
    lam = -(4 / dx ** 2) * sum(np.sin(np.pi * m / N) ** 2 for m in modes)
    y = O.Matrix(ndims, N).spmv(phi.reshape(-1), alpha=1.0 / dx ** 2)
    assert np.abs(y - lam * phi.reshape(-1)).max() <= 1e-12 * max(1.0, abs(lam))


@pytest.mark.parametrize("stencil", [1, 2, 3])
def test_higher_order_eigenvalue(stencil):
    # stencil_setting 1..3 (s_integrate.cpp:11-13): eigenvalue of a cosine
    # mode is c0 + sum_axes sum_l 2 c_l cos(l theta_axis), coefficients in
    # float as the reference's table
    f = np.float32
    tab = {1: [f(-5) / f(2), f(4) / f(3), f(-1) / f(12)],
           2: [f(-49) / f(18), f(3) / f(2), f(-3) / f(20), f(1) / f(90)],
           3: [f(-205) / f(72), f(8) / f(5), f(-1) / f(5), f(8) / f(315), f(-1) / f(560)]}[stencil]
    N, ndims, modes = 24, 3, (2, 5, 9)
    c0 = float(tab[0] * f(ndims))
    cl = [float(x) for x in tab[1:]]
    grids = np.meshgrid(*[np.arange(N)] * ndims, indexing="ij")
    thetas = [2 * np.pi * m / N for m in modes]
    phi = np.cos(sum(t * g for t, g in zip(thetas, grids)))
    lam = c0 + sum(2 * c * np.cos((l + 1) * t) for t in thetas for l, c in enumerate(cl))
    y = O.Matrix(ndims, N, stencil).spmv(phi.reshape(-1))
    assert np.abs(y - lam * phi.reshape(-1)).max() <= 1e-12
