"""Pins the oracle's FFT and power spectrum (output_powerspec,
pk_estimation.cpp:37-214), the background sums (s_run.cpp:135-150) and the
energy components."""
import numpy as np
import pytest

import oracle as O
from fields import kinv_field


@pytest.mark.parametrize("ndims,N", [(2, 8), (2, 12), (3, 6), (3, 16), (2, 30)])
def test_fft_matches_numpy(ndims, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N,) * ndims) + 1j * rng.standard_normal((N,) * ndims)
    got = O.fft(x, ndims, N, -1).reshape(x.shape)
    np.testing.assert_allclose(got, np.fft.fftn(x), rtol=0, atol=1e-10 * N ** ndims)
    back = O.fft(got, ndims, N, +1).reshape(x.shape) / N ** ndims
    np.testing.assert_allclose(back, x, rtol=0, atol=1e-12)


def _brute_counts(ndims, N):
    # count lattice modes with 0 < |k| < N/2 per bin floor(|k|) - 1
    f = np.arange(N) - N // 2
    g = np.meshgrid(*[f] * ndims, indexing="ij")
    k = np.sqrt(sum(x.astype(np.float64) ** 2 for x in g)).reshape(-1)
    sel = (k > 0) & (k < N // 2)
    b = np.floor(k[sel]).astype(int) - 1
    return np.bincount(b, minlength=N // 2)[: N // 2]


@pytest.mark.parametrize("ndims,N", [(2, 32), (3, 16)])
def test_mode_counts(ndims, N):
    p1, _ = kinv_field(ndims, N, seed=2)
    ks, pk, cnt = O.powerspec(p1, ndims, N, 0.5)
    np.testing.assert_array_equal(cnt, _brute_counts(ndims, N))


@pytest.mark.parametrize("ndims,N,mode", [(2, 32, (5, 0)), (2, 32, (3, 4)), (3, 16, (1, 2, 2))])
def test_single_cosine(ndims, N, mode):
    # phi = cos(2 pi n.x/N): power N^(2d)/4 in each of the two modes +-n,
    # both at |k| = |n|; bin floor(|n|) - 1 holds N^(2d)/2 over count modes
    dx = 0.5
    grids = np.meshgrid(*[np.arange(N)] * ndims, indexing="ij")
    if ndims == 2:
        theta = 2 * np.pi * (mode[0] * grids[1] + mode[1] * grids[0]) / N
    else:
        theta = 2 * np.pi * sum(m * g for m, g in zip(mode, grids)) / N
    phi = np.cos(theta).reshape(-1)
    ks, pk, cnt = O.powerspec(phi, ndims, N, dx)
    kn = np.sqrt(sum(m * m for m in mode))
    b = int(np.floor(kn)) - 1
    cnt_b = _brute_counts(ndims, N)[b]
    # mean |k| over the bin is data independent; recover the summed power
    kbar = ks[b] / (2 * np.pi / (N * dx))
    total = pk[b] * (2 * np.pi * kbar if ndims == 2 else 4 * np.pi * kbar ** 2)
    assert abs(total - N ** (2 * ndims) / 2) < 1e-8 * N ** (2 * ndims)
    others = [i for i in range(N // 2) if i != b and cnt[i] > 0]
    assert np.all(np.abs(pk[others]) < 1e-12 * N ** (2 * ndims))
    assert cnt[b] == cnt_b


def test_log_bins_parseval():
    # sum over log bins of count * mean|F|^2 plus the excluded modes
    # (k = 0 and |k| >= N/2) equals N^d sum |phi|^2
    ndims, N, nb = 3, 16, 12
    p1, _ = kinv_field(ndims, N, seed=4)
    ks, pk, cnt = O.powerspec_log(p1, ndims, N, 0.5, nb)
    F = np.fft.fftn(p1.reshape((N,) * ndims))
    f = np.fft.fftfreq(N, 1.0 / N)
    k2 = sum(x ** 2 for x in np.meshgrid(*[f] * ndims, indexing="ij"))
    excl = np.sum(np.abs(F[(k2 == 0) | (k2 >= (N // 2) ** 2)]) ** 2)
    total = np.sum(cnt * pk) + excl
    assert abs(total - N ** ndims * np.sum(p1 ** 2)) < 1e-9 * total
    # edges: bin b holds modes with (N/2)^(2b/nb) <= |k|^2 < (N/2)^(2(b+1)/nb)
    e2 = (N // 2) ** (2.0 * np.arange(nb + 1) / nb)
    kk = k2[(k2 >= 1) & (k2 < (N // 2) ** 2)]
    want = np.histogram(kk, bins=np.concatenate([e2[:-1], [(N // 2) ** 2]]))[0]
    np.testing.assert_array_equal(cnt, want)


def test_background_closed_forms():
    L = 1000
    a, b = np.full(L, 0.6), np.full(L, -0.8)
    out = O.background(a, b, np.full(L, 0.25), np.full(L, -2.0))
    np.testing.assert_allclose(out[:4], [0.6 * L, -0.8 * L, 0.25 * L, -2.0 * L], rtol=1e-13)
    assert abs(out[4] - L * abs(np.arctan2(0.6, -0.8))) < 1e-10
    assert abs(out[5] - L * 1.0) < 1e-10
    # antisymmetric field: phi and -phi cancel in the linear sums
    p1, p2 = kinv_field(2, 16, seed=9)
    s = O.background(np.concatenate([p1, -p1]), np.concatenate([p2, -p2]),
                     np.zeros(2 * p1.size), np.zeros(2 * p1.size))
    assert abs(s[0]) < 1e-12 and abs(s[1]) < 1e-12


def test_energy_uniform_field():
    N, nd = 8, 3
    L = N ** nd
    p = O.params(nd, N, lam=1.0, T_initial=4.0)
    tau = 1.7
    e = O.energy(p, tau, np.full(L, 0.3), np.full(L, 0.4), np.full(L, 1.5), np.full(L, -0.5))
    vol = 0.5 ** 3
    assert abs(e[0] - 0.5 * (1.5 ** 2 + 0.5 ** 2) * L * vol) < 1e-10
    assert e[1] == 0.0
    s = 0.25 - tau ** 2 + float(np.float32(4.0)) ** 2 / 3
    assert abs(e[2] - 1.0 / (4 * tau ** 2) * s * s * L * vol) < 1e-9


def test_energy_plane_wave_gradient():
    # sum_i (cos(th(i+1)) - cos(th i))^2 = 2 N sin^2(th/2) for th = 2 pi m/N
    N, m, dx = 32, 3, 0.5
    th = 2 * np.pi * m / N
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    phi = np.cos(th * i).reshape(-1)
    p = O.params(2, N, space_step=dx)
    e = O.energy(p, 1.0, phi, np.zeros(N * N), np.zeros(N * N), np.zeros(N * N))
    want = 0.5 * N * (2 * N * np.sin(th / 2) ** 2) / dx ** 2 * dx ** 2
    assert abs(e[1] - want) < 1e-10 * want


def test_energy_variation_is_the_force():
    # d(E_grad + E_pot)/d phi1(x) = -vol * ker1(x), ker = Lap phi - V'(phi):
    # the step's force is the gradient of this energy
    N, nd, dx = 8, 3, 0.5
    p = O.params(nd, N, lam=1.0, T_initial=4.0, space_step=dx)
    rng = np.random.default_rng(11)
    p1, p2 = rng.standard_normal(N ** nd), rng.standard_normal(N ** nd)
    tau = 1.3
    L = N ** nd
    z = np.zeros(L)
    A = O.Matrix(nd, N)
    s = p1 ** 2 + p2 ** 2 - tau ** 2 + float(np.float32(4.0)) ** 2 / 3
    ker1 = A.spmv(p1, 1 / dx ** 2) - 1.0 / tau ** 2 * p1 * s
    vol = dx ** nd
    for site in (0, 77, L - 1):
        h = 1e-5
        a = p1.copy(); a[site] += h
        b = p1.copy(); b[site] -= h
        ea = O.energy(p, tau, a, p2, z, z)
        eb = O.energy(p, tau, b, p2, z, z)
        d = ((ea[1] + ea[2]) - (eb[1] + eb[2])) / (2 * h)
        assert abs(d + vol * ker1[site]) < 1e-6 * max(1.0, abs(vol * ker1[site]))


@pytest.mark.parametrize("ndims,N,mode", [(2, 32, (5, 0)), (2, 32, (3, -4)), (3, 16, (1, -2, 2))])
def test_single_complex_exponential(ndims, N, mode):
    # phi = exp(i 2 pi n.x/N) (re = cos, im = sin): all power N^(2d) in the
    # one mode n -- not split between +-n as for a real cosine -- so its bin
    # holds N^(2d) in linear and in log binning, every other bin nothing
    dx = 0.5
    grids = np.meshgrid(*[np.arange(N)] * ndims, indexing="ij")
    if ndims == 2:
        theta = 2 * np.pi * (mode[0] * grids[1] + mode[1] * grids[0]) / N
    else:
        theta = 2 * np.pi * sum(m * g for m, g in zip(mode, grids)) / N
    re, im = np.cos(theta).reshape(-1), np.sin(theta).reshape(-1)
    ks, pk, cnt = O.powerspec(re, ndims, N, dx, field_imag=im)
    kn = np.sqrt(sum(m * m for m in mode))
    b = int(np.floor(kn)) - 1
    kbar = ks[b] / (2 * np.pi / (N * dx))
    total = pk[b] * (2 * np.pi * kbar if ndims == 2 else 4 * np.pi * kbar ** 2)
    assert abs(total - N ** (2 * ndims)) < 1e-8 * N ** (2 * ndims)
    others = [i for i in range(N // 2) if i != b and cnt[i] > 0]
    assert np.all(np.abs(pk[others]) < 1e-12 * N ** (2 * ndims))
    ks, pk, cnt = O.powerspec_log(re, ndims, N, dx, 10, field_imag=im)
    full = pk * cnt
    bl = int(np.argmax(full))
    assert abs(full[bl] - N ** (2 * ndims)) < 1e-8 * N ** (2 * ndims)
    assert np.all(np.abs(np.delete(full, bl)) < 1e-12 * N ** (2 * ndims))


@pytest.mark.parametrize("ndims,N", [(2, 32), (3, 16)])
def test_complex_spectrum_is_the_sum_of_its_parts(ndims, N):
    # every binned mode has |k| < N/2, so -k is binned with it: for
    # f = a + i b, |F(k)|^2 + |F(-k)|^2 = 2 (|A(k)|^2 + |B(k)|^2), and each bin
    # of the complex field's spectrum is the sum of its parts' bins
    a, b = kinv_field(ndims, N, seed=6)
    for spec in (lambda re, im=None: O.powerspec(re, ndims, N, 0.5, field_imag=im),
                 lambda re, im=None: O.powerspec_log(re, ndims, N, 0.5, 9, field_imag=im)):
        ks, pk, cnt = spec(a, b)
        _, pa, ca = spec(a)
        _, pb, cb = spec(b)
        np.testing.assert_array_equal(cnt, ca)
        # bins above pi/(4 dx) hold only transform rounding: floor 1e-9 of the
        # peak (an empty bin is 0/0 in the reference's normalisation: NaN)
        np.testing.assert_allclose(pk, pa + pb, rtol=1e-12, atol=1e-9 * np.nanmax(pk))
