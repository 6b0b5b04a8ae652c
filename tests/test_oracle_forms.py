"""Pins the oracle's alternative forms: the spectral Laplacian
(compute_laplacian_fft, s_integrate.cpp:179-296), the field-rescaled step
with it (the #else branch, :326-328) and the Hamiltonian-form step
(vvsl_hamiltonian_form, :358-395)."""
import numpy as np
import pytest

import oracle as O
from fields import kinv_field, random_state
from test_oracle_step import leapfrog_closed_form


def _numpy_spectral_laplacian(f, ndims, N, dx):
    k = 2 * np.pi * np.fft.fftfreq(N, d=dx)          # f = -N/2 .. N/2-1, times 2 pi / L
    grids = np.meshgrid(*[k] * ndims, indexing="ij")
    k2 = sum(g * g for g in grids)
    F = np.fft.fftn(f.reshape([N] * ndims))
    return np.fft.ifftn(-k2 * F).real.reshape(-1)


@pytest.mark.parametrize("ndims,N", [(2, 16), (2, 24), (3, 8), (3, 12)])
def test_laplacian_fft_matches_numpy(ndims, N):
    # numpy's FFT is independent of the oracle's; the (-1)^(i+j+k) shifts
    # of the reference only relabel the modes, so the operator is -k^2 on
    # every mode f in [-N/2, N/2)
    dx = 0.5
    rng = np.random.default_rng(7)
    f = rng.standard_normal(N ** ndims)
    got = O.laplacian_fft(f, ndims, N, dx)
    ref = _numpy_spectral_laplacian(f, ndims, N, dx)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("ndims,N,mode", [(2, 16, (3, 1)), (2, 16, (8, 0)), (3, 8, (1, 2, 4)),
                                          (3, 8, (4, 4, 4))])
def test_laplacian_fft_plane_waves(ndims, N, mode):
    # cosine modes are eigenfunctions with the continuum eigenvalue -k^2,
    # including the Nyquist frequency N/2 (whose k the reference takes as
    # -N/2 * 2 pi / L, same square)
    dx = 0.5
    L = N * dx
    grids = np.meshgrid(*[np.arange(N)] * ndims, indexing="ij")
    if ndims == 2:          # m = i + N*j: the flat index runs over grids[1] fastest
        theta = 2 * np.pi * (mode[0] * grids[1] + mode[1] * grids[0]) / N
    else:
        theta = 2 * np.pi * sum(m * g for m, g in zip(mode, grids)) / N
    f = np.cos(theta).reshape(-1)
    k2 = sum((2 * np.pi * m / L) ** 2 for m in mode)
    got = O.laplacian_fft(f, ndims, N, dx)
    assert np.abs(got + k2 * f).max() < 1e-11 * max(k2, 1.0)


@pytest.mark.parametrize("ndims,N,mode,nsteps", [(2, 32, (3, 1), 200), (3, 16, (1, 2, 3), 100)])
def test_spectral_step_free_plane_wave(ndims, N, mode, nsteps):
    # lambda = 0: a cosine mode oscillates with w^2 = k^2 (continuum) and
    # the step is the same leapfrog as the stencil form (zero initial kernel)
    dx = 0.5
    dt = float(np.float32(0.1))
    grids = np.meshgrid(*[np.arange(N)] * ndims, indexing="ij")
    if ndims == 2:
        theta = 2 * np.pi * (mode[0] * grids[1] + mode[1] * grids[0]) / N
    else:
        theta = 2 * np.pi * sum(m * g for m, g in zip(mode, grids)) / N
    shape = np.cos(theta).reshape(-1)
    w2 = sum((2 * np.pi * m / (N * dx)) ** 2 for m in mode)
    L = N ** ndims
    p = O.params(ndims, N, lam=0.0, T_initial=0.0)
    st = O.State(shape, 0.5 * shape, np.zeros(L), np.zeros(L))
    st.run_form(p, None, nsteps, eom=0, laplacian=1)
    amp = leapfrog_closed_form(w2, dt, nsteps)
    assert np.abs(st.phi1 - amp * shape).max() < 1e-10
    assert np.abs(st.phi2 - 0.5 * amp * shape).max() < 1e-10


def test_spectral_step_equals_stencil_step_on_uniform_fields():
    # both Laplacians vanish on a uniform field (the spectral one to
    # rounding), so the drift, potential, kick and tau bookkeeping of the
    # spectral step must reproduce the pinned stencil step
    N, ndims = 16, 2
    L = N * N
    p = O.params(ndims, N, lam=1.0, T_initial=3.0)
    init = (np.full(L, 0.8), np.full(L, -0.3), np.full(L, 0.2), np.full(L, 0.1),
            np.full(L, 0.05), np.full(L, -0.02))
    a = O.State(*init)
    a.run(p, O.Matrix(ndims, N), 25)
    b = O.State(*init)
    b.run_form(p, None, 25, eom=0, laplacian=1)
    assert a.tau == b.tau
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        np.testing.assert_allclose(getattr(b, k), getattr(a, k), rtol=0, atol=1e-12)


def test_hamiltonian_first_step_drift():
    # tau advances before the drift (:364) and the drift factor is
    # dt / tau^2 (:368): with the zero initial kernel the first step moves
    # phi by dt / (tau0 + dt)^2 * pi exactly
    N = 16
    p1, p2, v1, v2, _, _ = random_state(2, N)
    p = O.params(2, N)
    st = O.State(p1, p2, v1, v2, tau=1.25)
    st.run_form(p, O.Matrix(2, N), 1, eom=1)
    dt = float(np.float32(0.1))
    t = 1.25 + dt
    cd = dt / (t * t)
    np.testing.assert_array_equal(st.phi1, p1 + cd * (v1 + float(np.float32(0.05)) * 0.0))
    np.testing.assert_array_equal(st.phi2, p2 + cd * (v2 + float(np.float32(0.05)) * 0.0))
    assert st.tau == t


def test_hamiltonian_vacuum_is_fixed():
    # T0 = 0: phi = 1 is a vacuum of the Hamiltonian form (|phi|^2 - 1 = 0,
    # no gradients), so phi and pi stay exactly put while tau advances
    N = 8
    L = N ** 3
    p = O.params(3, N, lam=1.0, T_initial=0.0)
    st = O.State(np.ones(L), np.zeros(L), np.zeros(L), np.zeros(L))
    st.run_form(p, O.Matrix(3, N), 50, eom=1)
    assert (st.phi1 == 1.0).all() and (st.phi2 == 0.0).all()
    assert (st.phidot1 == 0.0).all() and (st.phidot2 == 0.0).all()


def test_hamiltonian_form_reproduces_field_rescaled_physics():
    # phi_fr = tau phi_H and pi = tau^2 phi_H' turn the field-rescaled EOM
    # into the Hamiltonian one, so the two schemes approach the same
    # solution: the Hamiltonian step (first order in dt, its drift uses the
    # end-of-step tau) against the pinned field-rescaled step at a fine dt,
    # both from consistent initial kernels
    N, dx, lam, T0 = 16, 0.5, 1.0, 1.0
    p1, p2 = kinv_field(2, N, seed=5)
    L = N * N
    A = O.Matrix(2, N)

    def force_fr(f1, f2, tau):
        s = f1 ** 2 + f2 ** 2 - tau ** 2 + T0 ** 2 / 3
        return (A.spmv(f1, 1 / dx ** 2) - lam / tau ** 2 * f1 * s,
                A.spmv(f2, 1 / dx ** 2) - lam / tau ** 2 * f2 * s)

    def force_h(f1, f2, tau):
        s = f1 ** 2 + f2 ** 2 - 1 + T0 ** 2 / (3 * tau ** 2)
        return (A.spmv(f1, tau ** 2 / dx ** 2) - lam * tau ** 2 * f1 * s,
                A.spmv(f2, tau ** 2 / dx ** 2) - lam * tau ** 2 * f2 * s)

    def run(dt, eom):
        n = int(round(1.0 / dt))
        p = O.params(2, N, time_step=dt, T_initial=T0)
        if eom == 0:     # phi_fr = phi_H (tau0 = 1), phi_fr' = phi_H + tau phi_H' = phi_H
            st = O.State(p1, p2, p1, p2, *force_fr(p1, p2, 1.0))
        else:            # pi = 0
            st = O.State(p1, p2, np.zeros(L), np.zeros(L), *force_h(p1, p2, 1.0))
        st.run_form(p, A, n, eom=eom)
        return st

    ref = run(0.003125, 0)
    target = ref.phi1 / ref.tau
    errs = []
    for dt in (0.1, 0.05, 0.025):
        h = run(dt, 1)
        assert h.tau == pytest.approx(ref.tau, abs=1e-6)
        errs.append(np.linalg.norm(h.phi1 - target) / np.linalg.norm(target))
    assert errs[0] < 0.05
    for a, b in zip(errs, errs[1:]):
        assert 1.8 < a / b < 2.2, errs
