"""Pins the oracle's velocity-Verlet step (vvsl_field_rescaled,
s_integrate.cpp:305-348) with closed-form solutions."""
import numpy as np
import pytest

import oracle as O
from fields import kinv_field, random_state


def test_first_step_has_no_force_term():
    # Q1: ker_curr starts at zero (s_run.cpp:44-47), so the first drift is
    # phi_1 = phi_0 + dt * phidot_0 exactly
    N = 16
    p1, p2, v1, v2, _, _ = random_state(2, N)
    p = O.params(2, N)
    st = O.State(p1, p2, v1, v2)
    st.run(p, O.Matrix(2, N), 1)
    dt = float(np.float32(0.1))
    np.testing.assert_array_equal(st.phi1, p1 + dt * (v1 + 0.5 * 0.0 * dt))
    np.testing.assert_array_equal(st.phi2, p2 + dt * (v2 + 0.5 * 0.0 * dt))
    assert st.tau == 1.0 + dt


@pytest.mark.parametrize("ndims,N", [(2, 16), (3, 8)])
def test_vacuum_is_exact_solution(ndims, N):
    # with T0 = 0, phi = tau e^{i theta}, phidot = e^{i theta} solves
    # phi'' = Lap phi - lambda tau^-2 phi (|phi|^2 - tau^2) exactly: the force
    # vanishes and velocity-Verlet is exact for motion linear in tau
    theta = 0.7
    L = N ** ndims
    p = O.params(ndims, N, lam=1.0, T_initial=0.0)
    st = O.State(np.full(L, np.cos(theta)), np.full(L, np.sin(theta)),
                 np.full(L, np.cos(theta)), np.full(L, np.sin(theta)))
    st.run(p, O.Matrix(ndims, N), 200)
    tau = st.tau
    assert abs(tau - (1.0 + 200 * float(np.float32(0.1)))) < 1e-11
    assert np.abs(st.phi1 - tau * np.cos(theta)).max() < 1e-11 * tau
    assert np.abs(st.phi2 - tau * np.sin(theta)).max() < 1e-11 * tau
    assert np.abs(st.phidot1 - np.cos(theta)).max() < 1e-11
    assert np.abs(st.phidot2 - np.sin(theta)).max() < 1e-11


def leapfrog_closed_form(w2, dt, n):
    """phi_n for phi'' = -w2 phi, phi_0 = 1, phidot_0 = 0, with the
    reference's zero initial kernel: phi_1 = 1; from then on Stormer-Verlet,
    whose positions obey x_{m+1} - 2c x_m + x_{m-1} = 0, c = 1 - w2 dt^2/2,
    with x_0 = phi_1 = 1 and x_1 = phi_2 = 1 - w2 dt^2."""
    if n == 0:
        return 1.0
    m = n - 1
    c = 1 - w2 * dt * dt / 2
    Om = np.arccos(c)
    B = (1 - w2 * dt * dt - c) / np.sin(Om)
    return np.cos(m * Om) + B * np.sin(m * Om)


@pytest.mark.parametrize("ndims,N,mode,nsteps", [(2, 32, (3, 1), 300), (3, 16, (1, 2, 3), 150)])
def test_free_plane_wave_dispersion(ndims, N, mode, nsteps):
    # lambda = 0: each cosine mode is an independent oscillator with
    # w^2 = (4/dx^2) sum sin^2(pi n_d / N); the field after n steps follows
    # the leapfrog closed form above
    dx = 0.5
    dt = float(np.float32(0.1))
    grids = np.meshgrid(*[np.arange(N)] * ndims, indexing="ij")
    if ndims == 2:
        theta = 2 * np.pi * (mode[0] * grids[1] + mode[1] * grids[0]) / N
    else:
        theta = 2 * np.pi * sum(m * g for m, g in zip(mode, grids)) / N
    shape = np.cos(theta).reshape(-1)
    w2 = (4 / dx ** 2) * sum(np.sin(np.pi * m / N) ** 2 for m in mode)
    L = N ** ndims
    p = O.params(ndims, N, lam=0.0, T_initial=0.0)
    st = O.State(shape, 0.5 * shape, np.zeros(L), np.zeros(L))
    st.run(p, O.Matrix(ndims, N), nsteps)
    amp = leapfrog_closed_form(w2, dt, nsteps)
    assert np.abs(st.phi1 - amp * shape).max() < 1e-10
    assert np.abs(st.phi2 - 0.5 * amp * shape).max() < 1e-10


def _force(A, phi1, phi2, tau, lam, T0, dx):
    # ker = Lap phi - lambda tau^-2 phi (|phi|^2 - tau^2 + T0^2/3)
    s = phi1 ** 2 + phi2 ** 2 - tau ** 2 + T0 ** 2 / 3
    k1 = A.spmv(phi1, 1 / dx ** 2) - lam / tau ** 2 * phi1 * s
    k2 = A.spmv(phi2, 1 / dx ** 2) - lam / tau ** 2 * phi2 * s
    return k1, k2


@pytest.mark.parametrize("consistent_ker,lo,hi", [(True, 3.3, 4.7), (False, 1.7, 2.4)])
def test_order_in_dt(consistent_ker, lo, hi):
    # Velocity-Verlet positions converge as dt^2 when the initial kernel is
    # the force at t=0.  With the reference's zero initial kernel (Q1) the
    # first kick misses dt/2 * F_0, a one-time O(dt) velocity error, and the
    # run is first order: halving dt halves the error.
    N, T, dx = 16, 2.0, 0.5
    p1, p2 = kinv_field(2, N, seed=3)
    L = N * N
    A = O.Matrix(2, N)
    k1, k2 = _force(A, p1, p2, 1.0, 1.0, 1.0, dx) if consistent_ker else (None, None)

    def run(dt):
        n = int(round(T / dt))
        p = O.params(2, N, time_step=dt, T_initial=1.0)
        st = O.State(p1, p2, np.zeros(L), np.zeros(L), k1, k2)
        st.run(p, A, n)
        return st.phi1

    ref = run(0.00625)
    e1 = np.linalg.norm(run(0.1) - ref)
    e2 = np.linalg.norm(run(0.05) - ref)
    ratio = e1 / e2
    assert lo < ratio < hi, ratio


def test_tau_accumulates_float_dt():
    # s_integrate.cpp:317: tau (double) += dt (float 0.1f promoted)
    N = 8
    L = N * N
    p = O.params(2, N)
    st = O.State(np.zeros(L), np.zeros(L), np.zeros(L), np.zeros(L))
    st.run(p, O.Matrix(2, N), 10)
    t = 1.0
    for _ in range(10):
        t = t + float(np.float32(0.1))
    assert st.tau == t


@pytest.mark.parametrize("ndims,N,stencil", [(2, 64, 0), (2, 48, 3), (3, 24, 0), (3, 20, 1),
                                             (3, 16, 3)])
def test_fused_cpu_stepper_equals_reference_structure(ndims, N, stencil):
    # the CPU baseline's fused stencil stepper (two sweeps, no matrix) and
    # the reference-structured stepper (five sweeps, CSR SpMV) agree bit for
    # bit, so the two CPU timings measure the same arithmetic
    s = random_state(ndims, N, seed=3, ker=True)
    p = O.params(ndims, N, stencil_setting=stencil)
    a = O.State(*s).run(p, O.Matrix(ndims, N, stencil), 7)
    b = O.State(*s).run_fused(p, 7)
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert a.tau == b.tau
