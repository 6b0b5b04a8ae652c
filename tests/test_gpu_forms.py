"""GPU parity of the reference's alternative step forms against the oracle:
the Hamiltonian form (vvsl_hamiltonian_form, s_integrate.cpp:358-395) bit
for bit, the field-rescaled step with the spectral Laplacian
(compute_laplacian_fft, :179-296, used at :326-328) to FFT rounding, and the
two Laplacian operators of axs_member_laplacian."""
import numpy as np
import pytest

import axs
import oracle as O
from fields import kinv_field, random_state, rel_l2

pytestmark = pytest.mark.gpu

# spectral forms: rocFFT and the oracle's FFT round differently; the
# tolerance is on the field's relative L2 difference after the run
SPECTRAL_TOL = 1e-11


def _oracle(nd, N, state, nsteps, eom=0, laplacian=0, **kw):
    p = O.params(nd, N, **kw)
    st = O.State(*state)
    A = O.Matrix(nd, N, kw.get("stencil_setting", 0)) if laplacian == 0 else None
    st.run_form(p, A, nsteps, eom=eom, laplacian=laplacian)
    return st


@pytest.mark.parametrize("nd,N,nsteps,stencil,kernel", [
    (2, 64, 60, 0, "auto"), (2, 48, 30, 2, "auto"), (3, 64, 20, 0, "march"),
    (3, 64, 12, 0, "simple"), (3, 24, 10, 1, "auto")])
def test_hamiltonian_form_bit_exact(nd, N, nsteps, stencil, kernel):
    state = random_state(nd, N, seed=21, ker=True)
    st = _oracle(nd, N, state, nsteps, eom=1, stencil_setting=stencil)
    lat = axs.Lattice(nd, N, eom="hamiltonian", stencil_setting=stencil)
    lat.set_kernel(kernel)
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    g = lat.get_state()
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=k)
    assert lat.tau == st.tau


def test_hamiltonian_form_ensemble_bit_exact():
    Ns = [32, 64, 48]
    lat = axs.Lattice(2, members=Ns, eom="hamiltonian")
    states = []
    for r, N in enumerate(Ns):
        s = random_state(2, N, seed=30 + r, ker=True)
        lat.set_state(*s, member=r)
        states.append(s)
    lat.tau = 1.0
    lat.step(25)
    for r, N in enumerate(Ns):
        st = _oracle(2, N, states[r], 25, eom=1)
        g = lat.get_state(member=r)
        for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
            np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=f"member {r} {k}")


@pytest.mark.parametrize("nd,N,nsteps", [(2, 64, 30), (2, 96, 12), (3, 32, 8), (3, 48, 4)])
def test_spectral_laplacian_step(nd, N, nsteps):
    p1, p2 = kinv_field(nd, N, seed=40)
    z = np.zeros(N ** nd)
    state = (p1, p2, 0.3 * p2, -0.2 * p1, z, z)
    st = _oracle(nd, N, state, nsteps, laplacian=1)
    lat = axs.Lattice(nd, N, laplacian="spectral")
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    g = lat.get_state()
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        assert rel_l2(g[k], getattr(st, k)) < SPECTRAL_TOL, k
    assert lat.tau == st.tau


def test_spectral_step_ensemble():
    Ns = [32, 64]
    lat = axs.Lattice(2, members=Ns, laplacian="spectral")
    states = []
    for r, N in enumerate(Ns):
        p1, p2 = kinv_field(2, N, seed=50 + r)
        z = np.zeros(N * N)
        s = (p1, p2, z, z, z, z)
        lat.set_state(*s, member=r)
        states.append(s)
    lat.tau = 1.0
    lat.step(10)
    for r, N in enumerate(Ns):
        st = _oracle(2, N, states[r], 10, laplacian=1)
        g = lat.get_state(member=r)
        assert rel_l2(g["phi1"], st.phi1) < SPECTRAL_TOL
        assert rel_l2(g["phidot2"], st.phidot2) < SPECTRAL_TOL


@pytest.mark.parametrize("nd,N,precision", [(2, 64, "fp64"), (3, 32, "fp64"), (3, 32, "fp32")])
def test_member_laplacian_both_operators(nd, N, precision):
    rng = np.random.default_rng(3)
    p1, p2 = rng.standard_normal(N ** nd), rng.standard_normal(N ** nd)
    z = np.zeros(N ** nd)
    lat = axs.Lattice(nd, N, precision=precision)
    lat.set_state(p1, p2, z, z, tau=1.0)
    if precision == "fp32":         # the lattice holds the fields in float
        p1 = p1.astype(np.float32).astype(np.float64)
        p2 = p2.astype(np.float32).astype(np.float64)
    s1, s2 = lat.laplacian("stencil")
    A = O.Matrix(nd, N)
    alpha = 1.0 / (0.5 * 0.5)
    # the stencil operator in the matrix's row order: identical bits
    np.testing.assert_array_equal(s1, A.spmv(p1, alpha))
    np.testing.assert_array_equal(s2, A.spmv(p2, alpha))
    f1, f2 = lat.laplacian("spectral")
    r1, r2 = O.laplacian_fft(p1, nd, N), O.laplacian_fft(p2, nd, N)
    assert np.abs(f1 - r1).max() < 1e-12 * np.abs(r1).max()
    assert np.abs(f2 - r2).max() < 1e-12 * np.abs(r2).max()


def test_stencil_and_spectral_laplacians_agree_on_smooth_fields():
    # the validation use: on the k^-1 field (modes below pi / (4 dx)) the
    # 7-point stencil's eigenvalue (4/dx^2) sin^2(k dx / 2) differs from
    # -k^2 by at most 1 - (sin(pi/8) / (pi/8))^2 = 5%
    N = 64
    p1, p2 = kinv_field(3, N, seed=8)
    z = np.zeros(N ** 3)
    lat = axs.Lattice(3, N)
    lat.set_state(p1, p2, z, z, tau=1.0)
    s1, _ = lat.laplacian("stencil")
    f1, _ = lat.laplacian("spectral")
    d = rel_l2(s1, f1)
    assert 1e-4 < d < 0.051, d


@pytest.mark.parametrize("kw,match", [
    (dict(precision="fp32", eom="hamiltonian"), "AXS_FP64"),
    (dict(precision="fp32", laplacian="spectral"), "AXS_FP64"),
    (dict(eom="hamiltonian", laplacian="spectral"), "Hamiltonian"),
])
def test_unsupported_form_combinations(kw, match):
    with pytest.raises(axs.AxsError, match=match):
        axs.Lattice(2, 32, **kw)


def test_energy_rejected_for_hamiltonian_form():
    lat = axs.Lattice(2, 32, eom="hamiltonian")
    z = np.zeros(32 * 32)
    lat.set_state(z + 1, z, z, z, tau=1.0)
    with pytest.raises(axs.AxsError, match="field-rescaled"):
        lat.energy()


@pytest.mark.parametrize("depth,rows,flags,nsteps", [(2, 2, 19, 9), (3, 3, 18, 8)])
def test_hamiltonian_form_multi_step_bit_exact(depth, rows, flags, nsteps):
    nd, N = 3, 64
    state = random_state(nd, N, seed=22, ker=True)
    st = _oracle(nd, N, state, nsteps, eom=1)
    lat = axs.Lattice(nd, N, eom="hamiltonian")
    lat.set_kernel("march", rows=rows, depth=depth, flags=flags)
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    g = lat.get_state()
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=k)


def test_hamiltonian_form_2d_multi_step_bit_exact():
    nd, N, nsteps = 2, 128, 15
    state = random_state(nd, N, seed=23, ker=True)
    st = _oracle(nd, N, state, nsteps, eom=1)
    lat = axs.Lattice(nd, N, eom="hamiltonian")
    lat.set_kernel("march", depth=4, flags=3)
    lat.set_state(*state, tau=1.0)
    lat.step(nsteps)
    g = lat.get_state()
    for k in ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2"):
        np.testing.assert_array_equal(g[k], getattr(st, k), err_msg=k)
