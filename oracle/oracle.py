"""ctypes binding of the CPU oracle (liboracle.so built from axs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, as the checker.  Nothing in the product path
imports it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_DP = POINTER(c_double)
_IP = POINTER(c_int32)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("ndims", c_int), ("N", c_int), ("space_step", c_float), ("time_step", c_float),
        ("lambda_", c_float), ("T_initial", c_float), ("stencil_setting", c_int),
        ("thr", c_int),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_length": ([c_int, c_int], c_int64),
        "orc_matrix_create": ([c_int, c_int, c_int], c_void_p),
        "orc_matrix_destroy": ([c_void_p], None),
        "orc_matrix_nnz": ([c_void_p], c_int64),
        "orc_matrix_dense": ([c_void_p, _DP, c_int64], c_int),
        "orc_spmv": ([c_void_p, c_double, _DP, _DP], None),
        "orc_laplacian2D": ([_DP, c_int, c_int, c_float, c_int], c_double),
        "orc_laplacian3D": ([_DP, c_int, c_int, c_int, c_float, c_int], c_double),
        "orc_vvsl_step": ([POINTER(OrcParams), c_void_p] + [_DP] * 9, None),
        "orc_vvsl_run": ([POINTER(OrcParams), c_void_p, c_int64] + [_DP] * 9, None),
        "orc_vvsl_run_fused": ([POINTER(OrcParams), c_int64] + [_DP] * 7, None),
        "orc_axion": ([c_int64, _DP, _DP, _DP], None),
        "orc_atan2_cr": ([c_double, c_double], c_double),
        "orc_atan2_cr_array": ([c_int64, _DP, _DP, _DP], None),
        "orc_cores2": ([_DP, c_int, c_int, _IP, c_int64], c_int64),
        "orc_cores3": ([_DP, c_int, c_int, _IP, c_int64], c_int64),
        "orc_num_cores2": ([_IP, c_int64], c_int64),
        "orc_xi2": ([c_int64, c_double, c_int, c_float], c_float),
        "orc_xi3": ([c_int64, c_double, c_int, c_float], c_float),
        "orc_background": ([c_int64] + [_DP] * 5, None),
        "orc_energy": ([POINTER(OrcParams), c_double] + [_DP] * 5, None),
        "orc_fft": ([c_int, c_int, _DP, c_int], c_int),
        "orc_laplacian_fft": ([c_int, c_int, c_float, _DP, _DP], c_int),
        "orc_vvsl_run_form": ([POINTER(OrcParams), c_void_p, c_int, c_int, c_int64] + [_DP] * 9,
                              c_int),
        "orc_powerspec": ([c_int, c_int, c_float, _DP, _DP, _DP, _DP, _IP], None),
        "orc_powerspec_log": ([c_int, c_int, c_float, _DP, _DP, c_int, _DP, _DP, _IP], None),
    }
    for name, (argt, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = argt
        fn.restype = res
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(_DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def params(ndims, N, space_step=0.5, time_step=0.1, lam=1.0, T_initial=4.0,
           stencil_setting=0, thr=1) -> OrcParams:
    return OrcParams(ndims, N, space_step, time_step, lam, T_initial, stencil_setting, thr)


class Matrix:
    """build_coefficient_matrix (s_integrate.cpp:5-149) as CSR."""

    def __init__(self, ndims: int, N: int, stencil: int = 0):
        self.ndims, self.N = ndims, N
        self.L = N ** ndims
        self.h = lib().orc_matrix_create(ndims, N, stencil)
        if not self.h:
            raise ValueError("orc_matrix_create failed")

    def close(self):
        if self.h:
            lib().orc_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def nnz(self) -> int:
        return lib().orc_matrix_nnz(self.h)

    def dense(self) -> np.ndarray:
        out = np.zeros(self.L * self.L)
        if lib().orc_matrix_dense(self.h, _dp(out), out.size) != 0:
            raise ValueError("matrix too large for dense export")
        return out.reshape(self.L, self.L)

    def spmv(self, x, alpha: float = 1.0) -> np.ndarray:
        x = _f64(x)
        y = np.empty_like(x)
        lib().orc_spmv(self.h, alpha, _dp(x), _dp(y))
        return y


class State:
    """The reference's all_data arrays (s_internal.hpp:6-21) and tau."""

    def __init__(self, phi1, phi2, phidot1, phidot2, ker1=None, ker2=None, tau=1.0):
        self.phi1 = _f64(phi1).copy()
        self.phi2 = _f64(phi2).copy()
        self.phidot1 = _f64(phidot1).copy()
        self.phidot2 = _f64(phidot2).copy()
        L = self.phi1.size
        self.ker1 = np.zeros(L) if ker1 is None else _f64(ker1).copy()
        self.ker2 = np.zeros(L) if ker2 is None else _f64(ker2).copy()
        self._k1n = np.zeros(L)
        self._k2n = np.zeros(L)
        self.tau = float(tau)

    def run(self, p: OrcParams, A: Matrix, nsteps: int) -> "State":
        t = c_double(self.tau)
        lib().orc_vvsl_run(ctypes.byref(p), A.h, int(nsteps), _dp(self.phi1), _dp(self.phi2),
                           _dp(self.phidot1), _dp(self.phidot2), _dp(self.ker1),
                           _dp(self.ker2), _dp(self._k1n), _dp(self._k2n), ctypes.byref(t))
        self.tau = t.value
        return self

    def run_fused(self, p: OrcParams, nsteps: int) -> "State":
        """the same steps with the fused CPU stencil stepper (no matrix)"""
        t = c_double(self.tau)
        lib().orc_vvsl_run_fused(ctypes.byref(p), int(nsteps), _dp(self.phi1), _dp(self.phi2),
                                 _dp(self.phidot1), _dp(self.phidot2), _dp(self.ker1),
                                 _dp(self.ker2), ctypes.byref(t))
        self.tau = t.value
        return self

    def run_form(self, p: OrcParams, A, nsteps: int, eom: int = 0, laplacian: int = 0) -> "State":
        """eom 0: vvsl_field_rescaled, 1: vvsl_hamiltonian_form (phidot
        holds pi); laplacian 1: compute_laplacian_fft in the field-rescaled
        step (A unused, may be None)."""
        t = c_double(self.tau)
        rc = lib().orc_vvsl_run_form(ctypes.byref(p), A.h if A is not None else None, eom,
                                     laplacian, int(nsteps), _dp(self.phi1), _dp(self.phi2),
                                     _dp(self.phidot1), _dp(self.phidot2), _dp(self.ker1),
                                     _dp(self.ker2), _dp(self._k1n), _dp(self._k2n),
                                     ctypes.byref(t))
        if rc != 0:
            raise ValueError(f"orc_vvsl_run_form failed ({rc})")
        self.tau = t.value
        return self


def laplacian_fft(field, ndims: int, N: int, dx: float = 0.5) -> np.ndarray:
    """compute_laplacian_fft (s_integrate.cpp:179-296)."""
    f = _f64(field)
    out = np.empty_like(f)
    if lib().orc_laplacian_fft(ndims, N, dx, _dp(f), _dp(out)) != 0:
        raise ValueError("orc_laplacian_fft failed")
    return out


def atan2_cr(y, x) -> np.ndarray:
    """atan2 rounded correctly to double, elementwise"""
    a, b = np.broadcast_arrays(np.asarray(y, np.float64), np.asarray(x, np.float64))
    a, b = _f64(a), _f64(b)
    out = np.empty_like(a)
    lib().orc_atan2_cr_array(a.size, _dp(a), _dp(b), _dp(out))
    return out


def axion(phi1, phi2) -> np.ndarray:
    a, b = _f64(phi1), _f64(phi2)
    out = np.empty_like(a)
    lib().orc_axion(a.size, _dp(a), _dp(b), _dp(out))
    return out


def cores(axion_field, ndims: int, N: int, thr: int = 1):
    """(count, coords[count, ndims]) in ascending site order."""
    ax = _f64(axion_field)
    fn = lib().orc_cores2 if ndims == 2 else lib().orc_cores3
    n = fn(_dp(ax), N, thr, None, 0)
    buf = np.zeros((max(n, 1), ndims), np.int32)
    n2 = fn(_dp(ax), N, thr, buf.ctypes.data_as(_IP), n)
    assert n2 == n
    return n, buf[:n]


def num_cores2(coords) -> int:
    c = np.ascontiguousarray(coords, dtype=np.int32)
    return lib().orc_num_cores2(c.ctypes.data_as(_IP), c.shape[0])


def xi2(num_cores, tau, N, dx) -> float:
    return lib().orc_xi2(num_cores, tau, N, dx)


def xi3(n_plaquettes, tau, N, dx) -> float:
    return lib().orc_xi3(n_plaquettes, tau, N, dx)


def background(phi1, phi2, phidot1, phidot2) -> np.ndarray:
    a = [_f64(x) for x in (phi1, phi2, phidot1, phidot2)]
    out = np.zeros(6)
    lib().orc_background(a[0].size, *[_dp(x) for x in a], _dp(out))
    return out


def energy(p: OrcParams, tau, phi1, phi2, phidot1, phidot2) -> np.ndarray:
    a = [_f64(x) for x in (phi1, phi2, phidot1, phidot2)]
    out = np.zeros(3)
    lib().orc_energy(ctypes.byref(p), tau, *[_dp(x) for x in a], _dp(out))
    return out


def fft(data: np.ndarray, ndims: int, N: int, sign: int = -1) -> np.ndarray:
    d = np.ascontiguousarray(data, dtype=np.complex128).reshape(-1).copy()
    v = d.view(np.float64)
    if lib().orc_fft(ndims, N, v.ctypes.data_as(_DP), sign) != 0:
        raise ValueError("orc_fft failed")
    return d


def powerspec(field_real, ndims: int, N: int, dx: float, field_imag=None):
    re = _f64(field_real)
    im = None if field_imag is None else _f64(field_imag)
    nb = N // 2
    ks, pk, cnt = np.zeros(nb), np.zeros(nb), np.zeros(nb, np.int32)
    lib().orc_powerspec(ndims, N, dx, _dp(re), _dp(im), _dp(ks), _dp(pk),
                        cnt.ctypes.data_as(_IP))
    return ks, pk, cnt


def powerspec_log(field_real, ndims: int, N: int, dx: float, nbins: int, field_imag=None):
    re = _f64(field_real)
    im = None if field_imag is None else _f64(field_imag)
    ks, pk, cnt = np.zeros(nbins), np.zeros(nbins), np.zeros(nbins, np.int32)
    lib().orc_powerspec_log(ndims, N, dx, _dp(re), _dp(im), nbins, _dp(ks), _dp(pk),
                            cnt.ctypes.data_as(_IP))
    return ks, pk, cnt
