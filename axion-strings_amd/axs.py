"""ctypes binding of libaxs.so, the C-ABI declared in include/axs_gpu.h.

Used by the test suite and by bench.py.  The library is the HIP
implementation; if it is missing this module raises at load time -- there is
no CPU fallback.

Arrays cross the boundary as C-contiguous float64 numpy arrays in the
reference layout (2D m = i + N*j, 3D m = (i*N + j)*N + k); a 2D lattice may
be given as shape (N, N) indexed [j, i] and a 3D one as (N, N, N) indexed
[i, j, k], which are the same memory order.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaxs.so")

AXS_OK = 0
AXS_ERR_INVALID_ARG = 1
AXS_ERR_UNSUPPORTED = 2
AXS_ERR_OUT_OF_MEMORY = 3
AXS_ERR_HIP = 4
AXS_ERR_CAPACITY = 5
AXS_ERR_NO_DEVICE = 6
AXS_ERR_FFT = 7
AXS_ERR_STATE = 8

AXS_FP64 = 0
AXS_FP32 = 1
AXS_BIN_LINEAR = 0
AXS_BIN_LOG = 1
AXS_INIT_KINV = 0
AXS_INIT_THERMAL = 1
AXS_SNAP_PHI = 1
AXS_SNAP_VEL = 2
AXS_SNAP_KER = 4
AXS_FIELD_PHI2 = 0
AXS_FIELD_PHI1 = 1
AXS_FIELD_PHI = 2
AXS_FIELD_PHIDOT2 = 3
AXS_FIELD_PHIDOT1 = 4
AXS_FIELD_PHIDOT = 5
FIELDS = {"phi2": AXS_FIELD_PHI2, "phi1": AXS_FIELD_PHI1, "phi": AXS_FIELD_PHI,
          "phidot2": AXS_FIELD_PHIDOT2, "phidot1": AXS_FIELD_PHIDOT1, "phidot": AXS_FIELD_PHIDOT}
AXS_EOM_FIELD_RESCALED = 0
AXS_EOM_HAMILTONIAN = 1
AXS_LAPLACIAN_STENCIL = 0
AXS_LAPLACIAN_SPECTRAL = 1
EOMS = {"field_rescaled": AXS_EOM_FIELD_RESCALED, "hamiltonian": AXS_EOM_HAMILTONIAN}
LAPLACIANS = {"stencil": AXS_LAPLACIAN_STENCIL, "spectral": AXS_LAPLACIAN_SPECTRAL}

# every entry point declared in include/axs_gpu.h
EXPORTS = (
    "axs_create", "axs_create_batch", "axs_destroy", "axs_num_members",
    "axs_set_member_state", "axs_get_member_state", "axs_init_member", "axs_set_tau",
    "axs_get_tau",
    "axs_set_state", "axs_get_state", "axs_step", "axs_member_axion", "axs_member_strings",
    "axs_strings",
    "axs_string_counts", "axs_member_background", "axs_background", "axs_member_nan_check",
    "axs_member_energy", "axs_energy", "axs_member_energy_noalloc",
    "axs_member_powerspec", "axs_member_powerspec_field", "axs_powerspec",
    "axs_powerspec_all",
    "axs_last_step_ms", "axs_bytes_per_site", "axs_strerror", "axs_last_error",
    "axs_set_kernel_config", "axs_get_kernel_config", "axs_autotune", "axs_member_laplacian",
    "axs_snapshot_begin", "axs_snapshot_end",
)

AXS_KERNEL_AUTO = 0
AXS_KERNEL_SIMPLE = 1
AXS_KERNEL_MARCH = 2
KERNELS = {"auto": AXS_KERNEL_AUTO, "simple": AXS_KERNEL_SIMPLE, "march": AXS_KERNEL_MARCH}


class AxsKernelConfig(ctypes.Structure):
    """struct axs_kernel_config (include/axs_gpu.h)."""

    _fields_ = [("kernel", c_int), ("rows", c_int), ("chunk", c_int), ("depth", c_int),
                ("flags", c_int)]


AXS_MARCH_PREFETCH = 1
AXS_MARCH_NT_STORE = 2
AXS_MARCH_JFAST = 4
AXS_STEP_GRAPH = 8
AXS_MARCH_WG512 = 16
AXS_MARCH_WG1024 = 32


class AxsParams(ctypes.Structure):
    """struct axs_params (include/axs_gpu.h)."""

    _fields_ = [
        ("ndims", c_int),
        ("N", c_int),
        ("space_step", c_float),
        ("time_step", c_float),
        ("lambda_", c_float),
        ("T_initial", c_float),
        ("stencil_setting", c_int),
        ("thr", c_int),
        ("precision", c_int),
        ("device", c_int),
        ("eom", c_int),
        ("laplacian", c_int),
    ]


class AxsError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: status {code}: {msg}")
        self.code = code


_lib = None
_DP = POINTER(c_double)


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libaxs.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `make -C axion-strings_amd` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    ctx_pp = POINTER(c_void_p)
    sig = {
        "axs_create": ([POINTER(AxsParams), ctx_pp], c_int),
        "axs_create_batch": ([POINTER(AxsParams), c_int, POINTER(c_int), ctx_pp], c_int),
        "axs_destroy": ([c_void_p], c_int),
        "axs_num_members": ([c_void_p, POINTER(c_int)], c_int),
        "axs_set_member_state": ([c_void_p, c_int] + [_DP] * 6, c_int),
        "axs_get_member_state": ([c_void_p, c_int] + [_DP] * 6, c_int),
        "axs_init_member": ([c_void_p, c_int, c_int, ctypes.c_uint64], c_int),
        "axs_set_tau": ([c_void_p, c_double], c_int),
        "axs_get_tau": ([c_void_p, _DP], c_int),
        "axs_set_state": ([c_void_p] + [_DP] * 6 + [c_double], c_int),
        "axs_get_state": ([c_void_p] + [_DP] * 7, c_int),
        "axs_step": ([c_void_p, c_int64], c_int),
        "axs_member_axion": ([c_void_p, c_int, _DP], c_int),
        "axs_member_strings": ([c_void_p, c_int, POINTER(c_int32), c_int64, POINTER(c_int64)], c_int),
        "axs_strings": ([c_void_p, POINTER(c_int32), c_int64, POINTER(c_int64)], c_int),
        "axs_string_counts": ([c_void_p, POINTER(c_int64), POINTER(c_int64)], c_int),
        "axs_member_background": ([c_void_p, c_int, _DP], c_int),
        "axs_background": ([c_void_p, _DP], c_int),
        "axs_member_nan_check": ([c_void_p, c_int, POINTER(c_int64)], c_int),
        "axs_member_energy": ([c_void_p, c_int, _DP], c_int),
        "axs_energy": ([c_void_p, _DP], c_int),
        "axs_member_energy_noalloc": ([c_void_p, c_int, _DP], c_int),
        "axs_member_powerspec": ([c_void_p, c_int, c_int, c_int, _DP, _DP, POINTER(c_int32)], c_int),
        "axs_member_powerspec_field": ([c_void_p, c_int, c_int, c_int, c_int, _DP, _DP,
                                        POINTER(c_int32)], c_int),
        "axs_powerspec_all": ([c_void_p, c_int, c_int, _DP, _DP, POINTER(c_int32)], c_int),
        "axs_powerspec": ([c_void_p, c_int, c_int, _DP, _DP, POINTER(c_int32)], c_int),
        "axs_last_step_ms": ([c_void_p, _DP], c_int),
        "axs_bytes_per_site": ([c_void_p, _DP, _DP], c_int),
        "axs_member_laplacian": ([c_void_p, c_int, c_int, _DP, _DP], c_int),
        "axs_snapshot_begin": ([c_void_p, c_int, c_int, c_int64], c_int),
        "axs_snapshot_end": ([c_void_p] + [_DP] * 7, c_int),
        "axs_set_kernel_config": ([c_void_p, POINTER(AxsKernelConfig)], c_int),
        "axs_get_kernel_config": ([c_void_p, POINTER(AxsKernelConfig)], c_int),
        "axs_autotune": ([c_void_p, c_int, POINTER(AxsKernelConfig), _DP], c_int),
        "axs_strerror": ([c_int], ctypes.c_char_p),
        "axs_last_error": ([], ctypes.c_char_p),
    }
    for name, (argt, rest) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = argt
        fn.restype = rest
    _lib = lib
    return lib


def _check(code: int, where: str) -> None:
    if code != AXS_OK:
        lib = load_library()
        msg = lib.axs_last_error().decode() or lib.axs_strerror(code).decode()
        raise AxsError(code, where, msg)


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(_DP)


def _as_f64(a, L: int, name: str) -> np.ndarray:
    arr = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if arr.size != L:
        raise ValueError(f"{name}: expected {L} values, got {arr.size}")
    return arr


PRECISIONS = {"fp64": AXS_FP64, "fp32": AXS_FP32}


class Lattice:
    """A context of one lattice (members=None) or an ensemble of lattices.

    Mirrors the reference's run-state: the four field arrays and the two
    current kernels of all_data (src/standard/s_internal.hpp:6-21) plus the
    global tau (src/common/physics.cpp:3).
    """

    def __init__(self, ndims: int, N: int = 0, *, space_step: float = 0.5,
                 time_step: float = 0.1, lam: float = 1.0, T_initial: float = 4.0,
                 stencil_setting: int = 0, thr: int = 1, precision: str = "fp64",
                 device: int = 0, members=None, eom: str = "field_rescaled",
                 laplacian: str = "stencil"):
        self.lib = load_library()
        self.params = AxsParams(ndims, N if members is None else int(members[0]),
                                space_step, time_step, lam, T_initial, stencil_setting,
                                thr, PRECISIONS[precision], device, EOMS[eom],
                                LAPLACIANS[laplacian])
        self.ndims = ndims
        self.precision = precision
        self._ctx = c_void_p()
        if members is None:
            _check(self.lib.axs_create(ctypes.byref(self.params), ctypes.byref(self._ctx)),
                   "axs_create")
            self.Ns = [N]
        else:
            Ns = (c_int * len(members))(*[int(n) for n in members])
            _check(self.lib.axs_create_batch(ctypes.byref(self.params), len(members), Ns,
                                             ctypes.byref(self._ctx)), "axs_create_batch")
            self.Ns = [int(n) for n in members]

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            self.lib.axs_destroy(self._ctx)
            self._ctx = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def length(self, member: int = 0) -> int:
        return self.Ns[member] ** self.ndims

    # -- state -------------------------------------------------------------
    def set_state(self, phi1, phi2, phidot1, phidot2, ker1=None, ker2=None, *,
                  member: int = 0, tau: float | None = None) -> None:
        L = self.length(member)
        a = [_as_f64(x, L, n) for x, n in ((phi1, "phi1"), (phi2, "phi2"),
                                           (phidot1, "phidot1"), (phidot2, "phidot2"))]
        k1 = None if ker1 is None else _as_f64(ker1, L, "ker1")
        k2 = None if ker2 is None else _as_f64(ker2, L, "ker2")
        _check(self.lib.axs_set_member_state(self._ctx, member, _dp(a[0]), _dp(a[1]),
                                             _dp(a[2]), _dp(a[3]), _dp(k1), _dp(k2)),
               "axs_set_member_state")
        if tau is not None:
            self.tau = tau

    STATE_FIELDS = ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2")

    def get_state(self, member: int = 0, fields=STATE_FIELDS) -> dict:
        L = self.length(member)
        out = {k: np.empty(L, np.float64) for k in fields}
        _check(self.lib.axs_get_member_state(self._ctx, member,
                                             *[_dp(out.get(k)) for k in self.STATE_FIELDS]),
               "axs_get_member_state")
        return out

    def init(self, mode: str = "kinv", seed: int = 1998, member: int = 0) -> None:
        """device-side initial conditions (include/axs_gpu.h axs_init_member)"""
        m = AXS_INIT_KINV if mode == "kinv" else AXS_INIT_THERMAL
        _check(self.lib.axs_init_member(self._ctx, member, m, int(seed)), "axs_init_member")

    @property
    def tau(self) -> float:
        t = c_double()
        _check(self.lib.axs_get_tau(self._ctx, ctypes.byref(t)), "axs_get_tau")
        return t.value

    @tau.setter
    def tau(self, value: float) -> None:
        _check(self.lib.axs_set_tau(self._ctx, float(value)), "axs_set_tau")

    # -- step --------------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        _check(self.lib.axs_step(self._ctx, int(nsteps)), "axs_step")

    def set_kernel(self, kernel: str = "auto", rows: int = 0, chunk: int = 0, depth: int = 1,
                   flags: int = AXS_MARCH_PREFETCH | AXS_MARCH_NT_STORE):
        cfg = AxsKernelConfig(KERNELS[kernel], rows, chunk, depth, flags)
        _check(self.lib.axs_set_kernel_config(self._ctx, ctypes.byref(cfg)),
               "axs_set_kernel_config")

    def kernel_config(self) -> dict:
        cfg = AxsKernelConfig()
        _check(self.lib.axs_get_kernel_config(self._ctx, ctypes.byref(cfg)),
               "axs_get_kernel_config")
        return {"kernel": cfg.kernel, "rows": cfg.rows, "chunk": cfg.chunk, "depth": cfg.depth,
                "flags": cfg.flags}

    def autotune(self, trials: int = 0):
        """choose the fastest launch shape; returns (config dict, ms per step)"""
        cfg = AxsKernelConfig()
        ms = c_double()
        _check(self.lib.axs_autotune(self._ctx, trials, ctypes.byref(cfg), ctypes.byref(ms)),
               "axs_autotune")
        return ({"kernel": cfg.kernel, "rows": cfg.rows, "chunk": cfg.chunk, "depth": cfg.depth,
                 "flags": cfg.flags}, ms.value)

    @property
    def last_step_ms(self) -> float:
        v = c_double()
        _check(self.lib.axs_last_step_ms(self._ctx, ctypes.byref(v)), "axs_last_step_ms")
        return v.value

    def bytes_per_site(self):
        s, a = c_double(), c_double()
        _check(self.lib.axs_bytes_per_site(self._ctx, ctypes.byref(s), ctypes.byref(a)),
               "axs_bytes_per_site")
        return s.value, a.value

    # -- strings -----------------------------------------------------------
    def axion(self, member: int = 0) -> np.ndarray:
        """atan2(phi1, phi2) of every site (the reference's axion array),
        rounded correctly"""
        out = np.empty(self.length(member), np.float64)
        _check(self.lib.axs_member_axion(self._ctx, member, _dp(out)), "axs_member_axion")
        return out

    def strings(self, member: int = 0, coords: bool = True):
        """(count, coords[count, ndims] int32 or None), ascending site order."""
        n = c_int64()
        if not coords:
            _check(self.lib.axs_member_strings(self._ctx, member, None, 0, ctypes.byref(n)),
                   "axs_member_strings")
            return n.value, None
        _check(self.lib.axs_member_strings(self._ctx, member, None, 0, ctypes.byref(n)),
               "axs_member_strings")
        buf = np.empty((max(n.value, 1), self.ndims), np.int32)
        _check(self.lib.axs_member_strings(self._ctx, member,
                                           buf.ctypes.data_as(POINTER(c_int32)),
                                           n.value, ctypes.byref(n)), "axs_member_strings")
        return n.value, buf[: n.value]

    def string_counts(self):
        nm = len(self.Ns)
        a = np.zeros(nm, np.int64)
        b = np.zeros(nm, np.int64)
        _check(self.lib.axs_string_counts(self._ctx, a.ctypes.data_as(POINTER(c_int64)),
                                          b.ctypes.data_as(POINTER(c_int64))),
               "axs_string_counts")
        return a, b

    # -- reductions --------------------------------------------------------
    def background(self, member: int = 0) -> np.ndarray:
        out = np.zeros(6, np.float64)
        _check(self.lib.axs_member_background(self._ctx, member, _dp(out)),
               "axs_member_background")
        return out

    def nan_check(self, member: int = 0) -> int:
        """index of the first site whose phi1 or phi2 is NaN, -1 if none
        (the reference's debug scan, s_run.cpp:4-12)"""
        first = c_int64()
        _check(self.lib.axs_member_nan_check(self._ctx, member, ctypes.byref(first)),
               "axs_member_nan_check")
        return first.value

    def snapshot_begin(self, member: int = 0, fields=("phi1", "phi2", "phidot1", "phidot2",
                                                      "ker1", "ker2"), count: int = 0):
        """start an asynchronous copy of the current state (see
        axs_snapshot_begin); finish it with snapshot_end()"""
        bits = ((AXS_SNAP_PHI if {"phi1", "phi2"} & set(fields) else 0) |
                (AXS_SNAP_VEL if {"phidot1", "phidot2"} & set(fields) else 0) |
                (AXS_SNAP_KER if {"ker1", "ker2"} & set(fields) else 0))
        _check(self.lib.axs_snapshot_begin(self._ctx, member, bits, count), "axs_snapshot_begin")
        self._snap = (member, tuple(fields), count or self.Ns[member] ** self.ndims)

    def snapshot_end(self):
        """(dict of the requested arrays, tau) of the snapshot in flight"""
        # without a snapshot begun, the library reports the error
        member, fields, n = getattr(self, "_snap", None) or (0, (), 0)
        names = ("phi1", "phi2", "phidot1", "phidot2", "ker1", "ker2")
        arrs = {k: np.empty(n) for k in fields}
        tau = c_double()
        _check(self.lib.axs_snapshot_end(self._ctx, *[_dp(arrs.get(k)) for k in names],
                                         ctypes.byref(tau)), "axs_snapshot_end")
        self._snap = None
        return arrs, tau.value

    def laplacian(self, mode: str = "stencil", member: int = 0):
        """(lap1, lap2) of the current field: the stencil operator
        (1/dx^2) A phi or the spectral one (compute_laplacian_fft)."""
        L = self.Ns[member] ** self.ndims
        l1, l2 = np.empty(L), np.empty(L)
        _check(self.lib.axs_member_laplacian(self._ctx, member, LAPLACIANS[mode], _dp(l1),
                                             _dp(l2)), "axs_member_laplacian")
        return l1, l2

    def energy(self, member: int = 0) -> np.ndarray:
        out = np.zeros(3, np.float64)
        _check(self.lib.axs_member_energy(self._ctx, member, _dp(out)), "axs_member_energy")
        return out

    def energy_noalloc(self, member: int = 0) -> np.ndarray:
        """energy() through the entry point that allocates nothing (a fixed
        per-GPU workspace)"""
        out = np.zeros(3, np.float64)
        _check(self.lib.axs_member_energy_noalloc(self._ctx, member, _dp(out)),
               "axs_member_energy_noalloc")
        return out

    def powerspec(self, member: int = 0, binning: str = "linear", nbins: int = 64,
                  field: str = "phi2"):
        """(ks, pk, count) of phi2 (the reference's call) or of another field:
        phi1, phidot1, phidot2, or the complex phi = phi1 + i phi2 and
        phidot = phidot1 + i phidot2"""
        N = self.Ns[member]
        b = AXS_BIN_LINEAR if binning == "linear" else AXS_BIN_LOG
        nb = N // 2 if b == AXS_BIN_LINEAR else nbins
        ks = np.zeros(nb, np.float64)
        pk = np.zeros(nb, np.float64)
        cnt = np.zeros(nb, np.int32)
        if field == "phi2":
            _check(self.lib.axs_member_powerspec(self._ctx, member, b, nb, _dp(ks), _dp(pk),
                                                 cnt.ctypes.data_as(POINTER(c_int32))),
                   "axs_member_powerspec")
        else:
            _check(self.lib.axs_member_powerspec_field(self._ctx, member, FIELDS[field], b, nb,
                                                       _dp(ks), _dp(pk),
                                                       cnt.ctypes.data_as(POINTER(c_int32))),
                   "axs_member_powerspec_field")
        return ks, pk, cnt

    def powerspec_all(self, binning: str = "linear", nbins: int = 64):
        """[(ks, pk, count)] of every member from one call (members of one side
        share batched transforms)."""
        b = AXS_BIN_LINEAR if binning == "linear" else AXS_BIN_LOG
        nbs = [N // 2 if b == AXS_BIN_LINEAR else nbins for N in self.Ns]
        tot = sum(nbs)
        ks = np.zeros(tot, np.float64)
        pk = np.zeros(tot, np.float64)
        cnt = np.zeros(tot, np.int32)
        _check(self.lib.axs_powerspec_all(self._ctx, b, nbins, _dp(ks), _dp(pk),
                                          cnt.ctypes.data_as(POINTER(c_int32))),
               "axs_powerspec_all")
        out, o = [], 0
        for nb in nbs:
            out.append((ks[o:o + nb], pk[o:o + nb], cnt[o:o + nb]))
            o += nb
        return out


# ---------------------------------------------------------------------------
# host driver library (libaxs_host.so): parameter file, host initial
# conditions and FFT, run loop

HOST_LIB_PATH = os.path.join(_HERE, "libaxs_host.so")
DRIVER_PATH = os.path.join(_HERE, "axs_main")
_host = None


def load_host_library(path: str = HOST_LIB_PATH) -> ctypes.CDLL:
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `make -C axion-strings_amd`")
    h = ctypes.CDLL(path)
    h.axs_host_params_json.argtypes = [ctypes.c_char_p, ctypes.c_char_p, c_int]
    h.axs_host_params_json.restype = c_int
    h.axs_host_init_fields.argtypes = [c_int, c_int, c_int, c_float, c_float, c_float,
                                       ctypes.c_uint64] + [_DP] * 4
    h.axs_host_init_fields.restype = c_int
    h.axs_host_philox4x32_10.argtypes = [POINTER(ctypes.c_uint32)] * 3
    h.axs_host_philox4x32_10.restype = None
    h.axs_host_fft.argtypes = [_DP, c_int, c_int, c_int]
    h.axs_host_fft.restype = c_int
    h.axs_host_final_step.argtypes = [c_int, c_float, c_float]
    h.axs_host_final_step.restype = c_int
    h.axs_host_xi.argtypes = [c_int, c_int64, c_double, c_int, c_float]
    h.axs_host_xi.restype = c_float
    h.axs_host_run.argtypes = [ctypes.c_char_p]
    h.axs_host_run.restype = c_int
    _host = h
    return h


def read_params(path: str) -> dict:
    import json
    h = load_host_library()
    buf = ctypes.create_string_buffer(1 << 14)
    rc = h.axs_host_params_json(path.encode(), buf, len(buf))
    if rc != 0:
        raise ValueError(f"parameter file {path}: error {rc}")
    return json.loads(buf.value.decode())


def host_init(mode: str, ndims: int, N: int, space_step: float = 0.5, lam: float = 1.0,
              T_initial: float = 4.0, seed: int = 1998):
    """host initial fields: mode "kinv" or "thermal"; returns phi1, phi2, phidot1, phidot2"""
    h = load_host_library()
    L = N ** ndims
    out = [np.empty(L) for _ in range(4)]
    rc = h.axs_host_init_fields(0 if mode == "kinv" else 1, ndims, N, space_step, lam, T_initial,
                                seed, *[_dp(a) for a in out])
    if rc != 0:
        raise ValueError(f"axs_host_init_fields: {rc}")
    return out


def host_philox(ctr, key):
    """one Philox4x32-10 block of the initial-condition generators"""
    h = load_host_library()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    h.axs_host_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def host_fft(data, ndims: int, N: int, sign: int = -1) -> np.ndarray:
    h = load_host_library()
    d = np.ascontiguousarray(data, dtype=np.complex128).reshape(-1).copy()
    if h.axs_host_fft(d.view(np.float64).ctypes.data_as(_DP), ndims, N, sign) != 0:
        raise ValueError("axs_host_fft failed")
    return d
