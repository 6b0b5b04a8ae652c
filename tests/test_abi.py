"""The C-ABI library loads and exports every entry point include/axs_gpu.h
declares; argument validation runs before any device call (no GPU needed)."""
import ctypes
import os
import re

import pytest

import axs


def _declared(root):
    hdr = open(os.path.join(root, "include", "axs_gpu.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(axs_\w+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol(root):
    lib = axs.load_library()
    names = _declared(root)
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(axs.EXPORTS) == names


def test_strerror():
    lib = axs.load_library()
    assert lib.axs_strerror(0) == b"ok"
    assert lib.axs_strerror(axs.AXS_ERR_CAPACITY) == b"output buffer too small"
    assert lib.axs_strerror(12345) == b"unknown status"


@pytest.mark.parametrize("field,value", [("ndims", 4), ("N", 7), ("N", 2), ("stencil_setting", 5),
                                         ("precision", 9), ("time_step", 0.0), ("eom", 2),
                                         ("laplacian", 3), ("eom", -1)])
def test_invalid_params_rejected(field, value):
    lib = axs.load_library()
    p = axs.AxsParams(2, 16, 0.5, 0.1, 1.0, 4.0, 0, 1, 0, 0)
    setattr(p, field, value)
    ctx = ctypes.c_void_p()
    rc = lib.axs_create(ctypes.byref(p), ctypes.byref(ctx))
    assert rc in (axs.AXS_ERR_INVALID_ARG, axs.AXS_ERR_UNSUPPORTED)
    assert not ctx
    assert lib.axs_last_error()


def test_null_arguments():
    lib = axs.load_library()
    assert lib.axs_create(None, None) == axs.AXS_ERR_INVALID_ARG
    assert lib.axs_step(None, 1) == axs.AXS_ERR_INVALID_ARG
    assert lib.axs_destroy(None) == axs.AXS_OK
    n = ctypes.c_int()
    assert lib.axs_num_members(None, ctypes.byref(n)) == axs.AXS_ERR_INVALID_ARG


def test_batch_rejects_bad_member_size():
    lib = axs.load_library()
    p = axs.AxsParams(2, 16, 0.5, 0.1, 1.0, 4.0, 0, 1, 0, 0)
    Ns = (ctypes.c_int * 3)(16, 33, 64)
    ctx = ctypes.c_void_p()
    assert lib.axs_create_batch(ctypes.byref(p), 3, Ns, ctypes.byref(ctx)) == axs.AXS_ERR_INVALID_ARG


@pytest.mark.parametrize("precision,eom,lap", [(1, 1, 0), (1, 0, 1), (0, 1, 1)])
def test_unsupported_form_combinations_rejected(precision, eom, lap):
    # the Hamiltonian form and the spectral Laplacian are FP64 paths of the
    # reference, and it has no Hamiltonian step with the spectral Laplacian
    lib = axs.load_library()
    p = axs.AxsParams(2, 16, 0.5, 0.1, 1.0, 4.0, 0, 1, precision, 0, eom, lap)
    ctx = ctypes.c_void_p()
    assert lib.axs_create(ctypes.byref(p), ctypes.byref(ctx)) == axs.AXS_ERR_UNSUPPORTED
    assert not ctx
