"""The C-ABI library loads and exports every symbol include/smoq.h declares;
host-side argument checks and error plumbing work without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smoq.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(smoq_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("smoq_init", "smoq_state_alloc", "smoq_apply_named", "smoq_apply_tape",
                 "smoq_expect_pauli", "smoq_sample", "smoq_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from smoq_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsmoq_amd.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares a signature for every exported function
    assert set(declared_functions()) <= set(_lib.SIGNATURES), \
        set(declared_functions()) - set(_lib.SIGNATURES)


def test_version_and_error_paths_without_gpu():
    from smoq_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsmoq_amd.so not built")
    L = _lib.lib()
    assert L.smoq_version() == 1
    # null pointers are rejected before any device call
    assert L.smoq_init(None, 0) == _lib.SMOQ_EINVAL
    assert b"null" in L.smoq_last_error()
    assert L.smoq_state_set_basis(None, 0) == _lib.SMOQ_EINVAL
    assert L.smoq_sync(None) == _lib.SMOQ_EINVAL
    assert L.smoq_mcwf_evolve(None, None, 1, 0.1, 1, 0) == _lib.SMOQ_EINVAL
    assert L.smoq_state_set_product(None, None) == _lib.SMOQ_EINVAL
    assert L.smoq_lanczos_range(None, 0, None, None, None, 30, None, None, None) == _lib.SMOQ_EINVAL


def test_init_reports_missing_device_cleanly():
    import torch

    from smoq_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsmoq_amd.so not built")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    rc = _lib.lib().smoq_init(ctypes.byref(h), 0)
    assert rc in (_lib.SMOQ_EDEVICE, _lib.SMOQ_ENOTSUP)
    assert _lib.lib().smoq_last_error()
