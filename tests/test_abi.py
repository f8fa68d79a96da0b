"""The C-ABI library loads and exports every entry point include/mpsheff.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mpsheff.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from finitemptools_amd import _build
    if not os.path.exists(_build.lib_path()):
        _build.build()
    return ctypes.CDLL(_build.lib_path())


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ["mh_ac_prime", "mh_ac2_prime", "mh_c_prime", "mh_updateleft",
              "mh_updateright", "mh_lanczos_eig", "mh_lanczos_expm", "mh_heff_apply"]:
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from finitemptools_amd import _lib
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_null_handle_errors(lib):
    from finitemptools_amd import _lib
    L = _lib.load()
    assert L.mh_version() >= 1
    assert L.mh_destroy(None) == 1  # MH_ERR_ARG, no abort
    assert L.mh_last_error(None) == b"null handle"
    n = ctypes.c_size_t()
    op = _lib.HeffOp()
    op.kind, op.Dl, op.d1, op.d2, op.Dr = 2, 3, 2, 2, 5
    assert L.mh_heff_size(ctypes.byref(op), ctypes.byref(n)) == 0
    assert n.value == 3 * 2 * 2 * 5
