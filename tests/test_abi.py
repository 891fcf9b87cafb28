"""The C ABI without a GPU: the library loads, exports every function that
include/faces_amd.h declares, and rejects invalid arguments before touching
the device."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "faces_amd.h")
LIB = os.path.join(ROOT, "faces_amd", "lib", "libfaces_amd.so")

needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(faces_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import torch  # noqa: F401  (one HIP runtime for the process)
    return ctypes.CDLL(LIB)


def test_header_declares_the_plain_step():
    names = declared()
    for must in ["faces_plan_create", "faces_plan_destroy", "faces_share", "faces_strerror",
                 "faces_batch_create", "faces_batch_share", "faces_server_create",
                 "faces_server_submit", "faces_server_wait"]:
        assert must in names


@needs_lib
def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@needs_lib
def test_strerror_and_version(lib):
    lib.faces_strerror.restype = ctypes.c_char_p
    lib.faces_version.restype = ctypes.c_char_p
    assert lib.faces_strerror(0) == b"ok" or lib.faces_strerror(0)
    for st in range(-7, 0):
        assert lib.faces_strerror(st)
    assert lib.faces_version()


@needs_lib
def test_plan_create_rejects_bad_arguments(lib):
    vp = ctypes.c_void_p()
    f = lib.faces_plan_create
    f.argtypes = [ctypes.POINTER(ctypes.c_void_p)] + [ctypes.c_int] * 9 + [ctypes.c_long]
    lib.faces_last_error.restype = ctypes.c_char_p
    cases = [(0, 1, 1, 1, 4, 4, 4, 1, 1, 0),     # n = 1
             (0, 1, 1, 1, 0, 4, 4, 5, 1, 0),     # mx = 0
             (4, 2, 1, 1, 4, 4, 4, 5, 1, 0),     # rank outside the grid
             (0, 1, 1, 1, 4, 4, 4, 5, 7, 0),     # unknown boundary
             (0, 0, 1, 1, 4, 4, 4, 5, 1, 0)]     # empty rank grid
    for c in cases:
        st = f(ctypes.byref(vp), *c, 0)
        assert st == -1, (c, st)
        assert lib.faces_last_error()
    assert f(None, 0, 1, 1, 1, 4, 4, 4, 5, 1, 0) == -1


@needs_lib
def test_field_doubles(lib):
    lib.faces_field_doubles.restype = ctypes.c_long
    lib.faces_field_doubles.argtypes = [ctypes.c_int] * 4 + [ctypes.c_long]
    assert lib.faces_field_doubles(5, 4, 4, 4, 0) == 8000
    assert lib.faces_field_doubles(7, 60, 60, 60, 0) == 74088000
    # padded planes: (mz - 1) strides plus one dense plane
    assert lib.faces_field_doubles(5, 4, 4, 4, 2048) == 3 * 2048 + 2000


@needs_lib
def test_batch_and_server_reject_bad_arguments(lib):
    vp = ctypes.c_void_p()
    ints = (ctypes.c_int * 1)
    lib.faces_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int] + \
        [ctypes.POINTER(ctypes.c_int)] * 4 + [ctypes.c_int]
    assert lib.faces_batch_create(ctypes.byref(vp), 0, ints(4), ints(2), ints(2), ints(2), 1) == -1
    assert lib.faces_batch_create(ctypes.byref(vp), 1, ints(1), ints(2), ints(2), ints(2), 1) == -1
    assert lib.faces_batch_create(ctypes.byref(vp), 1, ints(4), ints(2), ints(2), ints(2), 9) == -1
    lib.faces_server_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int]
    assert lib.faces_server_create(ctypes.byref(vp), 0.0, 4, 1) == -1
    assert lib.faces_server_create(ctypes.byref(vp), 1e9, 0, 1) == -1
    assert lib.faces_server_create(ctypes.byref(vp), 1e9, 4, 0) == -1
    assert lib.faces_batch_share(None, 1, None) == -1
    assert lib.faces_share(None, None, 1, None) == -1
