"""ctypes wrapper of the CPU oracle (faces_oracle.c).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  It checks results; it is never the thing
measured or shipped.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

FILL_INT, FILL_RAND, FILL_REF = 0, 1, 2

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} missing (run: make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i, l, u64 = ctypes.c_int, ctypes.c_long, ctypes.c_uint64
        L.orc_hash.argtypes = [u64, u64]
        L.orc_hash.restype = u64
        L.orc_field_doubles.argtypes = [i, i, i, i, l]
        L.orc_field_doubles.restype = l
        L.orc_fill.argtypes = [dp, i, i, i, i, l, i, u64]
        L.orc_fill.restype = None
        L.orc_share_sweep.argtypes = [dp, i, i, i, i, l, i, i]
        L.orc_share_sweep.restype = None
        L.orc_share_staged.argtypes = [dp, i, i, i, i, l, i, i]
        L.orc_share_staged.restype = i
        L.orc_share_mugs26.argtypes = [dp, i, i, i, i, l, i, i]
        L.orc_share_mugs26.restype = i
        L.orc_share_six.argtypes = [dp, i, i, i, i, l, i, i]
        L.orc_share_six.restype = i
        L.orc_pack.argtypes = [dp, i, i, i, i, l, dp, ctypes.POINTER(l)]
        L.orc_pack.restype = i
        L.orc_compare.argtypes = [dp, dp, i, i, i, i, l, dp]
        L.orc_compare.restype = i
        L.orc_opposite_slot.argtypes = [i]
        L.orc_opposite_slot.restype = i
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = i
        L.orc_set_num_threads.argtypes = [i]
        L.orc_set_num_threads.restype = None
        L.orc_omp_binding.argtypes = [ctypes.POINTER(i)]
        L.orc_omp_binding.restype = None
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def field_doubles(n, mx, my, mz, ps=0) -> int:
    return int(lib().orc_field_doubles(n, mx, my, mz, ps))


def new_field(n, mx, my, mz, kind=FILL_INT, seed=42, ps=0) -> np.ndarray:
    """Seeded field as a flat float64 array (padding slots zero)."""
    u = np.zeros(field_doubles(n, mx, my, mz, ps), dtype=np.float64)
    lib().orc_fill(_ptr(u), n, mx, my, mz, ps, kind, seed)
    return u


def fill(u: np.ndarray, n, mx, my, mz, kind=FILL_INT, seed=42, ps=0) -> None:
    lib().orc_fill(_ptr(u), n, mx, my, mz, ps, kind, seed)


def share_sweep(u: np.ndarray, n, mx, my, mz, periodic=True, compute=True, ps=0) -> None:
    """In place: gpu/Mugs.cc:136-203 (+ periodic wrap pairs)."""
    lib().orc_share_sweep(_ptr(u), n, mx, my, mz, ps, int(periodic), int(compute))


def share_six(u: np.ndarray, n, mx, my, mz, nbr_mask=63, compute=True, ps=0) -> None:
    """In place: the 6-message schedule, z then y then x phases
    (gpu/v-6-messages-basic/Faces.cc:65-172), self halos per nbr_mask."""
    if lib().orc_share_six(_ptr(u), n, mx, my, mz, ps, nbr_mask, int(compute)) != 0:
        raise ValueError("orc_share_six: bad arguments")


def share_staged(u: np.ndarray, n, mx, my, mz, nbr_mask=63, compute=True, ps=0) -> None:
    """In place: pack -> self-halo exchange -> tree over local + received copies."""
    st = lib().orc_share_staged(_ptr(u), n, mx, my, mz, ps, nbr_mask, int(compute))
    if st != 0:
        raise ValueError("bad shape")


def share_mugs26(u: np.ndarray, n, mx, my, mz, nbr_mask=63, compute=True, ps=0) -> None:
    """In place: multi-rank Mugs order (gpu/Mugs.cc:86-416) with self halos."""
    st = lib().orc_share_mugs26(_ptr(u), n, mx, my, mz, ps, nbr_mask, int(compute))
    if st != 0:
        raise ValueError("bad shape")


def pack(u: np.ndarray, n, mx, my, mz, ps=0):
    """The 26 send messages: (flat array, offsets[27])."""
    offs = (ctypes.c_long * 27)()
    lib().orc_pack(_ptr(u), n, mx, my, mz, ps, None, offs)
    out = np.zeros(offs[26], dtype=np.float64)
    lib().orc_pack(_ptr(u), n, mx, my, mz, ps, _ptr(out), offs)
    return out, [int(o) for o in offs]


def compare(u: np.ndarray, v: np.ndarray, n, mx, my, mz, ps=0) -> dict:
    """gpu/main.cc:39-105 criterion plus bit and ulp counts."""
    out = np.zeros(4, dtype=np.float64)
    ok = lib().orc_compare(_ptr(u), _ptr(v), n, mx, my, mz, ps, _ptr(out))
    return {"pass": bool(ok), "max_rel": out[0], "argmax": int(out[1]),
            "ndiff": int(out[2]), "max_ulp": out[3]}


def hash64(seed: int, idx: int) -> int:
    return int(lib().orc_hash(seed, idx))


def opposite_slot(k: int) -> int:
    return int(lib().orc_opposite_slot(k))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(t: int) -> None:
    lib().orc_set_num_threads(int(t))


def omp_binding() -> dict:
    """The OpenMP runtime's thread binding: proc_bind name and place count."""
    out = (ctypes.c_int * 2)()
    lib().orc_omp_binding(out)
    names = {0: "false", 1: "true", 2: "master", 3: "close", 4: "spread"}
    return {"proc_bind": names.get(out[0], str(out[0])), "places": int(out[1])}
