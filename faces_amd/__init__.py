"""faces_amd -- MI355X-native Faces exchange + DSS step.

Python binding of the C ABI in include/faces_amd.h (libfaces_amd.so, built for
gfx950 by ``python faces_amd/build.py``).  The API mirrors the reference's
step class (trey-ornl/faces gpu/v-26-messages/Faces.h:14-16):

    f = Faces(rank, lx, ly, lz, mx, my, mz, n, boundary="periodic")
    f.share(u, compute=True)

``u`` is a device buffer: a float64 torch tensor on the GPU or a raw device
address.  PyTorch is only used for device memory and streams.  Every call
goes to the HIP library; there is no CPU fallback, and importing this package
fails if the library is missing.
"""
from __future__ import annotations

import ctypes
import os
import re

__all__ = [
    "FacesError", "Faces", "Batch", "Server", "fill", "field_doubles", "stream_ceiling", "lib",
    "share_noalloc",
    "LIB_PATH", "OPEN", "PERIODIC", "PATH_AUTO", "PATH_FUSED", "PATH_STAGED",
    "PATH_STAGED_DIRECT", "PATH_SIX", "FILL_INT", "FILL_RAND", "FILL_REF", "header_symbols",
]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfaces_amd.so")
HEADER = os.path.join(ROOT, "include", "faces_amd.h")

OPEN, PERIODIC = 0, 1
PATH_AUTO, PATH_FUSED, PATH_STAGED, PATH_STAGED_DIRECT, PATH_SIX = 0, 1, 2, 3, 4
VARIANT_AUTO, VARIANT_DIRECT, VARIANT_LDS, VARIANT_STREAM, VARIANT_ITEMS = 0, 1, 2, 3, 4
(OPT_VARIANT, OPT_DIRECT_BLOCK, OPT_SEG, OPT_XCD, OPT_LDS_BLOCK, OPT_STREAM_BPC,
 OPT_STREAM_WALK, OPT_ACTIVE_KERNEL, OPT_STREAM_ROWS, OPT_STREAM_PLANES, OPT_FLOW,
 OPT_SURFACE_BLOCKS) = range(12)
FILL_INT, FILL_RAND, FILL_REF = 0, 1, 2

ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_ADMISSION, ERR_NODEV, ERR_TIMEOUT = (
    -1, -2, -3, -4, -5, -6, -7)


class FacesError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        super().__init__(f"{where}: {detail} (status {status})")
        self.status = status


if not os.path.exists(LIB_PATH):
    raise ImportError(f"faces_amd native library missing: {LIB_PATH} "
                      "(build it with: python faces_amd/build.py)")

# torch ships its own libamdhip64 (soname libamdhip64.so.7).  Loading torch
# first makes the dynamic loader resolve this library's HIP dependency to that
# same runtime, so device pointers and streams from torch are valid here.
try:
    import torch  # noqa: F401
except ImportError:  # the library also works without torch (raw device pointers)
    torch = None

_lib = ctypes.CDLL(LIB_PATH)
_c_int, _c_long, _c_double, _c_u64, _vp = (ctypes.c_int, ctypes.c_long, ctypes.c_double,
                                           ctypes.c_uint64, ctypes.c_void_p)
_dp = ctypes.POINTER(ctypes.c_double)


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("faces_strerror", ctypes.c_char_p, _c_int)
_sig("faces_last_error", ctypes.c_char_p)
_sig("faces_version", ctypes.c_char_p)
_sig("faces_plan_create", _c_int, ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _c_int,
     _c_int, _c_int, _c_int, _c_int, _c_long)
_sig("faces_plan_destroy", _c_int, _vp)
_sig("faces_share", _c_int, _vp, _vp, _c_int, _vp)
_sig("faces_share_noalloc", _c_int, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
     _c_int, _c_long, _c_int, _vp)
_sig("faces_plan_set_path", _c_int, _vp, _c_int)
_sig("faces_plan_set_tuning", _c_int, _vp, _c_int, _c_int, _c_int)
_sig("faces_plan_set_variant", _c_int, _vp, _c_int, _c_int)
_sig("faces_plan_set_option", _c_int, _vp, _c_int, _c_int)
_sig("faces_plan_get_option", _c_int, _vp, _c_int, ctypes.POINTER(_c_int))
_sig("faces_plan_tune", _c_int, _vp, _vp, _c_int, _vp, ctypes.POINTER(ctypes.c_double))
_sig("faces_plan_neighbor", _c_int, _vp, _c_int, ctypes.POINTER(_c_int))
_sig("faces_plan_message", _c_int, _vp, _c_int, _c_int, ctypes.POINTER(_vp), ctypes.POINTER(_c_long))
_sig("faces_pack", _c_int, _vp, _vp, _vp)
_sig("faces_exchange", _c_int, _vp, _vp)
_sig("faces_inner", _c_int, _vp, _vp, _vp)
_sig("faces_unpack", _c_int, _vp, _vp, _vp)
_sig("faces_inner_unpack", _c_int, _vp, _vp, _vp)
_sig("faces_fill", _c_int, _vp, _c_int, _c_int, _c_int, _c_int, _c_long, _c_int, _c_u64, _vp)
_sig("faces_field_doubles", _c_long, _c_int, _c_int, _c_int, _c_int, _c_long)
_sig("faces_stream_ceiling", _c_int, _vp, _c_long, _c_int, _vp, ctypes.POINTER(ctypes.c_double))


def lib() -> ctypes.CDLL:
    return _lib


def _check(status: int, where: str) -> None:
    if status != 0:
        detail = _lib.faces_last_error().decode() or _lib.faces_strerror(status).decode()
        raise FacesError(status, where, detail)


def _addr(u, need=None) -> int:
    """Device address of a torch tensor or an int.  A tensor must be a
    contiguous float64 GPU tensor of at least ``need`` elements (the kernels
    address it through raw pointers); a raw address is taken as is."""
    if isinstance(u, int):
        return u
    if hasattr(u, "data_ptr"):
        if not u.is_cuda:
            raise FacesError(ERR_ARG, "field", "tensor is not on the GPU")
        if torch is not None and u.dtype != torch.float64:
            raise FacesError(ERR_ARG, "field", f"tensor dtype {u.dtype} is not float64")
        if not u.is_contiguous():
            raise FacesError(ERR_ARG, "field", "tensor is not contiguous")
        if need is not None and u.numel() < need:
            raise FacesError(ERR_ARG, "field", f"tensor of {u.numel()} elements, the field needs {need}")
        return int(u.data_ptr())
    raise TypeError(f"cannot take a device address of {type(u)!r}")


def _stream(stream) -> int | None:
    if stream is None:
        try:
            import torch  # plumbing only: the current HIP stream
            if torch.cuda.is_available():
                return int(torch.cuda.current_stream().cuda_stream) or None
        except ImportError:
            pass
        return None
    if isinstance(stream, int):
        return stream or None
    return int(stream.cuda_stream) or None


def field_doubles(n, mx, my, mz, plane_stride=0) -> int:
    return int(_lib.faces_field_doubles(n, mx, my, mz, plane_stride))


def stream_ceiling(u, n=None, reps=20, stream=None) -> float:
    """In-place read+write GB/s of a plain streaming kernel over the first n
    doubles of u (all of a tensor by default); values are unchanged."""
    if n is None:
        n = u.numel()
    out = ctypes.c_double()
    _check(_lib.faces_stream_ceiling(_addr(u, n), n, reps, _stream(stream), ctypes.byref(out)),
           "faces_stream_ceiling")
    return out.value


def share_noalloc(u, rank, lx, ly, lz, mx, my, mz, n, boundary=PERIODIC, plane_stride=0,
                  compute=True, stream=None) -> None:
    """The staged step of one rank without a plan or any allocation
    (faces_share_noalloc): its messages live in the library's per-device
    scratch array, shared by every call on that device."""
    need = field_doubles(n, mx, my, mz, plane_stride)
    _check(_lib.faces_share_noalloc(_addr(u, need if need > 0 else None), rank, lx, ly, lz, mx, my, mz, n,
                                    boundary, plane_stride, int(bool(compute)), _stream(stream)),
           "faces_share_noalloc")


def fill(u, n, mx, my, mz, kind=FILL_INT, seed=42, plane_stride=0, stream=None) -> None:
    need = field_doubles(n, mx, my, mz, plane_stride)
    _check(_lib.faces_fill(_addr(u, need if need > 0 else None), n, mx, my, mz, plane_stride, kind,
                           seed, _stream(stream)),
           "faces_fill")


class Faces:
    """The plain step: mirrors class Faces (gpu/v-26-messages/Faces.h:6-39)."""

    def __init__(self, rank, lx, ly, lz, mx, my, mz, n, boundary=PERIODIC, plane_stride=0):
        if isinstance(boundary, str):
            boundary = {"open": OPEN, "periodic": PERIODIC}[boundary]
        h = _vp()
        _check(_lib.faces_plan_create(ctypes.byref(h), rank, lx, ly, lz, mx, my, mz, n, boundary,
                                      plane_stride), "faces_plan_create")
        self._h = h
        self.rank, self.lx, self.ly, self.lz = rank, lx, ly, lz
        self.mx, self.my, self.mz, self.n = mx, my, mz, n
        self.boundary, self.plane_stride = boundary, plane_stride
        self.doubles = field_doubles(n, mx, my, mz, plane_stride)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(_lib.faces_plan_destroy(self._h), "faces_plan_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def share(self, u, compute=True, stream=None) -> None:
        _check(_lib.faces_share(self._h, _addr(u, self.doubles), int(bool(compute)), _stream(stream)),
               "faces_share")

    def set_path(self, path: int) -> None:
        _check(_lib.faces_plan_set_path(self._h, path), "faces_plan_set_path")

    def set_tuning(self, block=256, seg_len=0, xcd_remap=1) -> None:
        _check(_lib.faces_plan_set_tuning(self._h, block, seg_len, xcd_remap),
               "faces_plan_set_tuning")

    def set_variant(self, variant: int, block: int = 0) -> None:
        _check(_lib.faces_plan_set_variant(self._h, variant, block), "faces_plan_set_variant")

    def set_option(self, option: int, value: int) -> None:
        _check(_lib.faces_plan_set_option(self._h, option, value), "faces_plan_set_option")

    def get_option(self, option: int) -> int:
        v = _c_int()
        _check(_lib.faces_plan_get_option(self._h, option, ctypes.byref(v)), "faces_plan_get_option")
        return v.value

    def tune(self, u, reps=5, stream=None) -> dict:
        """Pick the fastest launch options on this grid (shares u 1 + reps
        times per trial: tune on a scratch copy)."""
        out = (ctypes.c_double * 5)()
        _check(_lib.faces_plan_tune(self._h, _addr(u, self.doubles), reps, _stream(stream), out),
               "faces_plan_tune")
        return {"variant": int(out[0]), "walk": int(out[1]), "seg_len": int(out[2]),
                "blocks_per_cu": int(out[3]), "us_per_share": out[4]}

    def neighbor(self, slot: int) -> int:
        r = _c_int()
        _check(_lib.faces_plan_neighbor(self._h, slot, ctypes.byref(r)), "faces_plan_neighbor")
        return r.value

    def message(self, slot: int, recv: bool = False):
        p, c = _vp(), _c_long()
        _check(_lib.faces_plan_message(self._h, slot, int(recv), ctypes.byref(p), ctypes.byref(c)),
               "faces_plan_message")
        return int(p.value or 0), int(c.value)

    def pack(self, u, stream=None):
        _check(_lib.faces_pack(self._h, _addr(u, self.doubles), _stream(stream)), "faces_pack")

    def exchange(self, stream=None):
        _check(_lib.faces_exchange(self._h, _stream(stream)), "faces_exchange")

    def inner(self, u, stream=None):
        _check(_lib.faces_inner(self._h, _addr(u, self.doubles), _stream(stream)), "faces_inner")

    def unpack(self, u, stream=None):
        _check(_lib.faces_unpack(self._h, _addr(u, self.doubles), _stream(stream)), "faces_unpack")

    def inner_unpack(self, u, stream=None):
        """inner and unpack in one launch, once the messages have arrived."""
        _check(_lib.faces_inner_unpack(self._h, _addr(u, self.doubles), _stream(stream)),
               "faces_inner_unpack")


def header_symbols() -> list[str]:
    """Function names declared in include/faces_amd.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(faces_\w+)\s*\(", text, re.M)))


from .batch import Batch, Server  # noqa: E402  (binds the batched/served entry points)
