"""Batched and served forms of the step (bindings of the faces_batch_* and
faces_server_* entry points in include/faces_amd.h)."""
from __future__ import annotations

import ctypes

from . import FacesError, _addr, _check, _lib, _stream, field_doubles, PERIODIC, OPEN  # noqa: F401

_c_int, _c_long, _c_double, _c_u64, _vp = (ctypes.c_int, ctypes.c_long, ctypes.c_double,
                                           ctypes.c_uint64, ctypes.c_void_p)
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "faces_batch_create": (_c_int, ctypes.POINTER(_vp), _c_int, _ip, _ip, _ip, _ip, _c_int),
    "faces_batch_destroy": (_c_int, _vp),
    "faces_batch_set_fields": (_c_int, _vp, ctypes.POINTER(_vp)),
    "faces_batch_share": (_c_int, _vp, _c_int, _vp),
    "faces_batch_share_per_grid": (_c_int, _vp, _c_int, _vp),
    "faces_batch_set_tuning": (_c_int, _vp, _c_int, _c_int, _c_int),
    "faces_batch_tune": (_c_int, _vp, _c_int, _vp, _dp),
    "faces_batch_info": (_c_int, _vp, _dp, ctypes.POINTER(_c_long)),
    "faces_server_create": (_c_int, ctypes.POINTER(_vp), _c_double, _c_int, _c_int),
    "faces_server_destroy": (_c_int, _vp),
    "faces_server_submit": (_c_int, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp,
                            ctypes.POINTER(_c_u64)),
    "faces_server_wait": (_c_int, _vp, _c_u64, _c_double, _dp),
    "faces_server_run_trace": (_c_int, _vp, _c_int, _dp, _ip, _ip, _ip, _ip, _c_int, _c_int,
                               ctypes.POINTER(_vp), _dp, _dp),
    "faces_server_stats": (_c_int, _vp, _dp),
    "faces_server_set_tuning": (_c_int, _vp, _c_int),
}


def _fn(name):
    f = getattr(_lib, name)
    res, *args = _SIGS[name]
    f.restype = res
    f.argtypes = args
    return f


def _ints(xs):
    return (ctypes.c_int * len(xs))(*[int(x) for x in xs])


class Batch:
    """Many independent grids (dense fields) stepped by one launch."""

    def __init__(self, shapes, boundary=PERIODIC):
        """shapes: sequence of (n, mx, my, mz)."""
        shapes = [tuple(int(v) for v in s) for s in shapes]
        self.shapes = shapes
        h = _vp()
        ns, mxs, mys, mzs = zip(*shapes) if shapes else ((), (), (), ())
        _check(_fn("faces_batch_create")(ctypes.byref(h), len(shapes), _ints(ns), _ints(mxs),
                                         _ints(mys), _ints(mzs), boundary), "faces_batch_create")
        self._h = h
        self._fields = None

    def close(self):
        if getattr(self, "_h", None):
            _check(_fn("faces_batch_destroy")(self._h), "faces_batch_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, fields):
        if len(fields) != len(self.shapes):
            raise FacesError(-1, "faces_batch_set_fields",
                             f"{len(fields)} fields for {len(self.shapes)} grids")
        ptrs = (_vp * len(fields))(*[_addr(u, field_doubles(*sh)) for u, sh in zip(fields, self.shapes)])
        self._fields = list(fields)  # keep the tensors alive
        _check(_fn("faces_batch_set_fields")(self._h, ptrs), "faces_batch_set_fields")

    def share(self, compute=True, stream=None):
        _check(_fn("faces_batch_share")(self._h, int(bool(compute)), _stream(stream)),
               "faces_batch_share")

    def share_per_grid(self, compute=True, stream=None):
        _check(_fn("faces_batch_share_per_grid")(self._h, int(bool(compute)), _stream(stream)),
               "faces_batch_share_per_grid")

    def set_tuning(self, block=256, seg_len=0, blocks_per_cu=0):
        _check(_fn("faces_batch_set_tuning")(self._h, block, seg_len, blocks_per_cu),
               "faces_batch_set_tuning")

    def tune(self, reps=5, stream=None):
        out = (ctypes.c_double * 4)()
        _check(_fn("faces_batch_tune")(self._h, reps, _stream(stream), out), "faces_batch_tune")
        return {"block": int(out[0]), "seg_len": int(out[1]), "blocks_per_cu": int(out[2]),
                "us_per_step": out[3]}

    def info(self):
        b, it = ctypes.c_double(), _c_long()
        _check(_fn("faces_batch_info")(self._h, ctypes.byref(b), ctypes.byref(it)),
               "faces_batch_info")
        return {"bytes": b.value, "items": it.value}


class Server:
    """Continuous-batching request loop (faces_server_*)."""

    def __init__(self, footprint_cap=200e9, max_batch=256, nstreams=2):
        h = _vp()
        _check(_fn("faces_server_create")(ctypes.byref(h), float(footprint_cap), max_batch,
                                          nstreams), "faces_server_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _check(_fn("faces_server_destroy")(self._h), "faces_server_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, u, n, mx, my, mz, boundary=PERIODIC, compute=True) -> int:
        t = _c_u64()
        _check(_fn("faces_server_submit")(self._h, n, mx, my, mz, boundary, int(bool(compute)),
                                          _addr(u, field_doubles(n, mx, my, mz)), ctypes.byref(t)),
               "faces_server_submit")
        return t.value

    def wait(self, ticket, timeout_s=60.0) -> float:
        lat = ctypes.c_double()
        _check(_fn("faces_server_wait")(self._h, ticket, timeout_s, ctypes.byref(lat)),
               "faces_server_wait")
        return lat.value

    def run_trace(self, arrivals_us, shapes, fields, boundary=PERIODIC, compute=True):
        nreq = len(arrivals_us)
        arr = (ctypes.c_double * nreq)(*arrivals_us)
        ns, mxs, mys, mzs = zip(*shapes)
        ptrs = (_vp * nreq)(*[_addr(u, field_doubles(*sh)) for u, sh in zip(fields, shapes)])
        lat = (ctypes.c_double * nreq)()
        stats = (ctypes.c_double * 6)()
        _check(_fn("faces_server_run_trace")(self._h, nreq, arr, _ints(ns), _ints(mxs), _ints(mys),
                                             _ints(mzs), boundary, int(bool(compute)), ptrs, lat,
                                             stats), "faces_server_run_trace")
        return list(lat), {"wall_s": stats[0], "req_per_s": stats[1], "p50_us": stats[2],
                           "p99_us": stats[3], "max_us": stats[4], "launches": stats[5]}

    def set_tuning(self, blocks_per_cu):
        """Blocks per CU of batched launches (1-256), or -1: one block per item."""
        _check(_fn("faces_server_set_tuning")(self._h, blocks_per_cu), "faces_server_set_tuning")

    def stats(self):
        s = (ctypes.c_double * 4)()
        _check(_fn("faces_server_stats")(self._h, s), "faces_server_stats")
        return {"admitted": s[0], "completed": s[1], "launches": s[2], "peak_bytes": s[3]}
