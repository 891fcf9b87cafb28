"""Several ranks of an open decomposition on one GPU, each with its own plan
and field, exchanging their 26 messages through the staged pieces the way an
external transport would (faces_pack, a device copy of send[d] of each rank
into recv[opposite d] of its neighbour d, faces_inner, faces_unpack; the MPI
pattern of gpu/v-26-messages/Faces.cc:144-171, 386-417 with per-direction
slots).  The ranks together equal the sweep of the whole grid (open, or
periodic), bit for bit."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

import faces_amd as fa  # noqa: E402


def hip_runtime():
    """The HIP runtime torch loaded (the library's HIP calls resolve to it)."""
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line:
            lib = ctypes.CDLL(line.split()[-1])
            lib.hipMemcpy.restype = ctypes.c_int
            lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            return lib
    raise RuntimeError("HIP runtime not loaded")


def whole_grid(fields, lgrid, shape):
    """The ranks' fields as one field of the global grid (rank = (kz*ly + ky)*lx + kx)."""
    n, mx, my, mz = shape
    lx, ly, lz = lgrid
    blocks = [f.reshape(mz, my, mx, n, n, n) for f in fields]
    out = np.empty((mz * lz, my * ly, mx * lx, n, n, n))
    for r, b in enumerate(blocks):
        kz, rem = divmod(r, lx * ly)
        ky, kx = divmod(rem, lx)
        out[kz * mz:(kz + 1) * mz, ky * my:(ky + 1) * my, kx * mx:(kx + 1) * mx] = b
    return out


@pytest.mark.parametrize("lgrid,shape,boundary", [
    ((2, 1, 1), (5, 3, 2, 2), fa.OPEN), ((1, 2, 1), (4, 2, 3, 2), fa.OPEN),
    ((1, 1, 2), (7, 2, 2, 3), fa.OPEN), ((2, 2, 2), (6, 2, 3, 2), fa.OPEN),
    ((3, 2, 1), (3, 4, 1, 2), fa.OPEN),
    # periodic: with two ranks along an axis both directions lead to the
    # same neighbour, and the per-direction slots keep the messages apart
    ((2, 1, 1), (5, 3, 2, 2), fa.PERIODIC), ((2, 2, 2), (4, 2, 2, 2), fa.PERIODIC),
    ((1, 3, 1), (6, 2, 2, 3), fa.PERIODIC)])
@pytest.mark.parametrize("finish", ["inner+unpack", "inner_unpack", "inner_unpack blocks"])
def test_ranks_exchange_through_pieces(torch_cuda, lgrid, shape, boundary, finish):
    """After the exchange each rank finishes with faces_inner then
    faces_unpack, or with faces_inner_unpack: one launch whose items sum the
    rank surface from the messages other ranks sent (the default), or with
    the surface sums as extra blocks of that launch."""
    n, mx, my, mz = shape
    lx, ly, lz = lgrid
    nr = lx * ly * lz
    print("case transport", lgrid, shape, boundary, finish, flush=True)
    hip = hip_runtime()
    host = [oracle.new_field(*shape, fa.FILL_RAND, seed=50 + r) for r in range(nr)]
    dev = [torch_cuda.from_numpy(h.copy()).cuda() for h in host]
    plans = [fa.Faces(r, lx, ly, lz, mx, my, mz, n, boundary=boundary) for r in range(nr)]
    for p, d in zip(plans, dev):
        p.pack(d)
    torch_cuda.cuda.synchronize()
    for r, p in enumerate(plans):
        for k in range(26):
            q = p.neighbor(k)
            if q < 0:
                continue
            src, cnt = p.message(k)
            dst, cnt2 = plans[q].message(oracle.opposite_slot(k), recv=True)
            assert cnt == cnt2 and cnt > 0
            assert hip.hipMemcpy(dst, src, 8 * cnt, 3) == 0  # device to device
    torch_cuda.cuda.synchronize()
    for p, d in zip(plans, dev):
        if finish == "inner+unpack":
            p.inner(d)
            p.unpack(d)
        else:
            if finish.endswith("blocks"):
                p.set_option(fa.OPT_SURFACE_BLOCKS, 2)
            p.inner_unpack(d)
    torch_cuda.cuda.synchronize()
    got = [d.cpu().numpy() for d in dev]
    for p in plans:
        p.close()
    want = whole_grid(host, lgrid, shape)
    gshape = (n, mx * lx, my * ly, mz * lz)
    flat = np.ascontiguousarray(want).reshape(-1)
    oracle.share_sweep(flat, *gshape, periodic=boundary == fa.PERIODIC)
    got_whole = whole_grid(got, lgrid, shape).reshape(-1)
    r = oracle.compare(got_whole, flat, *gshape)
    assert r["ndiff"] == 0, r
