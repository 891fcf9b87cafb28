"""GPU parity of the plain step (faces_share and its staged pieces) against the
oracle, through the C ABI.  Bar: bit-exact -- the kernels evaluate the same
x-then-y-then-z tree as the oracle, so random FP64 fields match bit for bit,
not only integer-valued ones."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests import _brute as brute

pytestmark = pytest.mark.gpu

import faces_amd as fa  # noqa: E402


def to_dev(torch, u: np.ndarray):
    return torch.from_numpy(u).to("cuda")


def to_host(torch, d) -> np.ndarray:
    torch.cuda.synchronize()
    return d.cpu().numpy()


def run_gpu(torch, u, shape, boundary=fa.PERIODIC, compute=True, path=fa.PATH_AUTO, ps=0,
            rank=0, lgrid=(1, 1, 1), nshare=1, tuning=None, variant=None):
    n, mx, my, mz = shape
    d = to_dev(torch, u)
    f = fa.Faces(rank, *lgrid, mx, my, mz, n, boundary=boundary, plane_stride=ps)
    f.set_path(path)
    if tuning:
        f.set_tuning(*tuning)
    if variant:
        f.set_variant(*variant)
    for _ in range(nshare):
        f.share(d, compute=compute)
    out = to_host(torch, d)
    f.close()
    return out


def run_oracle(u, shape, periodic=True, compute=True, ps=0, nshare=1):
    v = u.copy()
    for _ in range(nshare):
        oracle.share_sweep(v, *shape, periodic=periodic, compute=compute, ps=ps)
    return v


def assert_same(a, b, shape, ps=0):
    r = oracle.compare(a, b, *shape, ps=ps)
    assert r["ndiff"] == 0, r


SHAPES = [(2, 1, 1, 1), (2, 3, 2, 2), (3, 2, 3, 1), (4, 3, 2, 2), (5, 4, 4, 4), (6, 2, 1, 3),
          (7, 3, 3, 2), (8, 2, 2, 2), (9, 2, 2, 1), (10, 1, 2, 2), (11, 2, 1, 1), (12, 2, 2, 2),
          (13, 2, 1, 2), (5, 1, 1, 1), (3, 7, 1, 1)]


@pytest.mark.parametrize("kind", [fa.FILL_INT, fa.FILL_RAND, fa.FILL_REF])
def test_fill_matches_oracle(torch_cuda, kind):
    shape = (7, 5, 3, 2)
    want = oracle.new_field(*shape, kind, seed=42)
    d = torch_cuda.zeros(want.size, dtype=torch_cuda.float64, device="cuda")
    fa.fill(d, *shape, kind=kind, seed=42)
    assert np.array_equal(to_host(torch_cuda, d), want)


def test_fill_padded(torch_cuda):
    shape = (5, 3, 2, 3)
    n, mx, my, mz = shape
    ps = n ** 3 * mx * my + 512
    want = oracle.new_field(*shape, fa.FILL_RAND, seed=3, ps=ps)
    d = torch_cuda.zeros(want.size, dtype=torch_cuda.float64, device="cuda")
    fa.fill(d, *shape, kind=fa.FILL_RAND, seed=3, plane_stride=ps)
    assert np.array_equal(to_host(torch_cuda, d), want)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", [fa.FILL_INT, fa.FILL_RAND])
def test_periodic_fused_bitexact(torch_cuda, shape, kind):
    print("case periodic fused", shape, kind, flush=True)
    u = oracle.new_field(*shape, kind, seed=101)
    assert_same(run_gpu(torch_cuda, u, shape), run_oracle(u, shape), shape)


@pytest.mark.parametrize("shape", SHAPES[:-3] + [(5, 1, 1, 1), (3, 7, 1, 1), (4, 1, 5, 1),
                                                  (4, 1, 1, 6), (6, 3, 1, 1), (2, 1, 1, 2)])
@pytest.mark.parametrize("variant", [(fa.VARIANT_DIRECT, 0), (fa.VARIANT_LDS, 64),
                                     (fa.VARIANT_LDS, 256), (fa.VARIANT_LDS, 512),
                                     (fa.VARIANT_STREAM, 1), (fa.VARIANT_STREAM, 8),
                                     (fa.VARIANT_STREAM, -1), (fa.VARIANT_ITEMS, 1),
                                     (fa.VARIANT_ITEMS, 4), (fa.VARIANT_ITEMS, 32),
                                     (fa.VARIANT_ITEMS, 256)])
def test_periodic_variants_bitexact(torch_cuda, shape, variant):
    print("case periodic variant", shape, variant, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=202)
    assert_same(run_gpu(torch_cuda, u, shape, variant=variant), run_oracle(u, shape), shape)


@pytest.mark.parametrize("walk,rows", [(0, 1), (1, 1), (2, 1), (2, 2), (2, 3), (2, 64), (3, 1), (3, 2),
                                       (3, 3), (3, 4)])
@pytest.mark.parametrize("seg,bpc,xcd", [(1, 0, 1), (2, 1, 0), (3, 2, 1), (4, 3, 1), (4, -2, 1), (2, -3, 0),
                                         (3, -4, 1)])
@pytest.mark.parametrize("shape", [(7, 12, 6, 3), (8, 4, 3, 5), (5, 6, 1, 2), (4, 3, 4, 1),
                                   (6, 2, 2, 2)])
def test_stream_walks_bitexact(torch_cuda, walk, rows, seg, bpc, xcd, shape):
    n, mx, my, mz = shape
    print("case stream walk", walk, rows, seg, bpc, xcd, shape, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=303)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
    f.set_option(fa.OPT_FLOW, 0)  # the streaming kernel's walks, not the one-item kernel
    f.set_option(fa.OPT_STREAM_WALK, walk)
    f.set_option(fa.OPT_STREAM_ROWS, rows)
    f.set_option(fa.OPT_SEG, seg)
    f.set_option(fa.OPT_STREAM_BPC, bpc)
    f.set_option(fa.OPT_XCD, xcd)
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == fa.VARIANT_STREAM
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, nshare=2), shape)


# --------------------------------- the one-item kernel (the default periodic step)

FLOW_SHAPES = [(7, 12, 6, 3), (7, 10, 3, 2), (8, 4, 3, 5), (5, 10, 2, 3), (5, 6, 1, 2), (4, 16, 4, 1),
               (6, 5, 2, 2), (2, 32, 2, 1), (3, 16, 3, 2), (9, 3, 2, 2), (10, 2, 3, 1), (11, 2, 1, 2),
               (12, 1, 2, 2), (7, 1, 1, 1), (7, 15, 9, 2), (3, 64, 2, 2), (2, 128, 2, 1)]


@pytest.mark.parametrize("flow", [1, 2, 3, 4])
@pytest.mark.parametrize("xcd", [0, 1, 2])
@pytest.mark.parametrize("shape", FLOW_SHAPES)
@pytest.mark.parametrize("seg", [0, 1, 3, 5])
def test_flow_bitexact(torch_cuda, flow, xcd, shape, seg):
    """The one-item kernel: 8- and 16-byte loads, each XCD order, item
    lengths whose first element starts on an even or an odd double (odd n,
    odd segment starts), planes with fewer items than XCDs; two shares of
    a random field, bit-exact against the oracle."""
    n, mx, my, mz = shape
    print("case flow", flow, xcd, shape, seg, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=313)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
    f.set_option(fa.OPT_FLOW, flow)
    f.set_option(fa.OPT_XCD, xcd)
    f.set_option(fa.OPT_SEG, seg)
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == fa.VARIANT_STREAM
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, nshare=2), shape)


@pytest.mark.parametrize("flow", [1, 2, 3])
@pytest.mark.parametrize("shape,extra", [((7, 6, 5, 4), 1000), ((7, 5, 3, 3), 1), ((6, 4, 3, 2), 3),
                                         ((5, 10, 2, 3), 512), ((8, 3, 3, 2), 7)])
def test_flow_padded_planes(torch_cuda, flow, shape, extra):
    """Padded plane strides, odd ones included (planes then start on odd
    doubles, so the 16-byte loads change phase from plane to plane)."""
    n, mx, my, mz = shape
    ps = n ** 3 * mx * my + extra
    print("case flow padded", flow, shape, ps, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=17, ps=ps)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC, plane_stride=ps)
    f.set_option(fa.OPT_FLOW, flow)
    for _ in range(3):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert np.array_equal(got.view(np.uint64), run_oracle(u, shape, ps=ps, nshare=3).view(np.uint64))


@pytest.mark.parametrize("flow", [2, 3, 4])
def test_flow_misaligned_field(torch_cuda, flow):
    """A field that starts 8 bytes into its allocation: flow 2 does not take
    its 16-byte aligned loads (it falls back to 8-byte lanes), flows 3 and 4
    move their pairs from any double; results exact, the bytes around
    untouched."""
    shape = (7, 10, 4, 3)
    n, mx, my, mz = shape
    print("case flow misaligned", flow, shape, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=23)
    buf = torch_cuda.zeros(u.size + 2, dtype=torch_cuda.float64, device="cuda")
    buf[1:1 + u.size] = torch_cuda.from_numpy(u).to("cuda")
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    assert f.get_option(fa.OPT_FLOW) == 4  # the default for n = 7
    f.set_option(fa.OPT_FLOW, flow)
    f.share(int(buf.data_ptr()) + 8)
    got = to_host(torch_cuda, buf)
    f.close()
    assert got[0] == 0.0 and got[-1] == 0.0
    assert_same(got[1:1 + u.size].copy(), run_oracle(u, shape), shape)


def test_periodic_lds_padded_and_repeated(torch_cuda):
    shape = (7, 6, 5, 4)
    n, mx, my, mz = shape
    ps = n ** 3 * mx * my + 1000
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=9, ps=ps)
    got = run_gpu(torch_cuda, u, shape, ps=ps, nshare=3, variant=(fa.VARIANT_LDS, 256))
    assert np.array_equal(got, run_oracle(u, shape, ps=ps, nshare=3))


@pytest.mark.parametrize("shape", SHAPES)
def test_open_bitexact(torch_cuda, shape):
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=5)
    got = run_gpu(torch_cuda, u, shape, boundary=fa.OPEN)
    assert_same(got, run_oracle(u, shape, periodic=False), shape)


@pytest.mark.parametrize("walk,rows,bpc", [(0, 1, 2), (1, 1, 2), (2, 2, 2), (2, 0, -2), (3, 2, 2),
                                           (1, 1, -3)])
@pytest.mark.parametrize("shape,seg", [((7, 12, 6, 4), 4), ((7, 1, 1, 1), 1), ((5, 10, 3, 2), 5),
                                       ((4, 32, 2, 3), 16), ((8, 3, 4, 5), 3), ((12, 2, 2, 1), 1),
                                       ((2, 4, 1, 3), 4), ((6, 8, 5, 1), 8)])
@pytest.mark.parametrize("planes", [0, 1])
def test_open_rank_streaming_bitexact(torch_cuda, walk, rows, bpc, shape, seg, planes):
    """Open boundary on a 1x1x1 rank grid (the reference's own mode) in one
    streaming launch: each rank-surface node sums only the copies on its side
    of the surface; equal to the open sweep for every traversal, with 32-bit
    field offsets and with the per-plane addressing of fields of 4 GB and more."""
    n, mx, my, mz = shape
    print("case open streaming", shape, seg, walk, rows, bpc, planes, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=808)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.OPEN)
    f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
    f.set_option(fa.OPT_FLOW, 0)  # the streaming kernel's walks
    f.set_option(fa.OPT_SEG, seg)
    f.set_option(fa.OPT_STREAM_WALK, walk)
    f.set_option(fa.OPT_STREAM_ROWS, rows)
    f.set_option(fa.OPT_STREAM_BPC, bpc)
    f.set_option(fa.OPT_STREAM_PLANES, planes)
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == fa.VARIANT_STREAM
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, periodic=False, nshare=2), shape)


@pytest.mark.parametrize("shape", [(5, 4, 4, 4), (7, 3, 3, 2), (2, 2, 3, 1), (12, 2, 1, 2)])
@pytest.mark.parametrize("periodic", [True, False])
def test_compute_false(torch_cuda, shape, periodic):
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=8)
    got = run_gpu(torch_cuda, u, shape, boundary=fa.PERIODIC if periodic else fa.OPEN,
                  compute=False)
    assert_same(got, run_oracle(u, shape, periodic=periodic, compute=False), shape)


@pytest.mark.parametrize("path", [fa.PATH_FUSED, fa.PATH_STAGED, fa.PATH_STAGED_DIRECT])
@pytest.mark.parametrize("shape", [(5, 4, 4, 4), (7, 3, 2, 2), (2, 1, 2, 1), (8, 2, 3, 2)])
@pytest.mark.parametrize("compute", [True, False])
def test_paths_agree_periodic(torch_cuda, path, shape, compute):
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=13)
    got = run_gpu(torch_cuda, u, shape, compute=compute, path=path)
    assert_same(got, run_oracle(u, shape, compute=compute), shape)


@pytest.mark.parametrize("tuning", [(64, 0, 0), (128, 1, 1), (512, 2, 1), (1024, 3, 0),
                                    (256, 0, 1), (256, 6, 1), (256, 12, 1)])
def test_tuning_does_not_change_bits(torch_cuda, tuning):
    shape = (7, 5, 3, 4)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=77)
    got = run_gpu(torch_cuda, u, shape, tuning=tuning)
    assert_same(got, run_oracle(u, shape), shape)


def test_padded_plane_stride(torch_cuda):
    shape = (7, 3, 2, 3)
    n, mx, my, mz = shape
    plane = n ** 3 * mx * my
    ps = ((plane - 1) // 512 + 1) * 512  # gpu/DArray.h:95-99
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=19, ps=ps)
    for periodic in (True, False):
        got = run_gpu(torch_cuda, u, shape, ps=ps,
                      boundary=fa.PERIODIC if periodic else fa.OPEN)
        want = run_oracle(u, shape, periodic=periodic, ps=ps)
        assert np.array_equal(got, want)  # padding slots included


def test_pack_messages_match_oracle(torch_cuda):
    shape = (5, 3, 4, 2)
    n, mx, my, mz = shape
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=31)
    msg, off = oracle.pack(u, *shape)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    f.pack(d)
    torch_cuda.cuda.synchronize()
    base, _ = f.message(0)
    arena_end, cnt_end = f.message(25, recv=True)
    nbytes = arena_end + 8 * cnt_end - base
    import ctypes
    host = np.zeros(nbytes // 8)
    # copy the whole arena to the host through torch
    t = torch_cuda.empty(host.size, dtype=torch_cuda.float64, device="cuda")
    torch_cuda.cuda.synchronize()
    lib = ctypes.CDLL("libamdhip64.so.7")  # the runtime torch already loaded
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert lib.hipMemcpy(host.ctypes.data, base, nbytes, 2) == 0  # device -> host
    del t
    for k in range(26):
        p, c = f.message(k)
        o = (p - base) // 8
        assert c == off[k + 1] - off[k]
        assert np.array_equal(host[o:o + c], msg[off[k]:off[k + 1]]), k
    # exchange: recv[k] == send[opposite(k)]
    f.exchange()
    torch_cuda.cuda.synchronize()
    assert lib.hipMemcpy(host.ctypes.data, base, nbytes, 2) == 0
    for k in range(26):
        p, c = f.message(k, recv=True)
        o = (p - base) // 8
        ok = oracle.opposite_slot(k)
        assert np.array_equal(host[o:o + c], msg[off[ok]:off[ok + 1]]), k
    f.close()


@pytest.mark.parametrize("lgrid", [(3, 3, 3), (2, 1, 3), (1, 2, 1)])
def test_rank_positions_open(torch_cuda, lgrid):
    """One GPU standing in for each rank of an open decomposition: existing
    neighbours are self halos; missing ones are MPI_PROC_NULL."""
    lx, ly, lz = lgrid
    shape = (5, 2, 3, 2)
    n, mx, my, mz = shape
    for rank in range(lx * ly * lz):
        mask = brute.mask_for_rank(rank, lx, ly, lz)
        u = oracle.new_field(*shape, fa.FILL_RAND, seed=rank + 1)
        want = u.copy()
        oracle.share_staged(want, *shape, nbr_mask=mask)
        for path in (fa.PATH_AUTO, fa.PATH_FUSED, fa.PATH_STAGED, fa.PATH_STAGED_DIRECT):
            got = run_gpu(torch_cuda, u, shape, boundary=fa.OPEN, rank=rank, lgrid=lgrid,
                          path=path)
            assert_same(got, want, shape)


@pytest.mark.parametrize("walk,rows,bpc,seg,planes", [(1, 1, 2, 4, 0), (2, 0, 2, 4, 0), (2, 2, -2, 2, 0),
                                                      (3, 2, 2, 3, 0), (0, 1, -3, 1, 0), (2, 0, 2, 4, 1),
                                                      (1, 1, -2, 3, 1)])
def test_rank_positions_open_streaming(torch_cuda, walk, rows, bpc, seg, planes):
    """Every rank position of a 3x3x3 open decomposition through the one-launch
    fused path (neighbours as self halos), for each traversal: bit-equal to
    the staged 26-message form."""
    shape = (7, 12, 4, 3)
    n, mx, my, mz = shape
    print("case rank positions streaming", walk, rows, bpc, seg, flush=True)
    for rank in range(27):
        mask = brute.mask_for_rank(rank, 3, 3, 3)
        u = oracle.new_field(*shape, fa.FILL_RAND, seed=100 + rank)
        want = u.copy()
        oracle.share_staged(want, *shape, nbr_mask=mask)
        d = to_dev(torch_cuda, u)
        f = fa.Faces(rank, 3, 3, 3, mx, my, mz, n, boundary=fa.OPEN)
        f.set_path(fa.PATH_FUSED)
        f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
        f.set_option(fa.OPT_FLOW, 0)  # the streaming kernel's walks
        f.set_option(fa.OPT_SEG, seg)
        f.set_option(fa.OPT_STREAM_WALK, walk)
        f.set_option(fa.OPT_STREAM_ROWS, rows)
        f.set_option(fa.OPT_STREAM_BPC, bpc)
        f.set_option(fa.OPT_STREAM_PLANES, planes)
        f.share(d)
        got = to_host(torch_cuda, d)
        f.close()
        assert_same(got, want, shape)


@pytest.mark.parametrize("flow,xcd", [(2, 2), (1, 1), (2, 0), (3, 2), (3, 1), (4, 2), (4, 1)])
@pytest.mark.parametrize("shape,seg", [((7, 12, 4, 3), 0), ((5, 10, 3, 2), 5), ((8, 4, 3, 5), 0),
                                       ((4, 16, 2, 3), 0), ((6, 5, 2, 2), 1), ((2, 4, 1, 3), 0),
                                       ((12, 2, 2, 2), 0)])
def test_open_rank_flow_bitexact(torch_cuda, flow, xcd, shape, seg):
    """The reference's own mode (open boundary, 1x1x1 rank grid) on the
    one-item kernel: each rank-surface node sums only the copies on its side;
    two shares equal the open sweep."""
    n, mx, my, mz = shape
    print("case open flow", flow, xcd, shape, seg, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=818)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.OPEN)
    f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
    f.set_option(fa.OPT_FLOW, flow)
    f.set_option(fa.OPT_XCD, xcd)
    f.set_option(fa.OPT_SEG, seg)
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, periodic=False, nshare=2), shape)


@pytest.mark.parametrize("flow,xcd,seg", [(2, 2, 0), (1, 1, 3), (2, 0, 1), (3, 2, 0), (3, 1, 2), (4, 2, 0), (4, 1, 2)])
def test_rank_positions_open_flow(torch_cuda, flow, xcd, seg):
    """Every rank position of a 3x3x3 open decomposition through the
    one-item kernel (neighbours as self halos): bit-equal to the staged
    26-message form; the staged path's interior sums (the same kernel with
    the rank surface skipped) plus its surface sums give the same bits."""
    shape = (7, 10, 4, 3)
    n, mx, my, mz = shape
    print("case rank positions flow", flow, xcd, seg, flush=True)
    for rank in range(27):
        mask = brute.mask_for_rank(rank, 3, 3, 3)
        u = oracle.new_field(*shape, fa.FILL_RAND, seed=200 + rank)
        want = u.copy()
        oracle.share_staged(want, *shape, nbr_mask=mask)
        for path in (fa.PATH_FUSED, fa.PATH_STAGED_DIRECT):
            d = to_dev(torch_cuda, u)
            f = fa.Faces(rank, 3, 3, 3, mx, my, mz, n, boundary=fa.OPEN)
            f.set_path(path)
            f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
            f.set_option(fa.OPT_FLOW, flow)
            f.set_option(fa.OPT_XCD, xcd)
            f.set_option(fa.OPT_SEG, seg)
            f.share(d)
            got = to_host(torch_cuda, d)
            f.close()
            assert_same(got, want, shape)


@pytest.mark.parametrize("path", [fa.PATH_STAGED, fa.PATH_STAGED_DIRECT])
@pytest.mark.parametrize("rank", [13, 0, 22])
@pytest.mark.parametrize("sblocks", [3, 1, 2, 0])
def test_staged_surface_blocks_bitexact(torch_cuda, path, rank, sblocks):
    """The staged step's surface sums as blocks of the interior launch (before
    or after its items, running beside them) or as a launch of their own: on
    a field large enough for the blocks to overlap, three steps in a row stay
    bit-equal to the oracle's staged form (random values)."""
    shape = (7, 40, 30, 20)
    n, mx, my, mz = shape
    print("case staged surface blocks", path, rank, sblocks, flush=True)
    mask = brute.mask_for_rank(rank, 3, 3, 3)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=300 + rank)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(rank, 3, 3, 3, mx, my, mz, n, boundary=fa.OPEN)
    f.set_path(path)
    f.set_option(fa.OPT_SURFACE_BLOCKS, sblocks)
    assert f.get_option(fa.OPT_SURFACE_BLOCKS) == sblocks
    for _ in range(3):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    want = u.copy()
    for _ in range(3):
        oracle.share_staged(want, *shape, nbr_mask=mask)
    assert_same(got, want, shape)


@pytest.mark.parametrize("flow", [1, 2, 3])
@pytest.mark.parametrize("sblocks", [3, 1, 2])
def test_staged_surface_blocks_all_ranks(torch_cuda, flow, sblocks):
    """Every rank position of a 3x3x3 grid, staged direct, the surface sums
    as blocks of the interior launch at each load width: bit-equal to the
    oracle."""
    shape = (5, 10, 4, 3)
    n, mx, my, mz = shape
    print("case staged surface blocks ranks", flow, sblocks, flush=True)
    for rank in range(27):
        mask = brute.mask_for_rank(rank, 3, 3, 3)
        u = oracle.new_field(*shape, fa.FILL_RAND, seed=400 + rank)
        d = to_dev(torch_cuda, u)
        f = fa.Faces(rank, 3, 3, 3, mx, my, mz, n, boundary=fa.OPEN)
        f.set_path(fa.PATH_STAGED_DIRECT)
        f.set_option(fa.OPT_FLOW, flow)
        f.set_option(fa.OPT_SURFACE_BLOCKS, sblocks)
        f.share(d)
        got = to_host(torch_cuda, d)
        f.close()
        want = u.copy()
        oracle.share_staged(want, *shape, nbr_mask=mask)
        assert_same(got, want, shape)


def test_neighbor_table(torch_cuda):
    f = fa.Faces(13, 3, 3, 3, 2, 2, 2, 4, boundary=fa.OPEN)  # centre rank
    assert all(f.neighbor(k) >= 0 for k in range(26))
    assert f.neighbor(0) == 12 and f.neighbor(1) == 14  # x-, x+
    assert f.neighbor(18) == 0 and f.neighbor(25) == 26  # corners
    f.close()
    g = fa.Faces(0, 3, 3, 3, 2, 2, 2, 4, boundary=fa.OPEN)
    assert g.neighbor(0) == -1 and g.neighbor(1) == 1 and g.neighbor(25) == 13
    g.close()


def test_repeated_shares(torch_cuda):
    shape = (5, 4, 4, 4)  # the CPU-runnable configuration, nk=nj=ni=1 .. 3
    u = oracle.new_field(*shape, fa.FILL_INT, seed=42)
    got = run_gpu(torch_cuda, u, shape, nshare=3)
    assert_same(got, run_oracle(u, shape, nshare=3), shape)


def test_c1_configuration(torch_cuda):
    """1x1x1 periodic, 4^3 elements, n=5, integer [0,1023], seed 42."""
    shape = (5, 4, 4, 4)
    u = oracle.new_field(*shape, fa.FILL_INT, seed=42)
    got = run_gpu(torch_cuda, u, shape)
    assert_same(got, run_oracle(u, shape), shape)


def test_c3_configuration(torch_cuda):
    """One rank of 6x6x6: 32^3 elements, n=8, random FP64 [-1,1), self halos."""
    shape = (8, 32, 32, 32)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=42)
    want = run_oracle(u, shape)
    got = run_gpu(torch_cuda, u, shape)
    assert_same(got, want, shape)
    got2 = run_gpu(torch_cuda, u, shape, rank=0, lgrid=(6, 6, 6))
    assert_same(got2, want, shape)
    # an interior rank of the open 6x6x6 decomposition with self halos: its
    # six face neighbours exist, so it sums the periodic partners (fused path)
    got3 = run_gpu(torch_cuda, u, shape, boundary=fa.OPEN, rank=43, lgrid=(6, 6, 6))
    assert_same(got3, want, shape)
    ref = u.copy()
    oracle.share_staged(ref, *shape, nbr_mask=63)
    assert_same(got3, ref, shape)


def test_c2_configuration(torch_cuda):
    """1x1x1 periodic, 60^3 elements, n=7, integer [0,1023], ni=10 shares."""
    shape = (7, 60, 60, 60)
    n, mx, my, mz = shape
    total = n ** 3 * mx * my * mz
    d = torch_cuda.empty(total, dtype=torch_cuda.float64, device="cuda")
    fa.fill(d, *shape, kind=fa.FILL_INT, seed=42)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    for _ in range(10):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    del d
    want = oracle.new_field(*shape, fa.FILL_INT, seed=42)
    for _ in range(10):
        oracle.share_sweep(want, *shape, periodic=True)
    assert_same(got, want, shape)


@pytest.mark.parametrize("shape,walk,bpc", [((8, 48, 48, 48), 2, -2), ((12, 30, 24, 30), 2, -2),
                                            ((5, 60, 60, 60), 1, -3), ((8, 32, 32, 32), 2, 2),
                                            ((9, 30, 46, 40), 1, -2)])
def test_default_launch_large_fields(torch_cuda, shape, walk, bpc):
    """The default traversal and block mode of fields of a few hundred MB
    (one unit per block with a residency cap for n >= 8 from 256 MiB and for
    n <= 5; rows in x where my has no divisor from 4 to 20), random values,
    two shares, bit-exact."""
    n, mx, my, mz = shape
    print("case default launch", shape, flush=True)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == fa.VARIANT_STREAM
    assert (f.get_option(fa.OPT_STREAM_WALK), f.get_option(fa.OPT_STREAM_BPC)) == (walk, bpc)
    d = torch_cuda.empty(fa.field_doubles(*shape), dtype=torch_cuda.float64, device="cuda")
    fa.fill(d, *shape, kind=fa.FILL_RAND, seed=77)
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    del d
    torch_cuda.cuda.empty_cache()
    want = oracle.new_field(*shape, fa.FILL_RAND, seed=77)
    for _ in range(2):
        oracle.share_sweep(want, *shape, periodic=True)
    assert_same(got, want, shape)


@pytest.mark.parametrize("count", [2, 3, 1001, 2 ** 20 + 7])
def test_stream_ceiling_leaves_values(torch_cuda, count):
    """The in-place streaming probe reports a bandwidth and leaves every bit
    of the buffer as it was (odd lengths included)."""
    print("case stream ceiling", count, flush=True)
    d = torch_cuda.empty(count, dtype=torch_cuda.float64, device="cuda")
    d.uniform_(-1.0, 1.0)
    d[0] = -0.0
    before = to_host(torch_cuda, d)
    gbs = fa.stream_ceiling(d, reps=3)
    assert gbs > 0
    assert np.array_equal(to_host(torch_cuda, d).view(np.uint64), before.view(np.uint64))
    with pytest.raises(fa.FacesError) as e:
        fa.stream_ceiling(int(d.data_ptr()) + 8, n=count - 1, reps=1)  # not 16-byte aligned
    assert e.value.status == fa.ERR_ARG
    with pytest.raises(fa.FacesError):
        fa.stream_ceiling(d, n=1)


def test_invalid_arguments(torch_cuda):
    with pytest.raises(fa.FacesError) as e:
        fa.Faces(0, 1, 1, 1, 2, 2, 2, 1)
    assert e.value.status == fa.ERR_ARG
    with pytest.raises(fa.FacesError):
        fa.Faces(5, 2, 2, 1, 2, 2, 2, 4)
    with pytest.raises(fa.FacesError):
        fa.Faces(0, 1, 1, 1, 2, 2, 2, 4, plane_stride=10)
    f = fa.Faces(0, 2, 1, 1, 2, 2, 2, 4, boundary=fa.OPEN)
    f.set_path(fa.PATH_FUSED)
    d = torch_cuda.zeros(fa.field_doubles(4, 2, 2, 2), dtype=torch_cuda.float64, device="cuda")
    with pytest.raises(fa.FacesError):
        f.share(d, compute=False)  # surface-only sums of a rank with neighbours: staged path only
    with pytest.raises(fa.FacesError):
        f.set_tuning(100, 0, 0)
    for bad in (-1, -5, 33):
        with pytest.raises(fa.FacesError):
            f.set_option(fa.OPT_STREAM_BPC, bad)
    for opt, bad in ((fa.OPT_FLOW, 5), (fa.OPT_FLOW, -1), (fa.OPT_XCD, 3), (fa.OPT_STREAM_WALK, 4),
                     (fa.OPT_SURFACE_BLOCKS, 4), (fa.OPT_SURFACE_BLOCKS, -1)):
        with pytest.raises(fa.FacesError):
            f.set_option(opt, bad)
    f.close()


@pytest.mark.parametrize("shape,kernel", [((7, 60, 2, 2), fa.VARIANT_STREAM),
                                          ((7, 9, 2, 2), fa.VARIANT_STREAM),
                                          ((7, 13, 3, 2), fa.VARIANT_STREAM),
                                          ((8, 7, 2, 3), fa.VARIANT_STREAM),
                                          ((4, 23, 2, 1), fa.VARIANT_STREAM),
                                          ((9, 7, 2, 2), fa.VARIANT_ITEMS),
                                          ((13, 2, 2, 2), fa.VARIANT_DIRECT)])
def test_auto_kernel_choice_bitexact(torch_cuda, shape, kernel):
    """AUTO takes the work-list kernel (ragged rows) when its items are more
    than twice as long as the one-item kernel's, else the one-item kernel,
    whose rows the item length does not divide are split between two item
    lengths (13, 7 and 23 columns above); n=9, 7 columns: one-element items
    against the work-list kernel's 3."""
    n, mx, my, mz = shape
    print("case auto kernel", shape, kernel, flush=True)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == kernel
    f.close()
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=404)
    assert_same(run_gpu(torch_cuda, u, shape, nshare=2), run_oracle(u, shape, nshare=2), shape)


@pytest.mark.parametrize("ps_pad", [0, 512])
def test_items_padded_plane_stride(torch_cuda, ps_pad):
    shape = (6, 7, 3, 3)
    n, mx, my, mz = shape
    ps = n ** 3 * mx * my + ps_pad if ps_pad else 0
    print("case items padded", ps, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=405, ps=ps)
    got = run_gpu(torch_cuda, u, shape, ps=ps, variant=(fa.VARIANT_ITEMS, 2))
    assert_same(got, run_oracle(u, shape, ps=ps), shape, ps=ps)


@pytest.mark.parametrize("shape,seg", [((4, 32, 2, 2), 16), ((2, 64, 2, 1), 32), ((3, 16, 3, 2), 16),
                                       ((6, 16, 2, 2), 8), ((5, 12, 3, 1), 6), ((6, 24, 1, 2), 6),
                                       ((5, 8, 2, 2), 8), ((4, 16, 1, 1), 8), ((5, 10, 3, 2), 5),
                                       ((5, 20, 2, 2), 10), ((6, 10, 2, 2), 5), ((7, 15, 2, 3), 5),
                                       ((8, 5, 3, 2), 5)])
@pytest.mark.parametrize("walk", [0, 1, 2, 3])
def test_stream_long_segments_bitexact(torch_cuda, shape, seg, walk):
    """Long items of the streaming kernel for small n, and the 5- and
    10-element items of n = 5-8."""
    n, mx, my, mz = shape
    print("case stream long segments", shape, seg, walk, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=406)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
    f.set_option(fa.OPT_FLOW, 0)
    f.set_option(fa.OPT_SEG, seg)
    f.set_option(fa.OPT_STREAM_WALK, walk)
    f.set_option(fa.OPT_STREAM_ROWS, 2)
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, nshare=2), shape)


# ------------------------------------------- the 6-message schedule (z, y, x)

SIX_SHAPES = [(2, 1, 1, 1), (4, 3, 2, 2), (5, 1, 3, 2), (7, 4, 3, 5), (8, 2, 2, 2), (12, 2, 1, 3),
              (13, 2, 2, 1)]


@pytest.mark.parametrize("shape", SIX_SHAPES)
@pytest.mark.parametrize("kind", [fa.FILL_INT, fa.FILL_RAND])
@pytest.mark.parametrize("compute", [True, False])
def test_six_periodic_bitexact(torch_cuda, shape, kind, compute):
    """PATH_SIX against the oracle's 6-message schedule (bit for bit, random
    fields too) and, on integer fields, against the 26-message result."""
    n, mx, my, mz = shape
    print("case six periodic", shape, kind, compute, flush=True)
    u = oracle.new_field(*shape, kind, seed=501)
    got = run_gpu(torch_cuda, u, shape, compute=compute, path=fa.PATH_SIX, nshare=2)
    want = u.copy()
    for _ in range(2):
        oracle.share_six(want, *shape, nbr_mask=63, compute=compute)
    assert_same(got, want, shape)
    if kind == fa.FILL_INT:
        assert_same(got, run_oracle(u, shape, compute=compute, nshare=2), shape)


@pytest.mark.parametrize("lgrid,rank", [((3, 3, 3), 13), ((3, 3, 3), 0), ((3, 3, 3), 4),
                                        ((2, 1, 3), 3), ((1, 1, 1), 0)])
@pytest.mark.parametrize("kind", [fa.FILL_INT, fa.FILL_RAND])
def test_six_open_ranks_bitexact(torch_cuda, lgrid, rank, kind):
    shape = (4, 3, 2, 2)
    n, mx, my, mz = shape
    print("case six open", lgrid, rank, kind, flush=True)
    u = oracle.new_field(*shape, kind, seed=502)
    got = run_gpu(torch_cuda, u, shape, boundary=fa.OPEN, path=fa.PATH_SIX, rank=rank, lgrid=lgrid)
    mask = brute.mask_for_rank(rank, *lgrid)
    want = u.copy()
    oracle.share_six(want, *shape, nbr_mask=mask)
    assert_same(got, want, shape)
    if kind == fa.FILL_INT:
        ref = u.copy()
        oracle.share_staged(ref, *shape, nbr_mask=mask)
        assert_same(got, ref, shape)


def _six_gpu(torch_cuda, u, shape, boundary, rank, lgrid, flow=None, nshare=1):
    n, mx, my, mz = shape
    d = to_dev(torch_cuda, u)
    f = fa.Faces(rank, *lgrid, mx, my, mz, n, boundary=boundary)
    f.set_path(fa.PATH_SIX)
    if flow is not None:
        f.set_option(fa.OPT_FLOW, flow)
    for _ in range(nshare):
        f.share(d)
    out = to_host(torch_cuda, d)
    f.close()
    return out


@pytest.mark.parametrize("shape", [(5, 5, 3, 4), (4, 6, 2, 3), (7, 10, 3, 2), (12, 2, 2, 1)])
def test_six_one_launch_all_ranks_bitexact(torch_cuda, shape):
    """With compute = 1 the 6-message schedule runs as one launch of the
    one-item kernel with its tree in z, y, x order: every rank position of a
    3x3x3 grid (each neighbour mask), random fields, bit for bit against the
    oracle's three phases and against the three phase kernels."""
    n, mx, my, mz = shape
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=503)
    for rank in range(27):
        print("case six one launch", shape, rank, flush=True)
        got = _six_gpu(torch_cuda, u, shape, fa.OPEN, rank, (3, 3, 3))
        want = u.copy()
        oracle.share_six(want, *shape, nbr_mask=brute.mask_for_rank(rank, 3, 3, 3))
        assert_same(got, want, shape)
        if rank in (0, 13, 26):
            phases = _six_gpu(torch_cuda, u, shape, fa.OPEN, rank, (3, 3, 3), flow=0)
            assert np.array_equal(got.view(np.uint64), phases.view(np.uint64))


@pytest.mark.parametrize("flow", [0, 1, 2, 3])
def test_six_periodic_flow_options_bitexact(torch_cuda, flow):
    """Periodic 6-message schedule at every FACES_OPT_FLOW setting (0: the
    three phase kernels; else one launch, tree z, y, x): same bits."""
    shape = (7, 10, 6, 4)
    print("case six periodic flow", flow, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=504)
    got = _six_gpu(torch_cuda, u, shape, fa.PERIODIC, 0, (1, 1, 1), flow=flow, nshare=2)
    want = u.copy()
    for _ in range(2):
        oracle.share_six(want, *shape, nbr_mask=63)
    assert_same(got, want, shape)


def test_six_c2_size_integer(torch_cuda):
    """60^3 n=7 integer field, 3 shares of PATH_SIX equal the oracle sweep."""
    shape = (7, 60, 60, 60)
    print("case six c2", flush=True)
    u = oracle.new_field(*shape, fa.FILL_INT, seed=42)
    got = run_gpu(torch_cuda, u, shape, path=fa.PATH_SIX, nshare=3)
    assert_same(got, run_oracle(u, shape, nshare=3), shape)


# --------------------------------- per-plane addressing (fields of 4 GB and more)

@pytest.mark.parametrize("walk,rows", [(0, 1), (1, 1), (2, 1), (2, 3), (3, 2)])
@pytest.mark.parametrize("shape,seg", [((7, 12, 3, 4), 4), ((4, 32, 2, 3), 16), ((12, 3, 2, 2), 1),
                                       ((5, 6, 1, 1), 3), ((8, 4, 3, 1), 2), ((7, 10, 3, 2), 5),
                                       ((5, 10, 2, 3), 10)])
def test_stream_per_plane_bitexact(torch_cuda, walk, rows, shape, seg):
    """The per-plane descriptors that fields above 4 GB need, forced on small
    fields (every walk, wrap planes included)."""
    n, mx, my, mz = shape
    print("case stream per plane", shape, seg, walk, rows, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=601)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
    f.set_option(fa.OPT_STREAM_PLANES, 1)
    assert f.get_option(fa.OPT_STREAM_PLANES) == 1
    f.set_option(fa.OPT_SEG, seg)
    f.set_option(fa.OPT_STREAM_WALK, walk)
    f.set_option(fa.OPT_STREAM_ROWS, rows)
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, nshare=2), shape)


@pytest.mark.slow
def test_field_above_4gb_bitexact(torch_cuda):
    """A 4.6 GB field (n=8, 100x100x112 elements) takes the streaming kernel
    with per-plane descriptors; one share equals the oracle."""
    shape = (8, 100, 100, 112)
    n, mx, my, mz = shape
    print("case field above 4 GB", shape, flush=True)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == fa.VARIANT_STREAM
    assert f.get_option(fa.OPT_STREAM_PLANES) == 1
    u = oracle.new_field(*shape, fa.FILL_INT, seed=602)
    assert 8 * u.size > 4 * 2 ** 30
    d = to_dev(torch_cuda, u)
    f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    del d
    torch_cuda.cuda.empty_cache()
    oracle.share_sweep(u, *shape)
    assert np.array_equal(got, u)


@pytest.mark.parametrize("shape", [(7, 12, 6, 4), (5, 8, 4, 3), (8, 7, 3, 2), (4, 16, 4, 2)])
@pytest.mark.parametrize("boundary", [fa.PERIODIC, fa.OPEN])
def test_plan_tune_then_bitexact(torch_cuda, shape, boundary):
    """The autotuner picks a valid launch shape; a share with it is exact
    (periodic, and open on one rank: the one-launch open mode)."""
    n, mx, my, mz = shape
    print("case plan tune", shape, boundary, flush=True)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=boundary)
    scratch = torch_cuda.empty(fa.field_doubles(*shape), dtype=torch_cuda.float64, device="cuda")
    fa.fill(scratch, *shape, kind=fa.FILL_RAND, seed=9)
    res = f.tune(scratch, reps=2)
    assert res["variant"] in (fa.VARIANT_STREAM, fa.VARIANT_ITEMS) and res["us_per_share"] > 0
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == res["variant"]
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=901)
    d = to_dev(torch_cuda, u)
    f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, periodic=boundary == fa.PERIODIC), shape)


# ------------------ rows the item length does not divide: two item launches

# (n, mx, my, mz) where the longest item dividing mx is 3 or more elements
# shorter than the default item: items of the default length cover the first
# mx - mx % S columns, a second launch the rest (n=7 mx=11: 2 items of 5 and
# one of 1 per row; n=8 mx=9: 2 of 4, one of 1; n=6 mx=11: 2 of 5, one of 1;
# n=7 mx=13: 2 of 5, one of 3; n=5 mx=23: 2 of 10, one of 3; n=8 mx=7: one
# of 4, one of 3; n=6 mx=7: one of 5, one of 2; n=4 mx=17: one of 16, one of
# 1; n=2 and n=3 mx=35: one of 32, one of 3; n=2 mx=131: two of 64, one of
# 3).
SPLIT_SHAPES = [(7, 11, 3, 2), (8, 9, 3, 2), (6, 11, 2, 2), (7, 13, 2, 3), (5, 23, 2, 2), (8, 7, 3, 2),
                (6, 7, 2, 2), (4, 17, 2, 2), (2, 35, 2, 1), (3, 35, 2, 2), (2, 131, 2, 2)]


@pytest.mark.parametrize("flow", [1, 2, 3, 4])
@pytest.mark.parametrize("xcd", [0, 1, 2])
@pytest.mark.parametrize("shape", SPLIT_SHAPES)
def test_flow_split_rows_bitexact(torch_cuda, flow, xcd, shape):
    """Periodic step over rows split between two item lengths, at each load
    width and XCD order: two shares of a random field, bit-exact."""
    n, mx, my, mz = shape
    print("case flow split rows", flow, xcd, shape, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=317)
    d = to_dev(torch_cuda, u)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    f.set_option(fa.OPT_VARIANT, fa.VARIANT_STREAM)
    f.set_option(fa.OPT_FLOW, flow)
    f.set_option(fa.OPT_XCD, xcd)
    assert f.get_option(fa.OPT_ACTIVE_KERNEL) == fa.VARIANT_STREAM
    for _ in range(2):
        f.share(d)
    got = to_host(torch_cuda, d)
    f.close()
    assert_same(got, run_oracle(u, shape, nshare=2), shape)


@pytest.mark.parametrize("path,sblocks", [(fa.PATH_FUSED, 2), (fa.PATH_STAGED_DIRECT, 2),
                                          (fa.PATH_STAGED_DIRECT, 1), (fa.PATH_STAGED_DIRECT, 3),
                                          (fa.PATH_STAGED_DIRECT, 0), (fa.PATH_STAGED, 2)])
@pytest.mark.parametrize("shape", [(7, 11, 3, 2), (6, 7, 2, 2), (5, 23, 2, 2), (3, 35, 3, 2), (2, 70, 2, 2),
                                   (8, 9, 3, 2)])
def test_split_rows_open_ranks(torch_cuda, path, sblocks, shape):
    """Open ranks of a 3x3x3 grid over split rows: the fused sweep, and the
    staged step with its surface sums as blocks of the first launch (after
    or before its items), by the items from the received messages, or as a
    launch of their own; bit-equal to the oracle's staged form."""
    n, mx, my, mz = shape
    print("case split rows open", path, sblocks, shape, flush=True)
    for rank in (13, 0, 22, 5, 26):
        mask = brute.mask_for_rank(rank, 3, 3, 3)
        u = oracle.new_field(*shape, fa.FILL_RAND, seed=600 + rank)
        d = to_dev(torch_cuda, u)
        f = fa.Faces(rank, 3, 3, 3, mx, my, mz, n, boundary=fa.OPEN)
        f.set_path(path)
        f.set_option(fa.OPT_SURFACE_BLOCKS, sblocks)
        for _ in range(2):
            f.share(d)
        got = to_host(torch_cuda, d)
        f.close()
        want = u.copy()
        for _ in range(2):
            oracle.share_staged(want, *shape, nbr_mask=mask)
        assert_same(got, want, shape)


@pytest.mark.parametrize("shape", [(7, 11, 3, 2), (8, 7, 3, 2), (2, 70, 2, 2), (8, 9, 3, 2)])
def test_split_rows_six(torch_cuda, shape):
    """The 6-message schedule in one launch per item length (tree z, y, x)
    over split rows: open ranks and the periodic step, bit-exact."""
    n, mx, my, mz = shape
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=505)
    for rank in (13, 0, 26):
        print("case split rows six", shape, rank, flush=True)
        got = _six_gpu(torch_cuda, u, shape, fa.OPEN, rank, (3, 3, 3))
        want = u.copy()
        oracle.share_six(want, *shape, nbr_mask=brute.mask_for_rank(rank, 3, 3, 3))
        assert_same(got, want, shape)
    got = _six_gpu(torch_cuda, u, shape, fa.PERIODIC, 0, (1, 1, 1), nshare=2)
    want = u.copy()
    for _ in range(2):
        oracle.share_six(want, *shape, nbr_mask=63)
    assert_same(got, want, shape)
