"""GPU parity of the batched step (faces_batch_*): many independent periodic
grids of different n, mx, my, mz summed by one persistent launch (replayed
from a HIP graph) must equal, bit for bit, the oracle applied to each grid on
its own and the one-launch-per-grid path."""
from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

import faces_amd as fa  # noqa: E402
from faces_amd.batch import Batch  # noqa: E402

def mixed_shapes(count, seed, ns=(4, 6, 8), lo=2, hi=24):
    """Seeded shapes: mx, my, mz uniform in [lo, hi], n drawn from ns."""
    out = []
    for g in range(count):
        h = [oracle.hash64(seed, (g << 3) + k) for k in range(4)]
        m = [lo + (h[k] % (hi - lo + 1)) for k in range(3)]
        out.append((ns[h[3] % len(ns)], *m))
    return out


def make_fields(torch, shapes, kind=fa.FILL_RAND, seed0=1000):
    fields = []
    for i, s in enumerate(shapes):
        d = torch.empty(fa.field_doubles(*s), dtype=torch.float64, device="cuda")
        fa.fill(d, *s, kind=kind, seed=seed0 + i)
        fields.append(d)
    torch.cuda.synchronize()
    return fields


def host(torch, fields):
    torch.cuda.synchronize()
    return [d.cpu().numpy() for d in fields]


def oracle_steps(shapes, kind, seed0, steps, periodic=True, compute=True):
    out = []
    for i, s in enumerate(shapes):
        u = oracle.new_field(*s, kind, seed=seed0 + i)
        for _ in range(steps):
            oracle.share_sweep(u, *s, periodic=periodic, compute=compute)
        out.append(u)
    return out


def assert_all_same(got, want, shapes):
    for i, (a, b, s) in enumerate(zip(got, want, shapes)):
        r = oracle.compare(a, b, *s)
        assert r["ndiff"] == 0, (i, s, r)


# (n, S) routines of the batched kernel; a grid takes the smallest S that
# holds its whole row, else the largest (and then a shorter last item)
ROUTINES = {2: [1, 2, 4, 8, 16, 32], 3: [1, 2, 4, 8, 16], 4: [1, 2, 4, 8, 16], 5: [1, 2, 4, 8],
            6: [1, 2, 4, 8], 7: [1, 2, 4, 6], 8: [1, 2, 4], 9: [1, 2, 3], 10: [1, 2],
            11: [1, 2], 12: [1]}


def edge_shapes():
    """Every routine with a full row, a short row, and (largest S) ragged rows."""
    out = []
    for n, opts in ROUTINES.items():
        prev = 0
        for k, S in enumerate(opts):
            out.append((n, S, 1 + k % 2, 1 + (k + 1) % 3))
            if S - prev > 1:
                out.append((n, prev + 1, 2, 1))
            prev = S
        out.append((n, 2 * opts[-1] + 1, 2, 2))
        out.append((n, 3 * opts[-1] - 1, 1, 1))
    return out


EDGE_SHAPES = edge_shapes()


@pytest.mark.parametrize("kind", [fa.FILL_INT, fa.FILL_RAND])
@pytest.mark.parametrize("bpc", [0, -1, 32])
def test_batch_edge_shapes_bitexact(torch_cuda, kind, bpc):
    """Every routine, ragged rows included, with the default launch, one
    block per item, and chunks of the list."""
    print("case batch edge shapes", kind, bpc, flush=True)
    fields = make_fields(torch_cuda, EDGE_SHAPES, kind=kind, seed0=50)
    b = Batch(EDGE_SHAPES)
    b.set_fields(fields)
    if bpc:
        b.set_tuning(256, 0, bpc)
    b.share()
    assert_all_same(host(torch_cuda, fields), oracle_steps(EDGE_SHAPES, kind, 50, 1), EDGE_SHAPES)
    b.close()


@pytest.mark.parametrize("seg,bpc", [(1, 1), (2, 2), (3, 8), (4, 1), (8, 6), (64, 32), (1, -1), (4, -1),
                                     (64, -1)])
def test_batch_tuning_bitexact(torch_cuda, seg, bpc):
    print("case batch tuning", seg, bpc, flush=True)
    shapes = mixed_shapes(40, seed=7, ns=(2, 3, 4, 5, 6, 7, 8), lo=1, hi=9)
    fields = make_fields(torch_cuda, shapes, seed0=60)
    b = Batch(shapes)
    b.set_fields(fields)
    b.set_tuning(256, seg, bpc)
    b.share()
    b.share()
    assert_all_same(host(torch_cuda, fields), oracle_steps(shapes, fa.FILL_RAND, 60, 2), shapes)
    b.close()


def test_batch_graph_replays(torch_cuda):
    """Four replays of the captured graph with unrelated eager work and host
    synchronisation between them; every replay checked."""
    print("case batch graph replays", flush=True)
    shapes = mixed_shapes(24, seed=11, lo=1, hi=8)
    fields = make_fields(torch_cuda, shapes, seed0=70)
    want = [oracle.new_field(*s, fa.FILL_RAND, seed=70 + i) for i, s in enumerate(shapes)]
    b = Batch(shapes)
    b.set_fields(fields)
    junk = torch_cuda.ones(1 << 20, dtype=torch_cuda.float64, device="cuda")
    for rep in range(4):
        b.share()
        junk.mul_(1.5).add_(float(rep))
        torch_cuda.cuda.synchronize()
        for i, s in enumerate(shapes):
            oracle.share_sweep(want[i], *s)
        assert_all_same(host(torch_cuda, fields), want, shapes)
    b.close()


def test_batch_matches_per_grid_256(torch_cuda):
    """256 mixed grids (mx, my, mz in [2,24], n in {4,6,8}): batched launch
    equals one launch per grid and the oracle on every one of them."""
    print("case batch 256 mixed", flush=True)
    shapes = mixed_shapes(256, seed=2024)
    a = make_fields(torch_cuda, shapes, seed0=5000)
    c = [x.clone() for x in a]
    b1, b2 = Batch(shapes), Batch(shapes)
    b1.set_fields(a)
    b2.set_fields(c)
    b1.share()
    b2.share_per_grid()
    ga, gc = host(torch_cuda, a), host(torch_cuda, c)
    for i in range(256):
        assert np.array_equal(ga[i].view(np.uint64), gc[i].view(np.uint64)), (i, shapes[i])
    assert_all_same(ga, [oracle_one(shapes[i], 5000 + i) for i in range(256)], shapes)
    b1.close()
    b2.close()


def oracle_one(shape, seed, steps=1):
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=seed)
    for _ in range(steps):
        oracle.share_sweep(u, *shape)
    return u


def test_batch_tuner(torch_cuda):
    print("case batch tuner", flush=True)
    shapes = mixed_shapes(32, seed=5, lo=2, hi=10)
    fields = make_fields(torch_cuda, shapes, seed0=80)
    keep = [d.clone() for d in fields]
    b = Batch(shapes)
    b.set_fields(fields)
    res = b.tune(reps=2)
    assert res["seg_len"] in (2, 4, 8, 16, 64)
    assert res["blocks_per_cu"] in (-1, 2, 4, 8, 16, 32, 48, 64, 96, 128)
    assert res["us_per_step"] > 0
    # the tuner stepped the fields; restore them and check one step
    for d, k in zip(fields, keep):
        d.copy_(k)
    b.share()
    assert_all_same(host(torch_cuda, fields), oracle_steps(shapes, fa.FILL_RAND, 80, 1), shapes)
    b.close()


@pytest.mark.parametrize("boundary,compute", [(fa.OPEN, True), (fa.PERIODIC, False),
                                              (fa.OPEN, False)])
def test_batch_other_modes(torch_cuda, boundary, compute):
    """Open boundary and surface-only steps go one grid at a time."""
    print("case batch other modes", boundary, compute, flush=True)
    shapes = [(4, 3, 2, 2), (6, 2, 3, 1), (5, 1, 1, 1), (13, 2, 1, 1)]
    fields = make_fields(torch_cuda, shapes, seed0=90)
    b = Batch(shapes, boundary=boundary)
    b.set_fields(fields)
    b.share(compute=compute)
    want = oracle_steps(shapes, fa.FILL_RAND, 90, 1, periodic=boundary == fa.PERIODIC,
                        compute=compute)
    assert_all_same(host(torch_cuda, fields), want, shapes)
    b.close()


def test_batch_with_solo_grids(torch_cuda):
    """Grids with n above the batched kernel's range step on their own plan
    after the batched launch, within the same share call."""
    print("case batch solo grids", flush=True)
    shapes = [(4, 3, 2, 2), (13, 2, 1, 2), (7, 2, 2, 2), (14, 1, 1, 1)]
    fields = make_fields(torch_cuda, shapes, seed0=110)
    b = Batch(shapes)
    b.set_fields(fields)
    b.share()
    b.share()
    assert_all_same(host(torch_cuda, fields), oracle_steps(shapes, fa.FILL_RAND, 110, 2), shapes)
    b.close()
    only = [(13, 1, 2, 1)]
    fields = make_fields(torch_cuda, only, seed0=120)
    b = Batch(only)
    b.set_fields(fields)
    b.share()
    assert_all_same(host(torch_cuda, fields), oracle_steps(only, fa.FILL_RAND, 120, 1), only)
    b.close()


@pytest.mark.parametrize("shapes", [
    [(4, 2, 1, 1)],                                     # one item: seven XCD shares empty
    [(2, 1, 1, 1)] * 5 + [(3, 1, 2, 1)],                # fewer items than XCDs
    [(8, 24, 24, 24)] + [(2, 1, 1, 1)] * 40,            # one grid holds most of the cost
    [(2, 1, 1, 1)] * 40 + [(12, 3, 2, 2), (5, 17, 1, 9)],
])
@pytest.mark.parametrize("bpc", [-1, 4])
def test_batch_uneven_xcd_shares(torch_cuda, shapes, bpc):
    """Work lists whose XCD shares of equal cost hold very different item
    counts (or none): every item is summed once."""
    print("case batch uneven shares", len(shapes), bpc, flush=True)
    fields = make_fields(torch_cuda, shapes, seed0=300)
    b = Batch(shapes)
    b.set_fields(fields)
    b.set_tuning(256, 0, bpc)
    b.share()
    b.share()
    assert_all_same(host(torch_cuda, fields), oracle_steps(shapes, fa.FILL_RAND, 300, 2), shapes)
    b.close()


def test_batch_rebind_fields(torch_cuda):
    """New field pointers re-upload the descriptors and re-capture the graph."""
    print("case batch rebind", flush=True)
    shapes = [(4, 3, 2, 2), (7, 2, 2, 2)]
    b = Batch(shapes)
    for seed0 in (100, 200):
        fields = make_fields(torch_cuda, shapes, seed0=seed0)
        b.set_fields(fields)
        b.share()
        assert_all_same(host(torch_cuda, fields), oracle_steps(shapes, fa.FILL_RAND, seed0, 1), shapes)
    b.close()


def test_batch_invalid(torch_cuda):
    with pytest.raises(fa.FacesError):
        Batch([(65, 2, 2, 2)])
    with pytest.raises(fa.FacesError):
        Batch([(4, 0, 2, 2)])
    b = Batch([(4, 2, 2, 2)])
    with pytest.raises(fa.FacesError) as e:
        b.share()  # fields not bound
    assert e.value.status == fa.ERR_STATE
    with pytest.raises(fa.FacesError):
        b.set_tuning(128, 0, 0)
    b.close()


def test_batch_large_grid_on_own_plan(torch_cuda):
    """A grid of 64 MiB or more steps on its own plan within the same share."""
    print("case batch large grid", flush=True)
    shapes = [(4, 3, 2, 2), (7, 30, 30, 30), (6, 2, 3, 4)]  # the middle one is 74 MB
    fields = make_fields(torch_cuda, shapes, seed0=130)
    b = Batch(shapes)
    b.set_fields(fields)
    b.share()
    b.share()
    assert_all_same(host(torch_cuda, fields), oracle_steps(shapes, fa.FILL_RAND, 130, 2), shapes)
    b.close()
