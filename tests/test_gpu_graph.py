"""faces_share inside a captured HIP graph: the step allocates nothing and
records no memset, so a caller may capture it.  Each case replays the graph
several times with unrelated eager work and host synchronisation between the
replays, and checks every replay against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests import _brute as brute

pytestmark = pytest.mark.gpu

import faces_amd as fa  # noqa: E402


CASES = [
    # (shape, boundary, rank, lgrid, path, variant, compute)
    ((7, 12, 5, 3), fa.PERIODIC, 0, (1, 1, 1), fa.PATH_AUTO, 0, True),        # streaming
    ((7, 13, 3, 2), fa.PERIODIC, 0, (1, 1, 1), fa.PATH_AUTO, 0, True),        # work-list items
    ((5, 4, 3, 2), fa.PERIODIC, 0, (1, 1, 1), fa.PATH_AUTO, 0, False),        # surface only
    ((4, 3, 2, 2), fa.OPEN, 0, (3, 3, 3), fa.PATH_AUTO, 0, True),             # open streaming, corner rank
    ((4, 3, 2, 2), fa.OPEN, 0, (3, 3, 3), fa.PATH_STAGED, 0, True),           # staged, corner rank
    ((7, 3, 2, 2), fa.OPEN, 4, (3, 3, 3), fa.PATH_STAGED_DIRECT, 0, True),    # staged direct, face rank
    ((4, 3, 2, 2), fa.OPEN, 0, (1, 1, 1), fa.PATH_AUTO, 0, True),             # open 1x1x1
    ((6, 3, 2, 2), fa.PERIODIC, 0, (1, 1, 1), fa.PATH_SIX, 0, True),          # 6-message
    ((13, 2, 2, 2), fa.PERIODIC, 0, (1, 1, 1), fa.PATH_AUTO, 0, True),        # direct (n > 12)
]


def oracle_step(u, shape, boundary, rank, lgrid, path, compute):
    if path == fa.PATH_SIX:
        mask = 63 if boundary == fa.PERIODIC else brute.mask_for_rank(rank, *lgrid)
        oracle.share_six(u, *shape, nbr_mask=mask, compute=compute)
    elif boundary == fa.PERIODIC:
        oracle.share_sweep(u, *shape, periodic=True, compute=compute)
    else:
        oracle.share_staged(u, *shape, nbr_mask=brute.mask_for_rank(rank, *lgrid), compute=compute)


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_share_in_graph_replays(torch_cuda, case):
    shape, boundary, rank, lgrid, path, variant, compute = case
    n, mx, my, mz = shape
    torch = torch_cuda
    print("case graph", case, flush=True)
    u = oracle.new_field(*shape, fa.FILL_RAND, seed=701)
    d = torch.from_numpy(u.copy()).to("cuda")
    f = fa.Faces(rank, *lgrid, mx, my, mz, n, boundary=boundary)
    f.set_path(path)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        f.share(d, compute=compute, stream=s)
    torch.cuda.synchronize()
    # capture does not run the work: the field is unchanged
    assert np.array_equal(d.cpu().numpy(), u)
    want = u.copy()
    junk = torch.ones(1 << 18, dtype=torch.float64, device="cuda")
    for rep in range(3):
        g.replay()
        junk.mul_(1.25).add_(float(rep))
        torch.cuda.synchronize()
        oracle_step(want, shape, boundary, rank, lgrid, path, compute)
        r = oracle.compare(d.cpu().numpy(), want, *shape)
        assert r["ndiff"] == 0, (rep, r)
    f.close()
