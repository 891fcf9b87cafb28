"""GPU test of faces_share_noalloc, the staged step that allocates nothing
(its messages live in the library's per-device scratch array): through the C
ABI it gives the same bits as faces_share on a plan with the staged path
(FACES_PATH_STAGED_DIRECT), call after call, the later calls on smaller
fields than the earlier ones (the scratch keeps the last call's contents)."""
from __future__ import annotations

import numpy as np
import pytest

import faces_amd as fa

pytestmark = pytest.mark.gpu

# (rank, lx, ly, lz, mx, my, mz, n, boundary, compute), largest first
SEQUENCE = [
    (0, 1, 1, 1, 60, 60, 60, 7, fa.PERIODIC, True),  # 592.7 MB field, 8.5 MB of messages
    (13, 3, 3, 3, 12, 10, 8, 7, fa.OPEN, True),      # interior rank: all 26 neighbours
    (0, 3, 3, 3, 6, 5, 4, 5, fa.OPEN, True),         # corner rank: 7 neighbours
    (22, 3, 3, 3, 4, 3, 5, 6, fa.OPEN, False),       # face rank, rank surface only
    (0, 1, 1, 1, 3, 2, 2, 4, fa.PERIODIC, True),
]


def staged_reference(torch, case, seed):
    rank, lx, ly, lz, mx, my, mz, n, boundary, compute = case
    u = torch.empty(fa.field_doubles(n, mx, my, mz), dtype=torch.float64, device="cuda")
    fa.fill(u, n, mx, my, mz, kind=fa.FILL_RAND, seed=seed)
    f = fa.Faces(rank, lx, ly, lz, mx, my, mz, n, boundary=boundary)
    f.set_path(fa.PATH_STAGED_DIRECT)
    f.share(u, compute=compute)
    torch.cuda.synchronize()
    f.close()
    return u.cpu().numpy()


def noalloc(torch, case, seed):
    rank, lx, ly, lz, mx, my, mz, n, boundary, compute = case
    u = torch.empty(fa.field_doubles(n, mx, my, mz), dtype=torch.float64, device="cuda")
    fa.fill(u, n, mx, my, mz, kind=fa.FILL_RAND, seed=seed)
    torch.cuda.synchronize()
    fa.share_noalloc(u, rank, lx, ly, lz, mx, my, mz, n, boundary=boundary, compute=compute)
    torch.cuda.synchronize()
    return u.cpu().numpy()


def test_noalloc_equals_staged_share(torch_cuda):
    for i, case in enumerate(SEQUENCE):
        print("case noalloc", case, flush=True)
        want = staged_reference(torch_cuda, case, 300 + i)
        got = noalloc(torch_cuda, case, 300 + i)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), case


def test_noalloc_repeated_and_smaller(torch_cuda):
    """Two calls in a row, the second on a smaller field, then the first
    shape again: each equal to the plan-based staged share."""
    for case in (SEQUENCE[1], SEQUENCE[2], SEQUENCE[1]):
        print("case noalloc again", case, flush=True)
        want = staged_reference(torch_cuda, case, 77)
        got = noalloc(torch_cuda, case, 77)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), case


def test_noalloc_rejects_messages_beyond_scratch(torch_cuda):
    """95x94x93 n=12 needs 61 MB of messages, above the 16 MiB scratch: the
    call fails on the host, before any launch (the field is never touched)."""
    u = torch_cuda.empty(16, dtype=torch_cuda.float64, device="cuda")
    st = fa.lib().faces_share_noalloc(u.data_ptr(), 0, 1, 1, 1, 95, 94, 93, 12, fa.PERIODIC, 0, 1, None)
    assert st == fa.ERR_ARG
    assert "faces_share_noalloc" in fa.lib().faces_last_error().decode()
