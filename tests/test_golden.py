"""The golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from a numpy restatement of the generators and the
brute-force node-id model) reproduced bit for bit by the oracle (CPU) and by
the library (GPU, through the C ABI)."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from tests import _brute as brute

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz"),
               allow_pickle=False)

SHAPES = {"c1": (5, 4, 4, 4), "m2_3x1x2": (2, 3, 1, 2), "m3_2x3x1": (3, 2, 3, 1),
          "m8_3x2x2": (8, 3, 2, 2), "m6_5x1x3": (6, 5, 1, 3), "rank": (4, 2, 3, 2)}
SEEDS = {"c1": 42, "m2_3x1x2": 1, "m3_2x3x1": 2, "m8_3x2x2": 3, "m6_5x1x3": 4, "rank": 9}


def expected_cases():
    """(fixture key, input key, shape, shares, periodic, compute, nbr_mask or None)."""
    out = [("c1_periodic_1", "c1", 1, True, True, None),
           ("c1_periodic_10", "c1", 10, True, True, None),
           ("c1_periodic_nocompute_2", "c1", 2, True, False, None),
           ("c1_open_1", "c1", 1, False, True, None),
           ("c1_open_nocompute_1", "c1", 1, False, False, None)]
    for k in ("m2_3x1x2", "m3_2x3x1", "m8_3x2x2", "m6_5x1x3"):
        out += [(k + "_periodic_2", k, 2, True, True, None), (k + "_open_1", k, 1, False, True, None)]
    for r in (0, 1, 4, 13):
        out.append(("rank%d_open_1" % r, "rank", 1, False, True, brute.mask_for_rank(r, 3, 3, 3)))
    return out


CASES = expected_cases()


def test_fixture_inputs_match_the_generators():
    for k, shape in SHAPES.items():
        u = oracle.new_field(*shape, oracle.FILL_INT, seed=SEEDS[k])
        assert np.array_equal(u, GOLD[k + "_in"]), k
    assert np.array_equal(oracle.new_field(5, 3, 2, 2, oracle.FILL_REF), GOLD["ref_init_5_3x2x2"])
    assert [oracle.hash64(42, i) for i in range(16)] == [int(x) for x in GOLD["splitmix_seed42"]]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_golden(case):
    key, src, k, periodic, compute, mask = case
    shape = SHAPES[src]
    u = GOLD[src + "_in"].copy()
    for _ in range(k):
        if mask is None:
            oracle.share_sweep(u, *shape, periodic=periodic, compute=compute)
        else:
            oracle.share_staged(u, *shape, nbr_mask=mask, compute=compute)
    assert np.array_equal(u, GOLD[key])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_matches_golden(torch_cuda, case):
    import faces_amd as fa
    key, src, k, periodic, compute, mask = case
    print("case golden", key, flush=True)
    n, mx, my, mz = SHAPES[src]
    if mask is None:
        rank, lgrid = 0, (1, 1, 1)
    else:
        rank, lgrid = int(key[4:key.index("_")]), (3, 3, 3)
    d = torch_cuda.from_numpy(GOLD[src + "_in"].copy()).to("cuda")
    f = fa.Faces(rank, *lgrid, mx, my, mz, n, boundary=fa.PERIODIC if periodic else fa.OPEN)
    for _ in range(k):
        f.share(d, compute=compute)
    torch_cuda.cuda.synchronize()
    got = d.cpu().numpy()
    f.close()
    assert np.array_equal(got, GOLD[key])
