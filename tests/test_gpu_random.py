"""Seeded random combinations of shape, boundary, rank position, compute
flag, path, kernel variant and launch options, each checked bit for bit
against the oracle: the sweep for periodic grids, the staged 26-message
form for open ranks."""
from __future__ import annotations

import random

import numpy as np
import pytest

import oracle
from tests import _brute as brute

pytestmark = pytest.mark.gpu

import faces_amd as fa  # noqa: E402

NCASES = 240


def cases():
    rng = random.Random(2026)
    for case in range(NCASES):
        n = rng.randint(2, 12)
        shape = (n, rng.randint(1, 9), rng.randint(1, 7), rng.randint(1, 6))
        periodic = rng.random() < 0.5
        lgrid = (1, 1, 1) if periodic else rng.choice([(1, 1, 1), (2, 1, 1), (3, 3, 3), (1, 2, 3), (2, 2, 2)])
        rank = rng.randrange(lgrid[0] * lgrid[1] * lgrid[2])
        mask = 63 if periodic else brute.mask_for_rank(rank, *lgrid)
        compute = rng.random() < 0.85
        path = rng.choice([fa.PATH_AUTO, fa.PATH_AUTO, fa.PATH_FUSED, fa.PATH_STAGED, fa.PATH_STAGED_DIRECT])
        if path == fa.PATH_FUSED and mask not in (0, 63):
            path = fa.PATH_AUTO  # explicit fused on such a rank needs the streaming kernel (tested apart)
        opts = {
            fa.OPT_VARIANT: rng.choice([0, 0, fa.VARIANT_STREAM, fa.VARIANT_ITEMS, fa.VARIANT_DIRECT,
                                        fa.VARIANT_LDS]),
            fa.OPT_STREAM_WALK: rng.randint(0, 3),
            fa.OPT_STREAM_ROWS: rng.randint(0, 6),
            fa.OPT_STREAM_BPC: rng.choice([1, 2, 3, -2, -3, 0]),
            fa.OPT_SEG: rng.choice([0, 1, 2, 3, 4, 5, 6, 8]),
            fa.OPT_STREAM_PLANES: int(rng.random() < 0.2),
        }
        kind = rng.choice([fa.FILL_RAND, fa.FILL_INT])
        yield case, shape, periodic, lgrid, rank, mask, compute, path, opts, kind


def test_random_combinations_bitexact(torch_cuda):
    for case, shape, periodic, lgrid, rank, mask, compute, path, opts, kind in cases():
        n, mx, my, mz = shape
        print("case random", case, shape, periodic, lgrid, rank, compute, path, opts, flush=True)
        u = oracle.new_field(*shape, kind, seed=case)
        want = u.copy()
        if periodic:
            oracle.share_sweep(want, *shape, periodic=True, compute=compute)
        else:
            oracle.share_staged(want, *shape, nbr_mask=mask, compute=compute)
        f = fa.Faces(rank, *lgrid, mx, my, mz, n, boundary=fa.PERIODIC if periodic else fa.OPEN)
        f.set_path(path)
        for k, v in opts.items():
            f.set_option(k, v)
        d = torch_cuda.from_numpy(u.copy()).cuda()
        f.share(d, compute=compute)
        got = d.cpu().numpy()
        f.close()
        r = oracle.compare(got, want, *shape)
        assert r["ndiff"] == 0, (case, shape, periodic, lgrid, rank, compute, path, opts, r)


def flow_cases():
    """The one-item kernel's options: load width, XCD order, where the staged
    step's surface sums run, and the 6-message schedule."""
    rng = random.Random(4051)
    for case in range(NCASES):
        n = rng.randint(2, 12)
        shape = (n, rng.randint(1, 12), rng.randint(1, 7), rng.randint(1, 6))
        periodic = rng.random() < 0.4
        lgrid = (1, 1, 1) if periodic else rng.choice([(1, 1, 1), (2, 1, 1), (3, 3, 3), (1, 2, 3), (2, 2, 2)])
        rank = rng.randrange(lgrid[0] * lgrid[1] * lgrid[2])
        mask = 63 if periodic else brute.mask_for_rank(rank, *lgrid)
        compute = rng.random() < 0.9
        path = rng.choice([fa.PATH_AUTO, fa.PATH_STAGED, fa.PATH_STAGED_DIRECT, fa.PATH_SIX])
        opts = {
            fa.OPT_FLOW: rng.randint(0, 3),
            fa.OPT_XCD: rng.randint(0, 2),
            fa.OPT_SURFACE_BLOCKS: rng.randint(0, 3),
            fa.OPT_SEG: rng.choice([0, 0, 1, 2, 3, 4, 5, 6]),
        }
        kind = rng.choice([fa.FILL_RAND, fa.FILL_INT])
        yield case, shape, periodic, lgrid, rank, mask, compute, path, opts, kind


def test_random_flow_combinations_bitexact(torch_cuda):
    for case, shape, periodic, lgrid, rank, mask, compute, path, opts, kind in flow_cases():
        n, mx, my, mz = shape
        print("case random flow", case, shape, periodic, lgrid, rank, compute, path, opts, flush=True)
        u = oracle.new_field(*shape, kind, seed=10_000 + case)
        want = u.copy()
        if path == fa.PATH_SIX:
            oracle.share_six(want, *shape, nbr_mask=mask, compute=compute)
        elif periodic:
            oracle.share_sweep(want, *shape, periodic=True, compute=compute)
        else:
            oracle.share_staged(want, *shape, nbr_mask=mask, compute=compute)
        f = fa.Faces(rank, *lgrid, mx, my, mz, n, boundary=fa.PERIODIC if periodic else fa.OPEN)
        f.set_path(path)
        for k, v in opts.items():
            f.set_option(k, v)
        d = torch_cuda.from_numpy(u.copy()).cuda()
        f.share(d, compute=compute)
        got = d.cpu().numpy()
        f.close()
        r = oracle.compare(got, want, *shape)
        assert r["ndiff"] == 0, (case, shape, periodic, lgrid, rank, compute, path, opts, r)
