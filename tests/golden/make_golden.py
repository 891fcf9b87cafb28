"""Writes the golden fixtures in this folder (golden.npz).

The reference ships no vectors for the exchange + DSS step, so the fixtures
are produced here, independently of the oracle: inputs by a numpy
restatement of the generators (splitmix64 of seed ^ index*golden-ratio; the
positional formula of gpu/main.cc:18), expected outputs by the brute-force
node-id model of tests/_brute.py (every copy gets the id of the mesh node it
sits on, and the result is the sum over all copies of that id).  Inputs are
integer-valued wherever a share is applied, so the sums are exact in any
order and the expected outputs are exact.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests import _brute as brute  # noqa: E402

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix(seed: int, idx: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = np.uint64(seed) ^ (idx.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def field_int(n, mx, my, mz, seed):
    """Integer-valued [0,1023]: the top 10 bits of the hash of the dense index."""
    h = splitmix(seed, np.arange(n ** 3 * mx * my * mz))
    return (h >> np.uint64(54)).astype(np.float64)


def field_ref(n, mx, my, mz):
    """gpu/main.cc:18: jz*10000+jy*100+jx+iz*0.01+iy*0.0001+ix*0.000001, each
    product and sum rounded in turn, left to right (no fused multiply-add)."""
    jz, jy, jx, iz, iy, ix = [a.astype(np.float64) for a in np.indices((mz, my, mx, n, n, n))]
    v = jz * 10000.0
    v = v + jy * 100.0
    v = v + jx
    v = v + iz * 0.01
    v = v + iy * 0.0001
    v = v + ix * 0.000001
    return v.ravel()


def steps(u, shape, k, periodic=True, compute=True, mask=None):
    n, mx, my, mz = shape
    v6 = u.reshape(mz, my, mx, n, n, n)
    for _ in range(k):
        if mask is None:
            v6 = brute.share(v6, n, mx, my, mz, periodic=periodic, compute=compute)
        else:
            v6 = brute.share_selfhalo(v6, n, mx, my, mz, mask, compute=compute)
    return np.ascontiguousarray(v6).ravel()


def cases():
    out = {}
    # C1: 1x1x1 periodic, 4x4x4 elements, n=5, integer-valued, seed 42
    s = (5, 4, 4, 4)
    u = field_int(*s, 42)
    out["c1_in"] = u
    out["c1_periodic_1"] = steps(u, s, 1)
    out["c1_periodic_10"] = steps(u, s, 10)
    out["c1_periodic_nocompute_2"] = steps(u, s, 2, compute=False)
    out["c1_open_1"] = steps(u, s, 1, periodic=False)
    out["c1_open_nocompute_1"] = steps(u, s, 1, periodic=False, compute=False)
    # ragged shape, several n
    for s, seed in [((2, 3, 1, 2), 1), ((3, 2, 3, 1), 2), ((8, 3, 2, 2), 3), ((6, 5, 1, 3), 4)]:
        key = "m%d_%dx%dx%d" % s
        u = field_int(*s, seed)
        out[key + "_in"] = u
        out[key + "_periodic_2"] = steps(u, s, 2)
        out[key + "_open_1"] = steps(u, s, 1, periodic=False)
    # one rank of an open 3x3x3 grid with self halos: interior (13), corner (0),
    # edge (1), face (4) ranks
    s = (4, 2, 3, 2)
    u = field_int(*s, 9)
    out["rank_in"] = u
    for r in (0, 1, 4, 13):
        out["rank%d_open_1" % r] = steps(u, s, 1, mask=brute.mask_for_rank(r, 3, 3, 3))
    # the reference's positional init (values only; not summed)
    out["ref_init_5_3x2x2"] = field_ref(5, 3, 2, 2)
    # the hash itself on a few indices
    out["splitmix_seed42"] = splitmix(42, np.arange(16))
    return out


def main():
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **cases())
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
