"""Pins the CPU oracle (oracle/faces_oracle.c) without running the reference.

The reference repository holds no golden vectors, and its own check
(gpu/Mugs.cc) needs MPI, which this image lacks.  What fixes the oracle:
  * a brute-force node-id model (tests/_brute.py) -- connectivity, bit-exact on
    integer-valued fields, rounding-bounded on random ones;
  * closed forms: constant fields, multiplicities, conservation, delta fields;
  * symmetries: periodic shift equivariance;
  * two oracle formulations that must agree bit for bit on any data (the pass
    sweep of gpu/Mugs.cc:136-203 and the 26-message staged form);
  * the pack index maps against numpy slices of the 6-D field.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from tests import _brute as brute

SHAPES = [
    (2, 1, 1, 1), (2, 2, 3, 1), (3, 1, 2, 3), (4, 3, 2, 2), (5, 4, 4, 4),
    (5, 1, 1, 1), (6, 2, 1, 3), (7, 3, 3, 2), (8, 2, 2, 2), (3, 5, 1, 1),
]


def _run_sweep(u, shape, periodic, compute=True, ps=0):
    v = u.copy()
    oracle.share_sweep(v, *shape, periodic=periodic, compute=compute, ps=ps)
    return v


# ---------------------------------------------------------------- generators

def test_fill_int_range_and_determinism():
    u = oracle.new_field(5, 4, 4, 4, oracle.FILL_INT, seed=42)
    assert u.min() >= 0 and u.max() <= 1023
    assert np.all(u == np.floor(u))
    assert len(np.unique(u)) > 900
    assert np.array_equal(u, oracle.new_field(5, 4, 4, 4, oracle.FILL_INT, seed=42))
    assert not np.array_equal(u, oracle.new_field(5, 4, 4, 4, oracle.FILL_INT, seed=43))


def test_fill_rand_range():
    u = oracle.new_field(6, 3, 3, 3, oracle.FILL_RAND, seed=7)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.05


def test_splitmix_known_values():
    # splitmix64 (Steele, Lea & Flood 2014) from state 0: the generator's first
    # outputs are the published test values of the reference implementation.
    assert oracle.hash64(0, 0) == 0xE220A8397B1DCDAF
    # seed ^ idx*golden then one splitmix step == the idx-th output from seed 0
    assert oracle.hash64(0, 1) == 0x6E789E6AA1B965F4
    assert oracle.hash64(0, 2) == 0x06C45D188009454F


def test_fill_reference_formula():
    # gpu/main.cc:21-37: jz*10000+jy*100+jx+iz*0.01+iy*0.0001+ix*0.000001
    n, mx, my, mz = 3, 2, 3, 2
    u = oracle.new_field(n, mx, my, mz, oracle.FILL_REF)
    v6 = brute.as6(u, n, mx, my, mz)
    for jz, jy, jx, iz, iy, ix in [(0, 0, 0, 0, 0, 0), (1, 2, 1, 2, 1, 0), (1, 0, 1, 0, 2, 2)]:
        want = jz * 10000 + jy * 100 + jx + iz * 0.01 + iy * 0.0001 + ix * 0.000001
        assert v6[jz, jy, jx, iz, iy, ix] == want


# ------------------------------------------------------ brute-force agreement

@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("compute", [True, False])
def test_sweep_matches_bruteforce_integer(shape, periodic, compute):
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=11)
    got = brute.as6(_run_sweep(u, shape, periodic, compute), *shape)
    want = brute.share(brute.as6(u, *shape), *shape, periodic=periodic, compute=compute)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("periodic", [True, False])
def test_sweep_random_within_rounding(shape, periodic):
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=5)
    got = brute.as6(_run_sweep(u, shape, periodic), *shape)
    v6 = brute.as6(u, *shape)
    ids, _ = brute.node_ids(*shape, periodic=periodic)
    # exact per-node sums (math.fsum) and the classic bound for a sum of at
    # most 8 terms: |s_hat - s| <= 7u/(1-7u) * sum|c|, u = 2^-53
    groups = {}
    for i, x in zip(ids.ravel(), v6.ravel()):
        groups.setdefault(int(i), []).append(float(x))
    exact = {k: math.fsum(v) for k, v in groups.items()}
    mag = {k: math.fsum(abs(t) for t in v) for k, v in groups.items()}
    u_r = 2.0 ** -53
    for i, x in zip(ids.ravel(), got.ravel()):
        bound = 7 * u_r / (1 - 7 * u_r) * mag[int(i)]
        assert abs(x - exact[int(i)]) <= bound


def test_sweep_all_copies_equal_after_share():
    shape = (6, 3, 2, 4)
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=9)
    got = brute.as6(_run_sweep(u, shape, True), *shape)
    ids, _ = brute.node_ids(*shape, periodic=True)
    flat_ids, flat = ids.ravel(), got.ravel()
    order = np.argsort(flat_ids, kind="stable")
    s_ids, s_val = flat_ids[order], flat[order]
    starts = np.r_[0, np.flatnonzero(np.diff(s_ids)) + 1]
    first = np.repeat(s_val[starts], np.diff(np.r_[starts, len(s_ids)]))
    assert np.array_equal(s_val, first)  # bit-identical copies


# ------------------------------------------------------------ closed forms

@pytest.mark.parametrize("shape", [(5, 4, 4, 4), (7, 3, 2, 5), (2, 2, 2, 2)])
def test_constant_field_gives_multiplicity(shape):
    n, mx, my, mz = shape
    u = np.full(oracle.field_doubles(*shape), 3.0)
    got = brute.as6(_run_sweep(u, shape, True), *shape)
    # periodic: every node of an axis with i in {0, n-1} is shared by 2 copies
    ix = np.arange(n)
    sx = np.where((ix == 0) | (ix == n - 1), 2, 1)
    mult = sx[None, None, None, :, None, None] * sx[None, None, None, None, :, None] * \
        sx[None, None, None, None, None, :]
    assert np.array_equal(got, np.broadcast_to(3.0 * mult, got.shape))


def test_constant_field_open_multiplicity():
    shape = (4, 3, 3, 3)
    u = np.ones(oracle.field_doubles(*shape))
    got = brute.as6(_run_sweep(u, shape, False), *shape)
    assert np.array_equal(got, brute.multiplicity(*shape, periodic=False).astype(float))
    # corners of the rank hold one copy, the middle of the cube eight
    assert got[0, 0, 0, 0, 0, 0] == 1.0
    assert got[1, 1, 1, 0, 0, 0] == 8.0


@pytest.mark.parametrize("periodic", [True, False])
def test_conservation_integer(periodic):
    shape = (5, 4, 3, 2)
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=3)
    got = _run_sweep(u, shape, periodic)
    mult = brute.multiplicity(*shape, periodic=periodic).ravel()
    assert math.fsum(brute.as6(got, *shape).ravel() / mult) == math.fsum(u)


def test_delta_field_reaches_exactly_its_copies():
    shape = (4, 3, 3, 3)
    n, mx, my, mz = shape
    rng = np.random.default_rng(0)
    ids, _ = brute.node_ids(*shape, periodic=True)
    for _ in range(20):
        u = np.zeros(oracle.field_doubles(*shape))
        k = int(rng.integers(u.size))
        u[k] = 1.0
        got = _run_sweep(u, shape, True)
        assert np.array_equal(got != 0, (ids.ravel() == ids.ravel()[k]))
        assert np.all(got[got != 0] == 1.0)


def test_periodic_shift_equivariance():
    shape = (5, 4, 3, 2)
    n, mx, my, mz = shape
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=21)
    a = brute.as6(_run_sweep(u, shape, True), *shape)
    for axis in (0, 1, 2):  # jz, jy, jx
        us = np.ascontiguousarray(np.roll(brute.as6(u, *shape), 1, axis=axis)).ravel()
        b = brute.as6(_run_sweep(us, shape, True), *shape)
        assert np.array_equal(np.roll(a, 1, axis=axis), b)


# ------------------------------------------------- formulations that must agree

@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("compute", [True, False])
def test_staged_equals_sweep_bitwise_random(shape, compute):
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=17)
    a = _run_sweep(u, shape, True, compute)
    b = u.copy()
    oracle.share_staged(b, *shape, nbr_mask=63, compute=compute)
    assert np.array_equal(a, b)
    a = _run_sweep(u, shape, False, compute)
    b = u.copy()
    oracle.share_staged(b, *shape, nbr_mask=0, compute=compute)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", SHAPES)
def test_mugs26_equals_sweep_on_integers(shape):
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=2)
    a = _run_sweep(u, shape, True)
    b = u.copy()
    oracle.share_mugs26(b, *shape, nbr_mask=63)
    assert np.array_equal(a, b)


def test_mugs26_random_within_rounding():
    shape = (5, 3, 4, 2)
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=4)
    a = _run_sweep(u, shape, True)
    b = u.copy()
    oracle.share_mugs26(b, *shape, nbr_mask=63)
    mag = brute.share(np.abs(brute.as6(u, *shape)), *shape, periodic=True)
    u_r = 2.0 ** -53
    assert np.all(np.abs(brute.as6(a, *shape) - brute.as6(b, *shape)) <= 14 * u_r * mag)


@pytest.mark.parametrize("lgrid", [(3, 3, 3), (2, 1, 3), (1, 2, 1)])
def test_staged_rank_positions_match_bruteforce(lgrid):
    lx, ly, lz = lgrid
    shape = (4, 2, 3, 2)
    for rank in range(lx * ly * lz):
        mask = brute.mask_for_rank(rank, lx, ly, lz)
        u = oracle.new_field(*shape, oracle.FILL_INT, seed=rank)
        got = u.copy()
        oracle.share_staged(got, *shape, nbr_mask=mask)
        want = brute.share_selfhalo(brute.as6(u, *shape), *shape, mask)
        assert np.array_equal(brute.as6(got, *shape), want), (rank, mask)
        got2 = u.copy()
        oracle.share_mugs26(got2, *shape, nbr_mask=mask)
        assert np.array_equal(got2, got), (rank, mask)


def test_selfhalo_full_mask_is_periodic():
    shape = (5, 3, 2, 3)
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=8)
    want = brute.share(brute.as6(u, *shape), *shape, periodic=True)
    assert np.array_equal(brute.share_selfhalo(brute.as6(u, *shape), *shape, 63), want)


# ------------------------------------------------------------------ layout

def test_padded_plane_stride():
    shape = (5, 3, 2, 3)
    n, mx, my, mz = shape
    plane = n ** 3 * mx * my
    ps = ((plane - 1) // 512 + 1) * 512 + 512  # gpu/DArray.h:95-99 style padding
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=1, ps=ps)
    pad = np.ones(u.size, bool)
    for jz in range(mz):
        pad[jz * ps: jz * ps + plane] = False
    u[pad] = 12345.0
    got = u.copy()
    oracle.share_sweep(got, *shape, periodic=True, ps=ps)
    dense = oracle.new_field(*shape, oracle.FILL_RAND, seed=1)
    oracle.share_sweep(dense, *shape, periodic=True)
    assert np.array_equal(brute.as6(got, *shape, ps=ps), brute.as6(dense, *shape))
    assert np.all(got[pad] == 12345.0)


def test_pack_maps_against_slices():
    # gpu/v-26-messages/Faces.cc:243-293 as numpy slices of v[jz,jy,jx,iz,iy,ix]
    shape = (4, 3, 2, 5)
    n, mx, my, mz = shape
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=6)
    v = brute.as6(u, *shape)
    msg, off = oracle.pack(u, *shape)
    S = lambda k: msg[off[k]:off[k + 1]]  # noqa: E731
    # x faces xfs(iy,iz,jy,jz,s): iy fastest
    assert np.array_equal(S(0), v[:, :, 0, :, :, 0].ravel())
    assert np.array_equal(S(1), v[:, :, mx - 1, :, :, n - 1].ravel())
    assert np.array_equal(S(2), v[:, 0, :, :, 0, :].ravel())
    assert np.array_equal(S(3), v[:, my - 1, :, :, n - 1, :].ravel())
    assert np.array_equal(S(4), v[0, :, :, 0, :, :].ravel())
    assert np.array_equal(S(5), v[mz - 1, :, :, n - 1, :, :].ravel())
    for e in range(4):
        a, b = e & 1, e >> 1
        assert np.array_equal(S(6 + e), v[-b, -a, :, -b, -a, :].ravel())  # x-edge (sy,sz)=(a,b)
        assert np.array_equal(S(10 + e), v[-b, :, -a, -b, :, -a].ravel())  # y-edge (sx,sz)
        assert np.array_equal(S(14 + e), v[:, -b, -a, :, -b, -a].ravel())  # z-edge (sx,sy)
    for c in range(8):
        sx, sy, sz = c & 1, (c >> 1) & 1, c >> 2
        assert S(18 + c)[0] == v[-sz, -sy, -sx, -sz, -sy, -sx]
    assert [off[k + 1] - off[k] for k in range(26)] == \
        [n * n * my * mz] * 2 + [n * n * mx * mz] * 2 + [n * n * mx * my] * 2 + \
        [n * mx] * 4 + [n * my] * 4 + [n * mz] * 4 + [1] * 8


def test_opposite_slot_pairs_negated_directions():
    def direction(k):
        if k < 6:
            d = [0, 0, 0]
            d[k // 2] = 1 if k & 1 else -1
            return tuple(d)
        if k < 18:
            f, e = (k - 6) // 4, (k - 6) % 4
            a, b = (1 if e & 1 else -1), (1 if e >> 1 else -1)
            return [(0, a, b), (a, 0, b), (a, b, 0)][f]
        c = k - 18
        return tuple(1 if c >> i & 1 else -1 for i in range(3))
    for k in range(26):
        o = oracle.opposite_slot(k)
        assert oracle.opposite_slot(o) == k
        assert direction(o) == tuple(-x for x in direction(k))


# ------------------------------------------------------------ the compare

def test_compare_counts_and_reference_metric():
    shape = (3, 2, 2, 2)
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=1)
    v = u.copy()
    r = oracle.compare(u, v, *shape)
    assert r["pass"] and r["ndiff"] == 0 and r["max_ulp"] == 0
    v[5] = np.nextafter(v[5], 2.0)
    r = oracle.compare(u, v, *shape)
    assert r["ndiff"] == 1 and r["max_ulp"] == 1 and r["argmax"] == 5
    # gpu/main.cc:55-66: |u-v|/(|u|+|v|), skipped when the denominator is <= 1e-15
    w = u.copy()
    w[7] = -w[7]
    r = oracle.compare(u, w, *shape)
    assert r["max_rel"] == 1.0 and not r["pass"] and r["argmax"] == 7
    z, z2 = np.zeros_like(u), np.zeros_like(u)
    z2[3] = 1e-16
    r = oracle.compare(z, z2, *shape)
    assert r["max_rel"] == 0.0 and r["pass"] and r["ndiff"] == 1


# ------------------------------------ a planted mistake is caught (in the oracle output)

def test_planted_mistakes_are_detected():
    shape = (4, 3, 3, 3)
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=12)
    good = _run_sweep(u, shape, True)
    want = brute.share(brute.as6(u, *shape), *shape, periodic=True).ravel()
    assert np.array_equal(good, want)
    # a dropped term at one edge copy (an edge node missing its z partner)
    bad = good.copy()
    edge = np.flatnonzero(brute.multiplicity(*shape, periodic=True).ravel() == 4)[0]
    bad[edge] -= u[edge]
    assert not np.array_equal(bad, want)
    # a wrong wrap partner: sum with the next element instead of the last one
    bad2 = u.copy()
    oracle.share_sweep(bad2, *shape, periodic=False)
    assert not np.array_equal(bad2, want)


def test_planted_transposed_index_is_detected():
    """mx and my exchanged (a transposed element grid) on a non-cubic grid."""
    shape = (3, 4, 2, 3)
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=13)
    want = brute.share(brute.as6(u, *shape), *shape, periodic=True).ravel()
    bad = u.copy()
    oracle.share_sweep(bad, shape[0], shape[2], shape[1], shape[3], periodic=True)
    assert not np.array_equal(bad, want)
    good = u.copy()
    oracle.share_sweep(good, *shape, periodic=True)
    assert np.array_equal(good, want)


# ------------------------------------------- the 6-message schedule (z, y, x)

@pytest.mark.parametrize("shape", [(4, 3, 2, 2), (3, 1, 2, 3), (5, 2, 2, 2), (2, 1, 1, 1),
                                   (6, 3, 1, 2)])
@pytest.mark.parametrize("compute", [True, False])
def test_six_equals_sweep_and_staged_on_integers(shape, compute):
    """gpu/v-6-messages-basic/Faces.cc:65-172 sums the same copies as the
    26-message step: on integer fields every order gives the same bits."""
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=7)
    a = u.copy()
    oracle.share_six(a, *shape, nbr_mask=63, compute=compute)
    b = u.copy()
    oracle.share_sweep(b, *shape, periodic=True, compute=compute)
    assert np.array_equal(a, b)
    for mask in range(64):
        a = u.copy()
        oracle.share_six(a, *shape, nbr_mask=mask, compute=compute)
        b = u.copy()
        oracle.share_staged(b, *shape, nbr_mask=mask, compute=compute)
        assert np.array_equal(a, b), mask


def test_six_matches_bruteforce_integer():
    shape = (4, 3, 2, 3)
    u = oracle.new_field(*shape, oracle.FILL_INT, seed=17)
    a = u.copy()
    oracle.share_six(a, *shape)
    want = brute.share(brute.as6(u, *shape), *shape, periodic=True).ravel()
    assert np.array_equal(a, want)


def test_six_tree_is_z_then_y_then_x():
    """One periodic element of n=2: all 8 points are copies of one node.  The
    z phase pairs iz, then y pairs iy, then x pairs ix."""
    u = oracle.new_field(2, 1, 1, 1, oracle.FILL_RAND, seed=21)
    c = u.reshape(2, 2, 2)  # [iz][iy][ix]
    z = [[c[0, iy, ix] + c[1, iy, ix] for ix in range(2)] for iy in range(2)]
    y = [z[0][ix] + z[1][ix] for ix in range(2)]
    want = y[0] + y[1]
    a = u.copy()
    oracle.share_six(a, 2, 1, 1, 1)
    assert np.all(a == want)
    # the 26-message tree (x, then y, then z) differs on these values
    b = u.copy()
    oracle.share_sweep(b, 2, 1, 1, 1, periodic=True)
    x = [[c[iz, iy, 0] + c[iz, iy, 1] for iy in range(2)] for iz in range(2)]
    assert np.all(b == (x[0][0] + x[0][1]) + (x[1][0] + x[1][1]))


def test_six_random_within_rounding():
    shape = (5, 3, 2, 2)
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=23)
    a = u.copy()
    oracle.share_six(a, *shape)
    exact = brute.share(brute.as6(u, *shape), *shape, periodic=True).ravel()
    mult = brute.multiplicity(*shape, periodic=True).ravel()
    absum = brute.share(brute.as6(np.abs(u), *shape), *shape, periodic=True).ravel()
    bound = (mult - 1) * np.finfo(np.float64).eps * absum
    assert np.all(np.abs(a - exact) <= bound + 1e-300)


def _swap_xz(u, n, mx, my, mz):
    """The field with the x and z axes exchanged (element and point axes):
    a flat field of shape (n, mz, my, mx)."""
    v = brute.as6(u, n, mx, my, mz)  # [jz][jy][jx][iz][iy][ix]
    return np.ascontiguousarray(v.transpose(2, 1, 0, 5, 4, 3)).ravel()


def _swap_mask_xz(mask):
    return ((mask & 3) << 4) | (mask & 12) | ((mask >> 4) & 3)


@pytest.mark.parametrize("shape", [(4, 3, 2, 2), (3, 1, 2, 3), (5, 2, 3, 4), (2, 2, 1, 3)])
@pytest.mark.parametrize("compute", [True, False])
def test_six_is_the_staged_tree_with_x_and_z_exchanged(shape, compute):
    """The 6-message schedule (z pairs, then y, then x) evaluates, for every
    node and every rank position, the 26-message step's tree (x, then y, then
    z) with the x and z axes exchanged: on random fields, bit for bit, the
    six-phase result equals the staged step run on the axis-swapped field
    with the axis-swapped neighbour mask.  One launch of the 26-message sum
    kernel with its tree levels in z, y, x order therefore reproduces the
    6-message schedule exactly."""
    n, mx, my, mz = shape
    u = oracle.new_field(*shape, oracle.FILL_RAND, seed=31)
    for mask in range(64):
        a = u.copy()
        oracle.share_six(a, *shape, nbr_mask=mask, compute=compute)
        t = _swap_xz(u, *shape)
        oracle.share_staged(t, n, mz, my, mx, nbr_mask=_swap_mask_xz(mask), compute=compute)
        b = _swap_xz(t, n, mz, my, mx)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), mask
    # periodic: the sweep with x and z exchanged
    a = u.copy()
    oracle.share_six(a, *shape, nbr_mask=63, compute=compute)
    t = _swap_xz(u, *shape)
    oracle.share_sweep(t, n, mz, my, mx, periodic=True, compute=compute)
    assert np.array_equal(a.view(np.uint64), _swap_xz(t, n, mz, my, mx).view(np.uint64))
    # and the exchange is not the identity on these values: the 26-message
    # tree on the unswapped field differs somewhere (all nodes summed)
    if compute:
        c = u.copy()
        oracle.share_sweep(c, *shape, periodic=True)
        assert not np.array_equal(a, c)
