"""Independent brute-force model of the DSS step for tiny grids (numpy).

It never looks at passes, messages or owners: every copy gets the global id of
the mesh node it sits on, and the result at a copy is the sum over all copies
with the same id.  That fixes *which* values are summed; it says nothing about
the order, so it pins the oracle bit for bit only on integer-valued fields
(exact in any order) and within a rounding bound on others.
"""
from __future__ import annotations

import numpy as np


def as6(u: np.ndarray, n, mx, my, mz, ps=0) -> np.ndarray:
    """View of a flat field as v[jz, jy, jx, iz, iy, ix] (padding dropped)."""
    plane = n ** 3 * mx * my
    ps = ps or plane
    rows = [u[jz * ps: jz * ps + plane] for jz in range(mz)]
    return np.stack(rows).reshape(mz, my, mx, n, n, n)


def node_ids(n, mx, my, mz, periodic=True, nbr_mask=None):
    """Global node id of every copy, shape (mz,my,mx,n,n,n).

    nbr_mask (6 bits x-,x+,y-,y+,z-,z+) models self-halo neighbours of an open
    rank: when the neighbour across a face exists, that face's nodes are also
    copies of the node held by the neighbour -- which, for a self halo, is the
    opposite face of this same field.  Those received copies are returned
    separately by received_copies(); node ids stay those of the open grid.
    """
    K = n - 1
    jz, jy, jx, iz, iy, ix = np.indices((mz, my, mx, n, n, n))
    gx, gy, gz = jx * K + ix, jy * K + iy, jz * K + iz
    if periodic:
        Gx, Gy, Gz = mx * K, my * K, mz * K
        gx, gy, gz = gx % Gx, gy % Gy, gz % Gz
    else:
        Gx, Gy, Gz = mx * K + 1, my * K + 1, mz * K + 1
    return (gz * Gy + gy) * Gx + gx, (Gx, Gy, Gz)


def share(v6: np.ndarray, n, mx, my, mz, periodic=True, compute=True) -> np.ndarray:
    ids, (Gx, Gy, Gz) = node_ids(n, mx, my, mz, periodic)
    sums = np.bincount(ids.ravel(), weights=v6.ravel(), minlength=Gx * Gy * Gz)
    out = sums[ids]
    if not compute:
        surf = surface_mask(n, mx, my, mz, periodic)
        out = np.where(surf, out, v6)
    return out


def multiplicity(n, mx, my, mz, periodic=True) -> np.ndarray:
    ids, (Gx, Gy, Gz) = node_ids(n, mx, my, mz, periodic)
    cnt = np.bincount(ids.ravel(), minlength=Gx * Gy * Gz)
    return cnt[ids]


def surface_mask(n, mx, my, mz, periodic=True) -> np.ndarray:
    """Copies of nodes on the rank surface."""
    K = n - 1
    jz, jy, jx, iz, iy, ix = np.indices((mz, my, mx, n, n, n))
    gx, gy, gz = jx * K + ix, jy * K + iy, jz * K + iz
    if periodic:
        return (gx % (mx * K) == 0) | (gy % (my * K) == 0) | (gz % (mz * K) == 0)
    return ((gx == 0) | (gx == mx * K) | (gy == 0) | (gy == my * K) | (gz == 0) | (gz == mz * K))


def share_selfhalo(v6: np.ndarray, n, mx, my, mz, nbr_mask: int, compute=True) -> np.ndarray:
    """Open rank whose existing neighbours are stood in for by its own field.

    The neighbour across the x- face holds, at the shared nodes, the values
    this field has on its x+ face (self halo), and so on for every face, edge
    and corner direction whose face bits are all set.  The result at a local
    copy is the sum of the node's local copies plus the received copies.
    """
    K = n - 1
    ids, (Gx, Gy, Gz) = node_ids(n, mx, my, mz, periodic=False)
    sums = np.bincount(ids.ravel(), weights=v6.ravel(), minlength=Gx * Gy * Gz)
    has = [(nbr_mask >> b) & 1 for b in range(6)]
    # For every direction (dx,dy,dz) != 0 with all needed neighbours, the
    # neighbour's copies of our boundary nodes: a node at our low x surface
    # (gx == 0) is at the neighbour's high x surface, whose values (self halo)
    # are our own values at gx == Gx-1 with the same other coordinates.
    jz, jy, jx, iz, iy, ix = np.indices((mz, my, mx, n, n, n))
    gx, gy, gz = jx * K + ix, jy * K + iy, jz * K + iz
    extra = np.zeros(Gx * Gy * Gz)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if (dx, dy, dz) == (0, 0, 0):
                    continue
                ok = True
                for d, b0 in ((dx, 0), (dy, 2), (dz, 4)):
                    if d == -1 and not has[b0]:
                        ok = False
                    if d == 1 and not has[b0 + 1]:
                        ok = False
                if not ok:
                    continue
                # copies held by neighbour (dx,dy,dz): the neighbour's copy at
                # its side -d of our node at side d.  Self halo: that copy is
                # our own copy at side -d, other axes unchanged.
                sel = np.ones(v6.shape, bool)
                tgt = [gx, gy, gz]
                G = (Gx, Gy, Gz)
                for a, d in enumerate((dx, dy, dz)):
                    src = (gx, gy, gz)[a]
                    if d == -1:      # our node at g=0 <- our copy at g=G-1
                        sel &= src == G[a] - 1
                        tgt[a] = np.zeros_like(src)
                    elif d == 1:     # our node at g=G-1 <- our copy at g=0
                        sel &= src == 0
                        tgt[a] = np.full_like(src, G[a] - 1)
                tid = (tgt[2] * Gy + tgt[1]) * Gx + tgt[0]
                np.add.at(extra, tid[sel], v6[sel])
    total = sums + extra
    out = total[ids]
    if not compute:
        surf = surface_mask(n, mx, my, mz, periodic=False)
        out = np.where(surf, out, v6)
    return out


def mask_for_rank(rank, lx, ly, lz) -> int:
    """Face-neighbour bits of an open rank grid (gpu/v-26-messages/Faces.cc:92-103)."""
    kz = rank // (lx * ly)
    ky = (rank - kz * lx * ly) // lx
    kx = rank - (kz * ly + ky) * lx
    m = 0
    if kx > 0: m |= 1
    if kx < lx - 1: m |= 2
    if ky > 0: m |= 4
    if ky < ly - 1: m |= 8
    if kz > 0: m |= 16
    if kz < lz - 1: m |= 32
    return m
