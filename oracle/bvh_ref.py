"""Pure-Python restatement of the reference's binned-SAH BVH builder.

TEST INFRASTRUCTURE ONLY (the checker of the product's C++ builder; small inputs).
Follows rustracer-core/src/bvh.rs:
  AABB / surface_area            bvh.rs:8-72
  BVHBuilder::build              bvh.rs:135-151
  build_subtree (children first) bvh.rs:153-215
  find_best_split (12 bins)      bvh.rs:224-319
  reverse_nodes                  bvh.rs:322-335
Arithmetic is float32 (numpy scalars) in the reference's order. The partition
step (select_nth_unstable_by) is restated as a stable sort by centroid, so the
node boxes and leaf prim *sets* match the reference whenever the centroids at a
split boundary are distinct; the order of prims inside a leaf may differ.
"""
from __future__ import annotations

import numpy as np

F = np.float32
FMAX = F(np.finfo(np.float32).max)
BINS = 12
MAX_LEAF = 4


def _sa(mn, mx):
    e = mx - mn
    return F(2.0) * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0])


def _union(a, b):
    return np.fmin(a[0], b[0]), np.fmax(a[1], b[1])


EMPTY = (np.full(3, FMAX, F), np.full(3, -FMAX, F))


def build(centroids, box_min, box_max):
    """Returns (nodes, prims): nodes as dicts {min, max, left_first, right_child, prim_count}."""
    c = np.asarray(centroids, F)
    bmin = np.asarray(box_min, F)
    bmax = np.asarray(box_max, F)
    prims = list(range(len(c)))
    nodes = []

    def best_split(start, end, bb):
        bb_sa = _sa(*bb)
        if not bb_sa > 0:
            return None
        inv = F(1.0) / bb_sa
        idx = prims[start:end]
        cmin = np.min(c[idx], axis=0)
        cmax = np.max(c[idx], axis=0)
        best = (0, 0, F(np.inf))
        for axis in range(3):
            rng = cmax[axis] - cmin[axis]
            if not rng > 0:
                continue
            scale = F(BINS) / rng
            bins = [[EMPTY, 0] for _ in range(BINS)]
            for p in idx:
                v = (c[p][axis] - cmin[axis]) * scale
                b = int(v) if v > 0 else 0
                b = min(b, BINS - 1)
                bins[b][1] += 1
                bins[b][0] = _union(bins[b][0], (bmin[p], bmax[p]))
            la, lc, ra, rc = [F(0)] * BINS, [0] * BINS, [F(0)] * BINS, [0] * BINS
            acc, n = EMPTY, 0
            for b in range(BINS):
                acc = _union(acc, bins[b][0])
                n += bins[b][1]
                with np.errstate(over="ignore", invalid="ignore"):
                    la[b] = _sa(*acc)
                lc[b] = n
            acc, n = EMPTY, 0
            for b in reversed(range(BINS)):
                acc = _union(acc, bins[b][0])
                n += bins[b][1]
                with np.errstate(over="ignore", invalid="ignore"):
                    ra[b] = _sa(*acc)
                rc[b] = n
            for b in range(BINS - 1):
                if lc[b] == 0 or rc[b + 1] == 0:
                    continue
                cost = F(1.0) + (la[b] * F(lc[b]) + ra[b + 1] * F(rc[b + 1])) * inv
                if cost < best[2]:
                    best = (axis, lc[b], cost)
        if best[2] < F(end - start):
            return best[0], best[1]
        return None

    def subtree(start, end):
        count = end - start
        if count == 0:
            return 0
        bb = EMPTY
        for p in prims[start:end]:
            bb = _union(bb, (bmin[p], bmax[p]))
        split = best_split(start, end, bb) if count > MAX_LEAF else None
        if split is None:
            nodes.append(dict(min=bb[0], max=bb[1], left_first=start, right_child=0, prim_count=count))
            return len(nodes) - 1
        axis, left = split
        seg = sorted(prims[start:end], key=lambda p: c[p][axis])
        prims[start:end] = seg
        mid = start + left
        li = subtree(start, mid)
        ri = subtree(mid, end)
        nodes.append(dict(min=bb[0], max=bb[1], left_first=li, right_child=ri, prim_count=0))
        return len(nodes) - 1

    subtree(0, len(prims))
    n = len(nodes)
    if n > 1:
        nodes.reverse()
        for nd in nodes:
            if nd["prim_count"] == 0:
                nd["left_first"] = n - 1 - nd["left_first"]
                nd["right_child"] = n - 1 - nd["right_child"]
    return nodes, prims
