"""Parse the canonical tree bytes (include/pbrtm.h, pbrtm_bvh_export_canonical)
and check the structural invariants every HLBVH tree must satisfy."""
from __future__ import annotations

import numpy as np


class Node:
    __slots__ = ("leaf", "axis", "lo", "hi", "count", "prims", "left", "right")

    def __init__(self):
        self.left = self.right = None
        self.prims = None


def parse(buf: bytes):
    """Return the list of nodes in preorder (index 0 = root) with child links."""
    w = np.frombuffer(buf, np.uint32)
    nodes, pos = [], 0
    # each stack entry: (parent index, is_right)
    pending = [(-1, False)]
    while pending:
        parent, is_right = pending.pop()
        nd = Node()
        tag = int(w[pos])
        nd.leaf = tag == 1
        nd.axis = -1 if nd.leaf else tag >> 8
        nd.lo = w[pos + 1:pos + 4].view(np.float32).copy()
        nd.hi = w[pos + 4:pos + 7].view(np.float32).copy()
        nd.count = int(w[pos + 7])
        pos += 8
        idx = len(nodes)
        nodes.append(nd)
        if parent >= 0:
            if is_right:
                nodes[parent].right = idx
            else:
                nodes[parent].left = idx
        if nd.leaf:
            nd.prims = w[pos:pos + nd.count].copy()
            pos += nd.count
        else:
            pending.append((idx, True))
            pending.append((idx, False))
    assert pos == len(w), "trailing bytes"
    return nodes


def tri_boxes(tri: np.ndarray):
    return tri.min(axis=1), tri.max(axis=1)


def check_invariants(nodes, tri: np.ndarray):
    """Every primitive in exactly one leaf; leaf box = exact union of its
    triangles' boxes; interior box = exact union of the children; counts add up."""
    n = tri.shape[0]
    lo, hi = tri_boxes(tri)
    seen = np.zeros(n, np.int64)
    for nd in nodes:
        if nd.leaf:
            assert nd.count == len(nd.prims) >= 1
            seen[nd.prims] += 1
            assert np.array_equal(nd.lo, lo[nd.prims].min(axis=0) + 0.0)
            assert np.array_equal(nd.hi, hi[nd.prims].max(axis=0) + 0.0)
        else:
            L, R = nodes[nd.left], nodes[nd.right]
            assert nd.count == L.count + R.count
            assert np.array_equal(nd.lo, np.minimum(L.lo, R.lo))
            assert np.array_equal(nd.hi, np.maximum(L.hi, R.hi))
            assert nd.axis in (0, 1, 2)
    assert np.all(seen == 1), "a primitive is missing or duplicated"
    assert nodes[0].count == n


def subtree_prims(nodes, i):
    out, st = [], [i]
    while st:
        j = st.pop()
        nd = nodes[j]
        if nd.leaf:
            out.append(nd.prims)
        else:
            st += [nd.right, nd.left]
    return np.concatenate(out)


def depth(nodes) -> int:
    best, st = 0, [(0, 1)]
    while st:
        i, d = st.pop()
        best = max(best, d)
        nd = nodes[i]
        if not nd.leaf:
            st += [(nd.left, d + 1), (nd.right, d + 1)]
    return best
