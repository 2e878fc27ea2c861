"""Pins the oracle's HLBVH build (oracle/hlbvh_oracle.cpp, restating
accelerator/hlbvh.cu) without trusting its own formulas:

* Morton codes: de-interleaved by a bit loop and compared with the cell index
  computed in FP64 (at most one cell apart, equal almost everywhere), plus the
  special cases the reference code fixes (x = 1024 clamps to 1023, zero extent).
* Canonical sort: a permutation, codes non-decreasing, ties by input index.
* Tree: structural invariants (every primitive once, exact box unions, counts),
  the split rules (bottom: centroid < midpoint; top: SAH bucket choice is the
  first minimum of a cost re-evaluated in FP64), and byte equality with an
  independent naive Python restatement on tiny inputs.
"""
import numpy as np
import pytest

import bvh_canon as C
import hlbvh_py
import oracle as O
import pbrtm


def soup(n, seed=1):
    return pbrtm.gen_soup(n, seed=seed)


def deinterleave(code):
    xyz = [0, 0, 0]
    for bit in range(10):
        for k in range(3):
            xyz[k] |= ((int(code) >> (3 * bit + k)) & 1) << bit
    return xyz


def test_morton_matches_fp64_cells():
    tri = soup(4000, seed=3)
    codes = O.morton_codes(tri)
    assert codes.max() < (1 << 30)
    c = 0.5 * (tri.min(axis=1).astype(np.float64) + tri.max(axis=1))
    lo, hi = c.min(axis=0), c.max(axis=0)
    md = int(np.argmax(hi - lo))
    r = (hi[md] - lo[md]) / 2
    mid = (lo + hi) / 2
    alo = np.where(np.arange(3) == md, lo, mid - r)
    ahi = np.where(np.arange(3) == md, hi, mid + r)
    cell = np.clip(np.floor((c - alo) / (ahi - alo) * 1024), 0, 1023).astype(np.int64)
    got = np.array([deinterleave(x) for x in codes])
    diff = np.abs(got - cell)
    assert diff.max() <= 1
    assert (diff.sum(axis=1) == 0).mean() > 0.99


def test_morton_special_cases():
    # centroids at both ends of the extent: the max maps to cell 1024 -> 1023
    tri = np.zeros((3, 3, 3), np.float32)
    tri[1] += 1.0
    tri[2] += np.float32([0.5, 0.25, 0.75])
    codes = O.morton_codes(tri)
    assert deinterleave(codes[0]) == [0, 0, 0]
    assert deinterleave(codes[1]) == [1023, 1023, 1023]
    # flat in z: the cube-adjusted frame still places every centroid
    tri = soup(500, seed=9)
    tri[:, :, 2] = 0.25
    codes = O.morton_codes(tri)
    cells = np.array([deinterleave(x) for x in codes])
    assert np.all(cells[:, 2] == cells[0, 2])
    # all centroids equal: every code is 0
    same = np.repeat(tri[:1], 10, axis=0)
    assert np.all(O.morton_codes(same) == 0)


def test_canonical_sort_order():
    tri = soup(20000, seed=5)
    tri[100:200] = tri[50]  # equal codes: ties must keep input order
    codes, order = O.morton_sort(tri)
    by_input = O.morton_codes(tri)
    assert np.array_equal(np.sort(order), np.arange(len(tri)))
    assert np.array_equal(codes, by_input[order])
    ref = np.lexsort((np.arange(len(tri)), by_input))
    assert np.array_equal(order, ref)


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (40, 3), (300, 4)])
def test_tree_equals_python_restatement(n, seed):
    tri = soup(n, seed=seed)
    assert O.OracleBvh(tri).export_canonical() == hlbvh_py.build_canonical(tri)


def test_tree_equals_python_restatement_clustered():
    rng = np.random.default_rng(7)
    # a few tight clusters (multi-primitive treelets) plus duplicates (multi-primitive leaves)
    centers = rng.random((5, 3)).astype(np.float32)
    pts = centers[rng.integers(0, 5, 150)][:, None, :] + rng.normal(0, 1e-3, (150, 3, 3)).astype(np.float32)
    pts[10:14] = pts[9]
    tri = np.ascontiguousarray(pts.astype(np.float32))
    b = O.OracleBvh(tri)
    assert b.info.max_treelet > 1
    assert b.export_canonical() == hlbvh_py.build_canonical(tri)


@pytest.mark.parametrize("threads", [2, 5, 16])
def test_threaded_build_is_the_same_tree(threads):
    """The CPU baseline's threaded build (parallel Morton pass and sort, subtrees of
    the top-level SAH on threads of their own, parallel bottom splits) exports the
    single-threaded build's bytes, clustered duplicates included."""
    tri = soup(200_000, seed=11)
    tri[1000:1400] = tri[999]
    one = O.OracleBvh(tri)
    many = O.OracleBvh(tri, threads=threads)
    assert many.info.total_nodes == one.info.total_nodes and many.info.top_nodes == one.info.top_nodes
    assert many.export_canonical() == one.export_canonical()


def test_tree_invariants_and_split_rules():
    tri = soup(30000, seed=11)
    b = O.OracleBvh(tri)
    nodes = C.parse(b.export_canonical())
    C.check_invariants(nodes, tri)
    cen = 0.5 * (tri.min(axis=1) + tri.max(axis=1))
    codes = O.morton_codes(tri)
    treelet = codes >> 9
    checked = 0
    for i, nd in enumerate(nodes):
        if nd.leaf:
            if nd.count > 1:  # a multi-primitive leaf only when a midpoint split is empty
                assert len(set(treelet[nd.prims].tolist())) == 1
                c = cen[nd.prims]
                d = c.max(axis=0) - c.min(axis=0)
                ax = 0 if (d[0] > d[1] and d[0] > d[2]) else (1 if d[1] > d[2] else 2)
                split = np.float32(0.5) * (c[:, ax].min() + c[:, ax].max())
                left = c[:, ax] < split
                assert left.all() or not left.any()
            continue
        Lp = C.subtree_prims(nodes, nd.left)
        Rp = C.subtree_prims(nodes, nd.right)
        same_treelet = len(set(treelet[np.concatenate([Lp, Rp])].tolist())) == 1
        if same_treelet:  # bottom level: strict midpoint partition on the axis
            c = cen[np.concatenate([Lp, Rp])][:, nd.axis]
            split = np.float32(0.5) * (c.min() + c.max())
            assert np.all(cen[Lp][:, nd.axis] < split)
            assert np.all(cen[Rp][:, nd.axis] >= split)
            checked += 1
        else:  # top level: whole treelets on each side
            assert not (set(treelet[Lp].tolist()) & set(treelet[Rp].tolist()))
    assert checked > 100


def test_top_level_sah_is_first_minimum():
    """The chosen bucket split minimises the FP64 SAH cost up to FP32 rounding."""
    tri = soup(6000, seed=13)
    b = O.OracleBvh(tri)
    nodes = C.parse(b.export_canonical())
    codes = O.morton_codes(tri)
    treelet = codes >> 9
    lo, hi = tri.min(axis=1).astype(np.float64), tri.max(axis=1).astype(np.float64)
    cen = 0.5 * (lo + hi)

    def area(l, h):
        d = h - l
        return 2 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2])

    tested = 0
    for nd in nodes:
        if nd.leaf:
            continue
        Lp, Rp = C.subtree_prims(nodes, nd.left), C.subtree_prims(nodes, nd.right)
        allp = np.concatenate([Lp, Rp])
        tids = np.unique(treelet[allp])
        if len(tids) < 2:
            continue
        tb = {t: (lo[allp][treelet[allp] == t].min(axis=0), hi[allp][treelet[allp] == t].max(axis=0)) for t in tids}
        tc = np.array([0.5 * (tb[t][0] + tb[t][1]) for t in tids])
        ax = nd.axis
        base, span = tc[:, ax].min(), tc[:, ax].max() - tc[:, ax].min()
        bk = np.minimum((24 * (tc[:, ax] - base) / span).astype(np.int64), 23)
        cnt = np.array([np.sum(treelet[allp] == t) for t in tids])
        total = area(lo[allp].min(axis=0), hi[allp].max(axis=0))
        costs = []
        for s in range(23):
            l, r = bk <= s, bk > s
            if not l.any() or not r.any():
                costs.append(np.inf)
                continue
            bl = (np.min([tb[t][0] for t in tids[l]], axis=0), np.max([tb[t][1] for t in tids[l]], axis=0))
            br = (np.min([tb[t][0] for t in tids[r]], axis=0), np.max([tb[t][1] for t in tids[r]], axis=0))
            costs.append(0.125 + (cnt[l].sum() * area(*bl) + cnt[r].sum() * area(*br)) / total)
        costs = np.array(costs)
        left_t = set(treelet[Lp].tolist())
        chosen = [s for s in range(23) if set(tids[bk <= s].tolist()) == left_t]
        assert chosen, "left side is not a bucket prefix"
        assert costs[chosen[0]] <= costs.min() * (1 + 1e-5)
        tested += 1
    assert tested > 50
