"""GPU parity of the HLBVH path through the C ABI against the CPU oracle.

Bit-exact: Morton codes and canonical order; the canonical tree (node sets,
split axes, bound bit patterns); closest-hit primitive ids, t and barycentrics
on identical rays; any-hit booleans. Bounce directions go through cos and sin,
evaluated in double and rounded once to float on both sides, so they are
bit-identical unless the two double-precision libraries straddle a float
rounding boundary (a handful of rays at most); origins and directions are
also held to small tolerances, then the GPU's bounce rays are traced by both
sides and compared bit-exactly.
"""
import numpy as np
import pytest

import oracle as O
import pbrtm

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def soup(n, seed):
    return pbrtm.gen_soup(n, seed=seed)


def clustered(n, seed, far=True):
    """n triangles in a tiny cluster (+ one far triangle): few, huge treelets."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 1, 3)) * 1e-3
    tri = (c + rng.normal(0, 1e-5, (n, 3, 3))).astype(np.float32)
    if far:
        tri[0] += np.float32(10.0)
    return np.ascontiguousarray(tri)


def flat(tri):
    """All vertices on the plane z = 0.5: zero extent on one axis."""
    tri[:, :, 2] = np.float32(0.5)
    return tri


CASES = {
    "n1": lambda: soup(1, 0),
    "n2": lambda: soup(2, 1),
    "n3_dup": lambda: np.repeat(soup(1, 2), 3, axis=0),
    "n70": lambda: soup(70, 3),
    "n5000": lambda: soup(5000, 4),
    "n200k": lambda: soup(200_000, 5),
    "cluster3000": lambda: clustered(3000, 6),
    "cluster_dups": lambda: np.concatenate([clustered(500, 7), np.repeat(soup(1, 8), 200, axis=0)]),
    "flat_z": lambda: flat(soup(20000, 9)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_morton_sort_parity(ctx, case):
    tri = CASES[case]()
    log(f"morton_sort {case} n={len(tri)}")
    codes, order = ctx.morton_sort(tri)
    oc, oo = O.morton_sort(tri)
    assert np.array_equal(codes, oc)
    assert np.array_equal(order, oo)


@pytest.mark.parametrize("case", sorted(CASES))
def test_bvh_canonical_parity(ctx, case):
    tri = CASES[case]()
    log(f"bvh_build {case} n={len(tri)}")
    bvh = ctx.build(tri)
    got = bvh.export_canonical()
    want = O.OracleBvh(tri).export_canonical()
    assert len(got) == len(want)
    assert got == want
    lo = tri.min(axis=(0, 1))
    hi = tri.max(axis=(0, 1))
    assert np.array_equal(bvh.bounds(), np.r_[lo, hi])


def test_bvh_canonical_parity_1m(ctx):
    tri = soup(1_000_000, 0x5EED)
    log("bvh_build 1M")
    bvh = ctx.build(tri)
    log(f"stats {bvh.stats.as_dict()}")
    assert bvh.export_canonical() == O.OracleBvh(tri).export_canonical()


def test_intersect_and_bounce_parity_1m(ctx):
    tri = soup(1_000_000, 0x5EED)
    rays = pbrtm.gen_camera_rays(960, 540)
    log("intersect 1M primary")
    bvh = ctx.build(tri)
    ob = O.OracleBvh(tri)
    stats, bstats = np.zeros(3, np.uint64), np.zeros(3, np.uint64)
    prim, t, bary = bvh.intersect(rays, stats=stats)
    op, ot, obary, ostats = ob.intersect(rays)
    assert (op != O.NO_HIT).mean() > 0.1
    # binary traversal: the same node visits and triangle tests as the oracle
    bp, bt, bb = bvh.intersect(rays, stats=bstats, binary=True)
    assert np.array_equal(bstats[[0, 2]], ostats)
    assert np.array_equal(bp, op) and np.array_equal(bt, ot) and np.array_equal(bb, obary)
    # 8-wide traversal: fewer node visits; the quantized child boxes contain the
    # exact ones, so the triangle tests are a superset of the binary traversal's
    # (a few per cent more) and every hit is the same
    log(f"wide stats {stats.tolist()} binary {bstats.tolist()}")
    assert ostats[1] <= stats[2] <= 1.15 * ostats[1]
    assert stats[0] < ostats[0] / 2
    assert np.array_equal(prim, op)
    assert np.array_equal(t, ot)
    assert np.array_equal(bary, obary)
    log("ao_bounce")
    br = bvh.ao_bounce(rays, prim, bary)
    obr = O.ao_bounce(tri, rays, op, obary)
    active = obr[:, 3] > 0
    assert np.array_equal(br[:, 3] > 0, active)
    assert np.allclose(br[active, :3], obr[active, :3], rtol=2e-6, atol=1e-7)
    # directions go through cos/sin rounded from double; an ulp difference in a
    # rare case is amplified by z = sqrt(1 - x^2 - y^2) near grazing angles
    diff = np.abs(br[active, 4:7].astype(np.float64) - obr[active, 4:7]).max(axis=1)
    P = tri[op[active]].astype(np.float64)
    n = np.cross(P[:, 0] - P[:, 2], P[:, 1] - P[:, 2])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    cos_n = np.abs(np.sum(n * obr[active, 4:7], axis=1))
    log(f"bounce direction: max |diff| {diff.max():.3g}, {np.mean(diff > 0):.2e} not bit-equal")
    assert diff[cos_n > 0.05].max() < 5e-6
    assert diff.max() < 1e-3 and np.mean(diff > 0) < 1e-5
    log("intersect bounce rays")
    s2 = np.zeros(3, np.uint64)
    p2, t2, b2 = bvh.intersect(br, stats=s2)
    q2, u2, c2, os2 = ob.intersect(br)
    assert np.array_equal(p2, q2)
    assert np.array_equal(t2, u2)
    assert np.array_equal(b2, c2)
    assert os2[1] <= s2[2] <= 1.15 * os2[1]
    log("any hit")
    pa, _, _ = bvh.intersect(br, any_hit=True, want_bary=False)
    qa, _, _, _ = ob.intersect(br, any_hit=True)
    assert np.array_equal(pa != O.NO_HIT, qa != O.NO_HIT)


def test_headline_config_parity_10m(ctx):
    """The benchmark's exact workload: 10M triangles of seed 0x5EED, the canonical
    tree, then closest hits (prim, t, barycentrics) of the 1920x1080 primary rays
    and of their cosine bounce, all equal to the oracle (hlbvh.cu:384-624, :231-289)."""
    n = 10_000_000
    log(f"headline: soup {n}")
    tri = soup(n, 0x5EED)
    bvh = ctx.build(tri)
    log(f"headline: stats {bvh.stats.as_dict()}")
    ob = O.OracleBvh(tri)
    log("headline: canonical trees")
    assert bvh.export_canonical() == ob.export_canonical()
    rays = pbrtm.gen_camera_rays(1920, 1080)
    log("headline: primary rays")
    prim, t, bary = bvh.intersect(rays)
    op, ot, obary, _ = ob.intersect(rays, threads=16)
    assert (op != O.NO_HIT).mean() > 0.1
    assert np.array_equal(prim, op) and np.array_equal(t, ot) and np.array_equal(bary, obary)
    log("headline: bounce rays")
    br = bvh.ao_bounce(rays, prim, bary)
    p2, t2, b2 = bvh.intersect(br)
    q2, u2, c2, _ = ob.intersect(br, threads=16)
    assert np.array_equal(p2, q2) and np.array_equal(t2, u2) and np.array_equal(b2, c2)
    bvh.free()


def test_deep_tree_spills_and_matches_brute_force(ctx):
    """A tree deeper than the 12-entry LDS stack (wide depth > 12): rays along the
    chain push one entry per wide level, so the traversal's memory spill area is
    used; hits equal brute force (no exact t ties from x = -1e-12) and the oracle's
    traversal (ties resolved by traversal order, from x = -1)."""
    import scenes

    tri = scenes.deep_chain()
    bvh = ctx.build(tri)
    log(f"deep chain stats {bvh.stats.as_dict()}")
    assert bvh.stats.wide_depth > 12
    assert bvh.export_canonical() == O.OracleBvh(tri).export_canonical()
    rays = scenes.chain_rays(4096)
    prim, t, _ = bvh.intersect(rays)
    bp, bt = O.brute_intersect(tri, rays)
    assert (bp != O.NO_HIT).all()
    assert np.array_equal(prim, bp) and np.array_equal(t, bt)
    far = scenes.chain_rays(4096, seed=4, x0=-1.0)
    prim, t, bary = bvh.intersect(far)
    op, ot, ob, _ = O.OracleBvh(tri).intersect(far)
    assert np.array_equal(prim, op) and np.array_equal(t, ot) and np.array_equal(bary, ob)
    pa, _, _ = bvh.intersect(far, any_hit=True, want_bary=False)
    assert np.array_equal(pa != O.NO_HIT, op != O.NO_HIT)


def test_intersect_brute_force_small(ctx):
    tri = soup(3000, 21)
    rng = np.random.default_rng(1)
    k = rng.integers(0, 3000, 5000)
    w = (rng.dirichlet((1.0, 1.0, 1.0), 5000) - 1 / 3) * 1.1 + 1 / 3
    tgt = np.einsum("ij,ijk->ik", w, tri[k].astype(np.float64))
    org = rng.random((5000, 3)) * 1.4 - 0.2
    d = (tgt - org) / np.linalg.norm(tgt - org, axis=1, keepdims=True)
    rays = np.zeros((5000, 8), np.float32)
    rays[:, :3], rays[:, 3], rays[:, 4:7] = org, np.inf, d
    rays[::7, 3] = 0.4  # finite t_max
    rays[::11, 3] = 0.0  # inactive
    bvh = ctx.build(tri)
    prim, t, _ = bvh.intersect(rays)
    bp, bt = O.brute_intersect(tri, rays)
    assert np.array_equal(prim, bp)
    assert np.array_equal(t, bt)


def test_build_size_sequence_in_one_context():
    """Builds of growing and shrinking treelet counts in one fresh context, so that
    the scratch buffers are reallocated between builds (treelet records growing
    while the top-level lists keep their size, and the reverse), each equal to the
    oracle's tree."""
    c = pbrtm.Context(0)
    try:
        for k, n in enumerate((40, 100, 180, 250, 90, 3000, 200, 70_000, 130, 66, 1000)):
            tri = soup(n, 100 + k)
            log(f"build {k}: {n} triangles")
            assert c.build(tri).export_canonical() == O.OracleBvh(tri).export_canonical(), n
    finally:
        c.close()


def skewed(n, seed, power=3.0):
    """A thin sheet whose centroids crowd towards one corner (u^3, v^3): uneven SAH
    splits all the way down, so that nodes of 65..2048 treelets hold deep,
    lopsided subtrees."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 1, 3))
    c[:, :, :2] = c[:, :, :2] ** power
    c[:, :, 2] = 0.5 + 0.01 * c[:, :, 2]
    return np.ascontiguousarray((c + rng.normal(0, 2e-4, (n, 3, 3))).astype(np.float32))


@pytest.mark.parametrize("n", [2500, 6000, 40000])
def test_skewed_sheet_canonical_parity(ctx, n):
    """1.8k, 3.6k and 13.9k treelets: the root inside the 65..2048-treelet finisher,
    just above it, and several fused levels before it."""
    tri = skewed(n, 11)
    got = ctx.build(tri).export_canonical()
    assert got == O.OracleBvh(tri).export_canonical()


def test_empty_aggregate(ctx):
    tri = np.zeros((0, 3, 3), np.float32)
    bvh = ctx.build(tri)
    assert bvh.export_canonical() == b""
    rays = pbrtm.gen_camera_rays(8, 8)
    prim, t, _ = bvh.intersect(rays)
    assert np.all(prim == pbrtm.NO_HIT)


def test_nonfinite_vertex_is_an_error(ctx):
    tri = soup(100, 3)
    tri[17, 1, 2] = np.nan
    with pytest.raises(pbrtm.PbrtmError) as e:
        ctx.build(tri)
    assert e.value.code == pbrtm.ERR_DEVICE
    # the context stays usable after the error
    assert ctx.build(soup(10, 3)).export_canonical() == O.OracleBvh(soup(10, 3)).export_canonical()


def test_indexed_mesh_equals_soup(ctx):
    rng = np.random.default_rng(5)
    verts = rng.random((500, 3)).astype(np.float32)
    idx = rng.integers(0, 500, (2000, 3)).astype(np.uint32)
    idx[:, 1] = (idx[:, 0] + 1) % 500
    idx[:, 2] = (idx[:, 0] + 2) % 500
    tri = np.ascontiguousarray(verts[idx])
    a = ctx.build(verts, idx=idx).export_canonical()
    b = ctx.build(tri).export_canonical()
    assert a == b
    bad = idx.copy()
    bad[5, 0] = 500
    with pytest.raises(pbrtm.PbrtmError):
        ctx.build(verts, idx=bad)


def test_cpp_interface_example():
    """The C++ interface (host/hlbvh.hpp) end to end: a unit cube, six rays."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(pbrtm.LIB_PATH), "hlbvh_example")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0] == "bounds 0 0 0 1 1 1"
    assert lines[1:] == ["1 1 1", "3 1 1", "8 1 1", "11 1 1", "5 1 1", "-1 0 0"]
