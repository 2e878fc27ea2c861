"""Pins the oracle's traversal and triangle test (hlbvh.cu:179-289,
shapes/triangle.cu:217-328) and its ambient-occlusion bounce
(integrators/ambient_occlusion.cu:28-52):

* BVH closest hit == brute force over every triangle with the same test;
  any hit == "brute force found something";
* the triangle test agrees with an FP64 Moller-Trumbore (t within 1e-6) away
  from edges, and is watertight on shared edges/vertices;
* bounce rays are unit length, on the face-forwarded normal's side, cosine
  distributed (chi-square), with the origin inside a few error-bounds of the hit.
"""
import numpy as np

import oracle as O
import pbrtm


def aimed_rays(tri, n, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, tri.shape[0], n)
    w = (rng.dirichlet((1.0, 1.0, 1.0), n) - 1 / 3) * 1.1 + 1 / 3  # sums to 1; a few just outside
    target = np.einsum("ij,ijk->ik", w, tri[k].astype(np.float64))
    org = rng.random((n, 3)) * 1.4 - 0.2
    d = target - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = np.inf
    rays[:, 4:7] = d
    return rays


def test_closest_and_any_hit_equal_brute_force():
    tri = pbrtm.gen_soup(3000, seed=21)
    bvh = O.OracleBvh(tri)
    rays = aimed_rays(tri, 3000, 1)
    p1, t1, _, _ = bvh.intersect(rays)
    p2, t2 = O.brute_intersect(tri, rays)
    assert (p2 != O.NO_HIT).mean() > 0.5
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1, t2)
    pa, _, _, _ = bvh.intersect(rays, any_hit=True)
    assert np.array_equal(pa != O.NO_HIT, p2 != O.NO_HIT)
    # finite t_max: hits beyond it are excluded
    rays[:, 3] = 0.5
    p3, t3, _, _ = bvh.intersect(rays)
    p4, t4 = O.brute_intersect(tri, rays)
    assert np.array_equal(p3, p4) and np.all(t3[p3 != O.NO_HIT] <= 0.5)


def mt64(o, d, p0, p1, p2):
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = e1 @ pv
    if abs(det) < 1e-300:
        return None
    tv = o - p0
    u = (tv @ pv) / det
    qv = np.cross(tv, e1)
    v = (d @ qv) / det
    t = (e2 @ qv) / det
    if u < 0 or v < 0 or u + v > 1 or t <= 0:
        return None
    return t, u, v


def test_triangle_test_matches_fp64():
    rng = np.random.default_rng(3)
    agree = hits = 0
    for _ in range(4000):
        tri9 = rng.random(9).astype(np.float32)
        o = (rng.random(3) * 3 - 1).astype(np.float32)
        tgt = tri9.reshape(3, 3).mean(axis=0) + rng.normal(0, 0.15, 3)
        d = (tgt - o) / np.linalg.norm(tgt - o)
        ray = np.zeros(8, np.float32)
        ray[0:3], ray[4:7] = o, d
        hit, out = O.intersect_triangle(ray, tri9, np.inf)
        ref = mt64(ray[0:3].astype(np.float64), ray[4:7].astype(np.float64),
                   *tri9.reshape(3, 3).astype(np.float64))
        if ref is None or hit is False:
            # may only disagree when the ray passes within 1e-5 of an edge
            if (ref is None) != (not hit):
                t, u, v = ref if ref is not None else (0, 0.5, 0.5)
                if ref is not None:
                    assert min(u, v, 1 - u - v) < 1e-5
            else:
                agree += 1
            continue
        hits += 1
        agree += 1
        t, u, v = ref
        assert abs(out[3] - t) <= 1e-6 * t + 1e-7
        assert abs(out[1] - u) < 1e-5 and abs(out[2] - v) < 1e-5
    assert hits > 500 and agree > 3900


def test_watertight_shared_edges():
    # a fan of 8 triangles around the origin in the z = 1 plane; rays aimed
    # exactly at the centre vertex and along the shared edges never miss
    ang = np.linspace(0, 2 * np.pi, 9)[:-1]
    rim = np.stack([np.cos(ang), np.sin(ang), np.ones_like(ang)], 1).astype(np.float32)
    tri = np.zeros((8, 3, 3), np.float32)
    tri[:, 0] = [0, 0, 1]
    tri[:, 1] = rim
    tri[:, 2] = np.roll(rim, -1, axis=0)
    rays = []
    for k in range(8):
        for s in np.linspace(0.0, 0.9, 10):
            p = rim[k] * np.float32(s)
            p[2] = 1
            o = np.float32([0.013, -0.007, 0])
            d = (p - o) / np.linalg.norm(p - o)
            rays.append(np.r_[o, np.inf, d, 0])
    rays = np.asarray(rays, np.float32)
    prim, t = O.brute_intersect(tri, rays)
    assert np.all(prim != O.NO_HIT)
    bvh = O.OracleBvh(tri)
    p2, t2, _, _ = bvh.intersect(rays)
    assert np.all(p2 != O.NO_HIT)


def test_ao_bounce_properties():
    tri = pbrtm.gen_soup(2000, seed=4)
    tri *= np.float32(1.0)
    bvh = O.OracleBvh(tri)
    rays = aimed_rays(tri, 20000, 5)
    prim, t, bary, _ = bvh.intersect(rays)
    out = O.ao_bounce(tri, rays, prim, bary)
    hit = prim != O.NO_HIT
    assert hit.mean() > 0.5
    d = out[hit, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5)
    assert np.all(out[~hit, 3] == 0)
    P = tri[prim[hit]].astype(np.float64)
    n = np.cross(P[:, 0] - P[:, 2], P[:, 1] - P[:, 2])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    rd = rays[hit, 4:7].astype(np.float64)
    ff = np.where((np.sum(n * -rd, axis=1) < 0)[:, None], -n, n)
    cos = np.sum(d * ff, axis=1)
    assert np.all(cos >= -1e-6)
    # cosine-weighted: cos^2 is uniform on [0, 1]
    hist, _ = np.histogram(cos ** 2, bins=10, range=(0, 1))
    exp = len(cos) / 10
    chi2 = np.sum((hist - exp) ** 2 / exp)
    assert chi2 < 30  # 9 dof, p ~ 4e-4
    # origin: the barycentric hit point pushed off the surface by a tiny amount
    ph = np.einsum("ij,ijk->ik", bary[hit].astype(np.float64), P)
    off = out[hit, 0:3] - ph
    assert np.all(np.linalg.norm(off, axis=1) < 1e-5)
    side = np.sum(off * n, axis=1) * np.sum(d * n, axis=1)
    assert np.all(side >= 0)  # pushed to the side the bounce leaves from
