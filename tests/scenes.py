"""Scene descriptions shared by the oracle and the library in the tests (the
generators live in the package: pbrtm/scenes.py)."""
from pbrtm.scenes import GREY, GREEN, RED, batch_jobs, box, cornell34, quad, shapes_room, soup_room, sphere_grid  # noqa: F401


def cornell_emitters(width=32, height=32, spp=16, max_depth=4):
    """cornell34 with every triangle given its own diffuse albedo (U[0.1, 0.9]^3 from
    numpy's seeded generator) and three emitters: the ceiling quad (scale 20), a
    two-sided orange quad low on the back wall (6) and a two-sided blue quad on the
    left wall (9). For gradient tests."""
    import numpy as np

    sc = cornell34(width, height, spp=spp, max_depth=max_depth)
    extra = quad((150.0, 50.0, 559.0), (250.0, 50.0, 559.0), (250.0, 120.0, 559.0), (150.0, 120.0, 559.0)) + \
        quad((1.0, 300.0, 300.0), (1.0, 300.0, 400.0), (1.0, 380.0, 400.0), (1.0, 380.0, 300.0))
    tri = np.concatenate([np.asarray(sc["tri"], np.float32), np.array(extra, np.float32).reshape(-1, 9)])
    n = len(tri)
    rng = np.random.default_rng(0x5CE7E)
    sc.update(
        tri=tri,
        tri_material=np.arange(n, dtype=np.uint32),
        tri_light=np.concatenate([np.asarray(sc["tri_light"], np.int32), np.array([1, 1, 2, 2], np.int32)]),
        tri_reverse=np.concatenate([np.zeros(n - 4, np.uint8), np.ones(4, np.uint8)]),
        materials=None,
        albedo=rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32),
        lights=[{"rgb": (1.0, 1.0, 1.0), "scale": 20.0}, {"rgb": (1.0, 0.5, 0.2), "scale": 6.0, "two_sided": True},
                {"rgb": (0.3, 0.5, 1.0), "scale": 9.0, "two_sided": True}],
    )
    return sc


def deep_chain(levels=48, per_level=16):
    """A triangle set whose 8-wide tree is deeper than the traversal's 12-entry LDS
    stack, so that closest-hit rays along +x spill their stack to memory.

    Cluster k (k = 1..levels) holds `per_level` triangles with bounding-box centres
    at x = 0.5 * 2^-k * (1 - 0.03 j), y = z = 0.5: clusters do not interleave and no
    two centres coincide, so every leaf holds one triangle (the order of the
    triangles inside a multi-primitive leaf is left open by the reference, whose
    treelet sort places them with atomics, and with exact t ties that order picks
    the hit). Midpoint splits of the centroid bounds on x (the only axis with
    extent) then peel one cluster per binary level.
    Cluster k spans 0.8 * 0.9^k in y and z, more than every deeper cluster, so the
    greedy wide collapse (largest surface area first) fills a wide node with
    cluster k's own nodes before it opens the deeper chain: the wide tree is 40
    levels deep. A ray along +x through the middle
    enters the deeper chain first (near child) and pushes cluster slots at every
    level."""
    import numpy as np

    tri = []
    for k in range(1, levels + 1):
        c0 = 0.5 * 2.0 ** -k
        a = 0.4 * 0.9 ** k
        for j in range(per_level):
            c = c0 * (1.0 - 0.03 * j)
            d = 0.01 * c0
            tri.append([[c - d, 0.5 - a, 0.5 - 0.3 * a], [c + d, 0.5 + a, 0.5 - a], [c, 0.5 + 0.2 * a, 0.5 + a]])
    return np.ascontiguousarray(np.array(tri, np.float32))


def chain_rays(n, seed=3, x0=-1e-12):
    """Rays from x = x0 towards +x through the chain's middle, with small tilts.
    From x0 = -1e-12 the clusters' hit distances differ (no exact t ties); from
    x0 = -1 every cluster below 2^-25 is hit at t = 1.0 exactly."""
    import numpy as np

    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0] = x0
    rays[:, 1:3] = 0.5 + rng.uniform(-0.01, 0.01, (n, 2))
    rays[:, 3] = np.inf
    d = np.stack([np.ones(n), rng.uniform(-0.02, 0.02, n), rng.uniform(-0.02, 0.02, n)], 1)
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    return rays
