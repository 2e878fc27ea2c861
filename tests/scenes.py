"""Scene descriptions shared by the oracle and the library in the tests (the
generators live in the package: pbrtm/scenes.py)."""
from pbrtm.scenes import GREY, GREEN, RED, batch_jobs, box, cornell34, quad, soup_room, sphere_grid  # noqa: F401


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
