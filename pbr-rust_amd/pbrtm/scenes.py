"""Synthetic scenes for the path tracer, as scene dicts for pbrtm.Scene (and the
oracle). Deterministic: every random choice comes from PCG32 streams or a
64-bit integer hash, never from numpy's generators.

cornell34: the Cornell box of the reference's example scene
(example/cornell-box-specular.pbrt:1-91) with its sphere and rotated block
replaced by two axis-aligned boxes: 5 wall quads (10 triangles), a grey diffuse
short box and a specular silver (metal-Ag, roughness 0) tall box (12 triangles
each), and the example's emissive quad (rgb L 1, scale 20, 2 triangles).

sphere_grid: n x n UV spheres (tess x tess quads, 2 tess^2 triangles each,
radius 0.4, pitch 1.0) radially displaced by value noise of amplitude 0.1, on a
ground plane; each sphere's material is drawn from {diffuse rgb U[0.1, 0.9]^3,
conductor Cu roughness U[0.05, 0.4], dielectric eta 1.5 roughness 0}; 32
emissive quads (64 triangles) at y = 5 with L = rgb(1, 1, 1), scale 10. At
n = 32, tess = 32 it holds 2,097,218 triangles.

soup_room: the HLBVH benchmark's seeded triangle soup (centroids uniform in the
unit cube, log-normal edges) inside a closed room (x, y in [-0.5, 1.5], z in
[-1.5, 1.5]; grey floor, ceiling and end walls, red left and green right wall)
lit by one emissive quad under the ceiling (rgb L 1, scale 8); soup triangle k
takes diffuse material k mod 8 (albedo from PCG32). The camera sits inside the
room at z = -1.3 looking at the cube's centre.

grad_room: the differentiable-rendering workload: the soup_room geometry with
every triangle given its own diffuse material (albedo U[0.1, 0.9]^3 from a
64-bit hash) and the ceiling light split into 8 emitters (quads in a 4 x 2
grid, scales U[4, 8]); perturbed() scales every albedo and emitter scale by
1 + U[-a, a] (albedos clamped to [0.01, 0.99]) to make the target.

shapes_room: cornell34 with analytic shapes (Shape "sphere" / "disk") and object
instancing: a smooth glass sphere (the example scene's), a rough copper partial
sphere (zmin, zmax, phimax) under a rotation, a diffuse sphere with
ReverseOrientation, an emissive sphere and an emissive disk facing down, a
diffuse annular disk under a mirroring transform (Scale -1: the disk's normal
flips with the handedness), and the short box instanced twice more through
ObjectInstance transforms (rotation + translation). Each transform carries its
inverse as composed from the directives (Translate, Rotate, Scale).

batch_jobs: the mixed batch of the batched-render benchmark, drawn from one PCG32
stream (seed 0xBA7C): per job a width and a height each uniform in [lo, hi], spp
from {1, 4, 16, 64}, max depth uniform in 1..8 and a batch_scene of n triangles,
n log-uniform in [10^4, 10^6].

batch_scene: the HLBVH benchmark's triangle soup (centroids in the unit cube,
log-normal edges) seen by its camera, all diffuse with albedo 0.5, lit by one
emissive quad above the cube (facing down, rgb 1 scale 8).
"""
import numpy as np

GREY, RED, GREEN = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class Pcg32:
    """PCG32 (O'Neill), the reference's RNG (util/rng.h), scalar."""

    MULT = 0x5851F42D4C957F2D

    def __init__(self, seq, offset=None):
        def mix(v):
            v ^= v >> 31
            v = (v * 0x7FB5D329728EA185) & 0xFFFFFFFFFFFFFFFF
            v ^= v >> 27
            v = (v * 0x81DADEF4BC2DD44D) & 0xFFFFFFFFFFFFFFFF
            return v ^ (v >> 33)

        self.state = 0
        self.inc = ((seq << 1) | 1) & 0xFFFFFFFFFFFFFFFF
        self.next_u32()
        self.state = (self.state + (mix(seq) if offset is None else offset)) & 0xFFFFFFFFFFFFFFFF
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & 0xFFFFFFFFFFFFFFFF
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self):
        return min(float(np.float32(self.next_u32() * 2.0 ** -32)), float.fromhex('0x1.fffffep-1'))


def quad(a, b, c, d):
    """Triangles of indices [0 1 2 2 3 0] over the quad a b c d."""
    return [a + b + c, c + d + a]


def box(x0, x1, y0, y1, z0, z1):
    p = lambda x, y, z: (x, y, z)
    faces = [
        (p(x0, y0, z0), p(x1, y0, z0), p(x1, y0, z1), p(x0, y0, z1)),  # bottom
        (p(x0, y1, z0), p(x0, y1, z1), p(x1, y1, z1), p(x1, y1, z0)),  # top
        (p(x0, y0, z0), p(x0, y1, z0), p(x1, y1, z0), p(x1, y0, z0)),  # front (z0)
        (p(x0, y0, z1), p(x1, y0, z1), p(x1, y1, z1), p(x0, y1, z1)),  # back (z1)
        (p(x0, y0, z0), p(x0, y0, z1), p(x0, y1, z1), p(x0, y1, z0)),  # left (x0)
        (p(x1, y0, z0), p(x1, y1, z0), p(x1, y1, z1), p(x1, y0, z1)),  # right (x1)
    ]
    out = []
    for f in faces:
        out += quad(*f)
    return out


def cornell34(width=256, height=256, spp=4, max_depth=5):
    tris, mats, lights = [], [], []
    materials = [  # 0 grey, 1 red, 2 green, 3 silver
        {"type": "diffuse", "rgb": GREY},
        {"type": "diffuse", "rgb": RED},
        {"type": "diffuse", "rgb": GREEN},
        {"type": "conductor", "metal": "Ag", "roughness": 0.0, "remap": True},
    ]

    def add(ts, m, light=-1):
        for t in ts:
            tris.append(t)
            mats.append(m)
            lights.append(light)

    # light (example :18-29), grey material, emitter 0
    add(quad((343.0, 548.7, 227.0), (343.0, 548.7, 332.0), (213.0, 548.7, 332.0), (213.0, 548.7, 227.0)), 0, 0)
    add(quad((552.8, 0.0, 0.0), (549.6, 0.0, 559.2), (556.0, 548.8, 559.2), (556.0, 548.8, 0.0)), 1)  # red
    add(quad((0.0, 0.0, 559.2), (0.0, 0.0, 0.0), (0.0, 548.8, 0.0), (0.0, 548.8, 559.2)), 2)  # green
    add(quad((552.8, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 559.2), (549.6, 0.0, 559.2)), 0)  # floor
    add(quad((556.0, 548.8, 0.0), (556.0, 548.8, 559.2), (0.0, 548.8, 559.2), (0.0, 548.8, 0.0)), 0)  # ceiling
    add(quad((549.6, 0.0, 559.2), (0.0, 0.0, 559.2), (0.0, 548.8, 559.2), (556.0, 548.8, 559.2)), 0)  # back
    add(box(130.0, 290.0, 0.0, 165.0, 65.0, 225.0), 0)  # short box, grey
    add(box(265.0, 430.0, 0.0, 330.0, 270.0, 435.0), 3)  # tall box, silver
    return {
        "eye": (278.0, 273.0, -800.0), "look": (278.0, 273.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 37.5,
        "width": width, "height": height, "spp": spp, "max_depth": max_depth,
        "tri": np.array(tris, np.float32).reshape(-1, 9),
        "tri_material": np.array(mats, np.uint32), "tri_light": np.array(lights, np.int32),
        "materials": materials, "lights": [{"rgb": (1.0, 1.0, 1.0), "scale": 20.0}],
    }


def _hash01(ix, iy, iz, seed):
    """Lattice hash -> [0, 1): a SplitMix64 finaliser over packed coordinates (uint64 numpy)."""
    with np.errstate(over="ignore"):
        k = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ (iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) \
            ^ (iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)) ^ np.uint64(seed)
        k ^= k >> np.uint64(30)
        k *= np.uint64(0xBF58476D1CE4E5B9)
        k ^= k >> np.uint64(27)
        k *= np.uint64(0x94D049BB133111EB)
        k ^= k >> np.uint64(31)
    return (k >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _value_noise(p, seed, freq=3.0):
    """Trilinear value noise with smoothstep weights, in [0, 1)."""
    q = p * freq
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    w = f * f * (3 - 2 * f)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                h = _hash01(i0[:, 0] + dx + (1 << 20), i0[:, 1] + dy + (1 << 20), i0[:, 2] + dz + (1 << 20), seed)
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out += wx * wy * wz * h
    return out


def uv_sphere(tess):
    """Unit-sphere vertices [tess+1][tess+1][3] and the triangle index list (2 tess^2)."""
    th = np.linspace(0, np.pi, tess + 1)
    ph = np.linspace(0, 2 * np.pi, tess + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    v = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], axis=-1)
    idx = []
    for i in range(tess):
        for j in range(tess):
            a, b = i * (tess + 1) + j, i * (tess + 1) + j + 1
            c, d = (i + 1) * (tess + 1) + j + 1, (i + 1) * (tess + 1) + j
            idx += [(a, b, c), (c, d, a)]
    return v.reshape(-1, 3), np.array(idx, np.int64)


def sphere_grid(width=1920, height=1080, n=32, tess=32, seed=0xC0FFEE, spp=64, max_depth=8):
    rng = Pcg32(seed)
    unit, idx = uv_sphere(tess)
    materials, tris, tri_mat = [], [], []
    for gi in range(n):
        for gj in range(n):
            c = np.array([(gi - (n - 1) / 2) * 1.0, 0.5, (gj - (n - 1) / 2) * 1.0])
            kind = min(int(rng.uniform() * 3), 2)
            if kind == 0:
                materials.append({"type": "diffuse", "rgb": tuple(0.1 + 0.8 * rng.uniform() for _ in range(3))})
            elif kind == 1:
                materials.append({"type": "conductor", "metal": "Cu", "roughness": 0.05 + 0.35 * rng.uniform(),
                                  "remap": True})
            else:
                materials.append({"type": "dielectric", "eta": 1.5, "roughness": 0.0, "remap": True})
            r = 0.4 + 0.1 * (_value_noise(unit + c, seed + gi * n + gj) - 0.5)
            verts = c + unit * r[:, None]
            tris.append(verts[idx].reshape(-1, 9))
            tri_mat.append(np.full(len(idx), len(materials) - 1, np.uint32))
    tri = np.concatenate(tris).astype(np.float32)
    tri_mat = np.concatenate(tri_mat)
    ground = len(materials)
    materials.append({"type": "diffuse", "rgb": (0.5, 0.5, 0.5)})
    g = n * 1.5
    gt = np.array(quad((-g, 0.0, -g), (-g, 0.0, g), (g, 0.0, g), (g, 0.0, -g)), np.float32)
    lights_t = []
    for li in range(8):
        for lj in range(4):
            x0, z0 = -12.0 + 3.0 * li, -8.0 + 5.0 * lj
            # wound so the geometric normal faces down (-y)
            lights_t += quad((x0, 5.0, z0), (x0 + 1.0, 5.0, z0), (x0 + 1.0, 5.0, z0 + 1.0), (x0, 5.0, z0 + 1.0))
    lt = np.array(lights_t, np.float32)
    tri = np.concatenate([tri, gt, lt])
    tri_mat = np.concatenate([tri_mat, np.full(2 + len(lt), ground, np.uint32)])
    tri_light = np.full(len(tri), -1, np.int32)
    tri_light[-len(lt):] = 0
    return {
        "eye": (0.0, 3.5, -21.0), "look": (0.0, 0.0, -4.0), "up": (0.0, 1.0, 0.0), "fov": 45.0,
        "width": width, "height": height, "spp": spp, "max_depth": max_depth,
        "tri": np.ascontiguousarray(tri), "tri_material": tri_mat, "tri_light": tri_light,
        "materials": materials, "lights": [{"rgb": (1.0, 1.0, 1.0), "scale": 10.0}],
    }


def soup_room(n_tris, width=256, height=256, seed=0x5EED, spp=4, max_depth=5):
    from . import gen_soup
    soup = gen_soup(n_tris, seed=seed).reshape(-1, 9)
    rng = Pcg32(seed ^ 0x9E3779B9)
    materials = [{"type": "diffuse", "rgb": tuple(0.1 + 0.8 * rng.uniform() for _ in range(3))} for _ in range(8)]
    grey, red, green = len(materials), len(materials) + 1, len(materials) + 2
    materials += [{"type": "diffuse", "rgb": GREY}, {"type": "diffuse", "rgb": RED}, {"type": "diffuse", "rgb": GREEN}]
    walls = box(-0.5, 1.5, -0.5, 1.5, -1.5, 1.5)
    wall_mat = [grey, grey, grey, grey, grey, grey, grey, grey, red, red, green, green]
    light = quad((0.25, 1.49, 0.25), (0.75, 1.49, 0.25), (0.75, 1.49, 0.75), (0.25, 1.49, 0.75))
    tri = np.concatenate([soup, np.array(walls, np.float32), np.array(light, np.float32)]).astype(np.float32)
    tri_mat = np.concatenate([np.arange(n_tris, dtype=np.uint32) % 8, np.array(wall_mat, np.uint32),
                              np.full(2, grey, np.uint32)])
    tri_light = np.full(len(tri), -1, np.int32)
    tri_light[-2:] = 0
    return {
        "eye": (0.5, 0.5, -1.3), "look": (0.5, 0.5, 0.5), "up": (0.0, 1.0, 0.0), "fov": 50.0,
        "width": width, "height": height, "spp": spp, "max_depth": max_depth,
        "tri": np.ascontiguousarray(tri), "tri_material": tri_mat, "tri_light": tri_light,
        "materials": materials, "lights": [{"rgb": (1.0, 1.0, 1.0), "scale": 8.0}],
    }


def batch_scene(n_tris, width, height, seed=0x5EED, spp=4, max_depth=5):
    from . import gen_soup
    soup = gen_soup(n_tris, seed=seed).reshape(-1, 9)
    light = quad((0.25, 1.5, 0.25), (0.75, 1.5, 0.25), (0.75, 1.5, 0.75), (0.25, 1.5, 0.75))
    tri = np.concatenate([soup, np.array(light, np.float32)]).astype(np.float32)
    tri_light = np.full(len(tri), -1, np.int32)
    tri_light[-2:] = 0
    return {
        "eye": (0.5, 0.5, -1.5), "look": (0.5, 0.5, 0.5), "up": (0.0, 1.0, 0.0), "fov": 45.0,
        "width": width, "height": height, "spp": spp, "max_depth": max_depth,
        "tri": np.ascontiguousarray(tri), "tri_material": np.zeros(len(tri), np.uint32), "tri_light": tri_light,
        "materials": [{"type": "diffuse", "rgb": (0.5, 0.5, 0.5)}], "lights": [{"rgb": (1.0, 1.0, 1.0), "scale": 8.0}],
    }


def batch_jobs(n_jobs=64, seed=0xBA7C, res=(64, 1024), tris=(10_000, 1_000_000)):
    """[(scene dict, spp, max_depth)] of the mixed batch."""
    rng = Pcg32(seed)
    out = []
    span = res[1] - res[0]
    for k in range(n_jobs):
        w = res[0] + min(int(rng.uniform() * (span + 1)), span)
        h = res[0] + min(int(rng.uniform() * (span + 1)), span)
        spp = (1, 4, 16, 64)[min(int(rng.uniform() * 4), 3)]
        depth = 1 + min(int(rng.uniform() * 8), 7)
        lo, hi = np.log(tris[0]), np.log(tris[1])
        n = int(round(float(np.exp(lo + (hi - lo) * rng.uniform()))))
        out.append((batch_scene(n, w, h, seed=seed + 17 * k + 1, spp=spp, max_depth=depth), spp, depth))
    return out


def grad_room(n_tris, width=512, height=512, seed=0xD1FF, spp=16, max_depth=4):
    from . import gen_soup
    soup = gen_soup(n_tris, seed=seed).reshape(-1, 9)
    walls = np.array(box(-0.5, 1.5, -0.5, 1.5, -1.5, 1.5), np.float32)
    lights = []
    for i in range(4):
        for j in range(2):
            x0, z0 = 0.05 + 0.25 * i, 0.1 + 0.45 * j
            lights += quad((x0, 1.49, z0), (x0 + 0.2, 1.49, z0), (x0 + 0.2, 1.49, z0 + 0.35), (x0, 1.49, z0 + 0.35))
    lt = np.array(lights, np.float32)
    tri = np.ascontiguousarray(np.concatenate([soup, walls, lt]).astype(np.float32))
    n = len(tri)
    k = np.arange(n, dtype=np.uint64)
    alb = np.stack([0.1 + 0.8 * _hash01(k, np.full(n, c, np.uint64), np.zeros(n, np.uint64), seed) for c in range(3)], 1)
    rng = Pcg32(seed)
    tri_light = np.full(n, -1, np.int32)
    tri_light[-len(lt):] = np.repeat(np.arange(8, dtype=np.int32), 2)
    return {
        "eye": (0.5, 0.5, -1.3), "look": (0.5, 0.5, 0.5), "up": (0.0, 1.0, 0.0), "fov": 50.0,
        "width": width, "height": height, "spp": spp, "max_depth": max_depth,
        "tri": tri, "tri_material": np.arange(n, dtype=np.uint32), "tri_light": tri_light,
        "albedo": alb.astype(np.float32),
        "materials": None,
        "lights": [{"rgb": (1.0, 1.0, 1.0), "scale": 4.0 + 4.0 * rng.uniform()} for _ in range(8)],
    }


def perturbed(sc, seed=0x7A27, amount=0.1):
    """The scene with every albedo and emitter scale multiplied by 1 + U[-amount, amount]."""
    n = len(sc["albedo"])
    k = np.arange(n, dtype=np.uint64)
    f = np.stack([1.0 + amount * (2.0 * _hash01(k, np.full(n, c, np.uint64), np.ones(n, np.uint64), seed) - 1.0)
                  for c in range(3)], 1)
    rng = Pcg32(seed)
    out = dict(sc)
    out["albedo"] = np.clip(sc["albedo"] * f, 0.01, 0.99).astype(np.float32)
    out["lights"] = [dict(l, scale=l["scale"] * (1.0 + amount * (2.0 * rng.uniform() - 1.0))) for l in sc["lights"]]
    return out


def _translate(x, y, z):
    m, mi = np.eye(4, dtype=np.float32), np.eye(4, dtype=np.float32)
    m[:3, 3] = (x, y, z)
    mi[:3, 3] = (-np.float32(x), -np.float32(y), -np.float32(z))
    return m, mi


def _scale(x, y, z):
    m = np.diag([x, y, z, 1.0]).astype(np.float32)
    mi = np.diag([np.float32(1.0) / np.float32(x), np.float32(1.0) / np.float32(y), np.float32(1.0) / np.float32(z),
                  1.0]).astype(np.float32)
    return m, mi


def _rotate(deg, ax, ay, az):
    """Transform::rotate (transform.h:61-88): sin and cos rounded from double; the
    inverse is the transpose."""
    rad = np.float32(np.float64(np.float32(np.pi)) / 180.0 * np.float64(np.float32(deg)))
    s, c = np.float32(np.sin(np.float64(rad))), np.float32(np.cos(np.float64(rad)))
    a = np.array([ax, ay, az], np.float32)
    a = a / np.sqrt(np.float32(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]))
    x, y, z = (np.float64(v) for v in a)
    s, c = np.float64(s), np.float64(c)
    m = np.eye(4, dtype=np.float32)
    m[0, :3] = (x * x + (1.0 - x * x) * c, x * y * (1.0 - c) - z * s, x * z * (1.0 - c) + y * s)
    m[1, :3] = (x * y * (1.0 - c) + z * s, y * y + (1.0 - y * y) * c, y * z * (1.0 - c) - x * s)
    m[2, :3] = (x * z * (1.0 - c) - y * s, y * z * (1.0 - c) + x * s, z * z + (1.0 - z * z) * c)
    return m, np.ascontiguousarray(m.T)


def _fma_mul(a, b):
    """SquareMatrix<4>::operator* (squared_matrix.h:98-109): an FMA chain per entry."""
    r = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            acc = np.float32(0)
            for k in range(4):
                acc = np.float32(np.float64(a[i, k]) * np.float64(b[k, j]) + np.float64(acc))
            r[i, j] = acc
    return r


def _compose(*xfs):
    """CTM = x0 * x1 * ...; its inverse composed in reverse order, as Transform does."""
    m, mi = xfs[0]
    for x, xi in xfs[1:]:
        m, mi = _fma_mul(m, x), _fma_mul(xi, mi)
    return m, mi


def shapes_room(width=64, height=64, spp=4, max_depth=5):
    sc = cornell34(width, height, spp=spp, max_depth=max_depth)
    tri = [t for t in np.asarray(sc["tri"], np.float32)]
    mats = list(np.asarray(sc["tri_material"], np.uint32))
    lig = list(np.asarray(sc["tri_light"], np.int32))
    materials = list(sc["materials"]) + [
        {"type": "dielectric", "eta": 1.5, "roughness": 0.0, "remap": False},  # 4 glass
        {"type": "conductor", "metal": "Cu", "roughness": 0.05, "remap": True},  # 5 rough copper
        {"type": "diffuse", "rgb": (0.2, 0.35, 0.8)},  # 6 blue
    ]
    lights = list(sc["lights"]) + [{"rgb": (1.0, 0.6, 0.3), "scale": 8.0, "two_sided": False},
                                   {"rgb": (0.4, 0.7, 1.0), "scale": 5.0, "two_sided": False}]
    n0 = len(tri)
    rev = [0] * n0
    xforms, shapes = [], []
    tri_xform = [0xFFFFFFFF] * n0
    tri_shape = [0xFFFFFFFF] * n0

    def add_xf(x):
        xforms.append(x)
        return len(xforms) - 1

    def add_shape(q, mat, light=-1, reverse=0):
        shapes.append(q)
        tri.append(np.zeros(9, np.float32))
        mats.append(mat)
        lig.append(light)
        rev.append(reverse)
        tri_xform.append(q.get("xform") if q.get("xform") is not None else 0xFFFFFFFF)
        tri_shape.append(len(shapes) - 1)

    add_shape({"type": "sphere", "radius": 75.0, "xform": add_xf(_translate(400.0, 80.0, 150.0))}, 4)
    add_shape({"type": "sphere", "radius": 60.0, "zmin": -35.0, "zmax": 50.0, "phimax": 290.0,
               "xform": add_xf(_compose(_translate(430.0, 330.0 + 60.0, 350.0), _rotate(-70.0, 1.0, 0.3, 0.0)))}, 5)
    add_shape({"type": "sphere", "radius": 40.0, "xform": add_xf(_translate(70.0, 45.0, 120.0))}, 6, reverse=1)
    add_shape({"type": "sphere", "radius": 18.0, "xform": add_xf(_translate(90.0, 430.0, 300.0))}, 0, light=1)
    add_shape({"type": "disk", "radius": 45.0,
               "xform": add_xf(_compose(_translate(420.0, 548.0, 110.0), _rotate(90.0, 1.0, 0.0, 0.0)))}, 0, light=2)
    add_shape({"type": "disk", "radius": 70.0, "innerradius": 25.0, "phimax": 300.0, "height": 2.0,
               "xform": add_xf(_compose(_translate(300.0, 520.0, 480.0), _scale(1.0, -1.0, 1.0),
                                        _rotate(90.0, 1.0, 0.0, 0.0)))}, 6)
    # the short box (cornell34 entries 22..33) instanced twice
    boxes = list(range(22, 34))
    for xf in (_compose(_translate(80.0, 0.0, 300.0), _rotate(25.0, 0.0, 1.0, 0.0), _scale(0.5, 0.8, 0.5)),
               _compose(_translate(330.0, 220.0, 60.0), _rotate(-15.0, 0.0, 0.0, 1.0), _scale(0.35, 0.3, 0.35))):
        shapes.append({"type": "instance", "xform": add_xf(xf)})
        for b in boxes:
            tri.append(np.asarray(sc["tri"], np.float32)[b])
            mats.append(6)
            lig.append(-1)
            rev.append(0)
            tri_xform.append(0xFFFFFFFF)
            tri_shape.append(len(shapes) - 1)
    sc.update(
        tri=np.asarray(tri, np.float32).reshape(-1, 9), tri_material=np.asarray(mats, np.uint32),
        tri_light=np.asarray(lig, np.int32), tri_reverse=np.asarray(rev, np.uint8), materials=materials,
        lights=lights, xforms=np.stack([m for m, _ in xforms]), xforms_inv=np.stack([mi for _, mi in xforms]),
        tri_xform=np.asarray(tri_xform, np.uint32), shapes=shapes, tri_shape=np.asarray(tri_shape, np.uint32))
    return sc
