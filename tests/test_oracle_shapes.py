"""Pins of the oracle's analytic shapes and instancing to the mathematics.

The oracle restates Sphere, Disk and TransformedPrimitive from the reference
(shapes/sphere.{h,cu}, shapes/disk.{h,cu}, primitives/transformed_primitive.cu)
with the reference's interval arithmetic (util/interval.h) and Transform
operations; the GPU kernels are held to it bit for bit elsewhere. These tests
check the restatement against facts that do not depend on reading the
reference again:

* the directed-rounding primitives against exact rational arithmetic;
* ray/sphere and ray/disk hits against the FP64 quadratic and plane solutions,
  including partial spheres (zmin, zmax, phimax) and transformed shapes;
* hit points on the surface and outward normals;
* cone sampling of a sphere: pdf = 1 / solid angle, points on the visible cap,
  directions uniform in the cone; area sampling from inside: the pdf integrates
  to 1 over directions; the disk's pdf integrates to 1 and equals its sampling
  density; sample() and pdf() agree;
* whole renders: a Lambertian plane under an emissive sphere or disk shows
  albedo x Le x the form factor (next-event estimation with MIS through the
  shape's sampling and pdf);
* an instanced mesh renders as the same mesh placed by hand (TransformedPrimitive
  only shifts ray origins by the transform's error bounds).

A ray whose line passes through the sphere's centre (to within the interval
widths) gets a NaN distance in the reference: Vector3fi::length takes the square
root of x*x + y*y + z*z, whose lower bound is negative there, and the interval
square root does not clamp it. The oracle keeps that; the random rays below pass
the centre at distances many orders of magnitude larger.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(0x5AE7)
# The scenes' coordinate scale: FP32 distances carry errors of a few ulps of the
# ray origin's coordinates, so t is compared to 2e-5 of (t + SCALE).
SCALE = 4.0


def f32(x):
    return float(np.float32(x))


def next_up(x):
    return float(np.nextafter(np.float32(x), np.float32(np.inf)))


def next_down(x):
    return float(np.nextafter(np.float32(x), np.float32(-np.inf)))


# ------------------------------------------------------------ directed rounding
def _floats(n):
    mant = RNG.uniform(1, 2, n) * RNG.choice([-1, 1], n)
    exp = RNG.integers(-30, 30, n)
    return [f32(m * 2.0 ** e) for m, e in zip(mant, exp)]


def _check_bracket(lo, hi, exact):
    assert Fraction(lo) <= exact <= Fraction(hi)
    if Fraction(lo) != exact:
        assert Fraction(next_up(lo)) > exact  # lo is the largest float below
        assert Fraction(next_down(hi)) < exact  # hi the smallest float above
    else:
        assert lo == hi


def test_directed_rounding_is_exact():
    a, b = _floats(400), _floats(400)
    for x, y in zip(a, b):
        _check_bracket(*O.directed("mul", x, y), Fraction(x) * Fraction(y))
        _check_bracket(*O.directed("div", x, y), Fraction(x) / Fraction(y))
        _check_bracket(*O.directed("add", x, y), Fraction(x) + Fraction(y))
        ax = abs(x)
        lo, hi = O.directed("sqrt", ax)
        assert Fraction(lo) ** 2 <= Fraction(ax) <= Fraction(hi) ** 2
        if lo != hi:
            assert Fraction(next_up(lo)) ** 2 > Fraction(ax) and Fraction(next_down(hi)) ** 2 < Fraction(ax)
    # exact cases are not widened
    assert O.directed("mul", 3.0, 0.5) == (1.5, 1.5)
    assert O.directed("div", 1.0, 4.0) == (0.25, 0.25)
    assert O.directed("sqrt", 9.0) == (3.0, 3.0)


# ------------------------------------------------------------ transforms
def rot(axis, deg):
    axis = np.asarray(axis, np.float64)
    axis /= np.linalg.norm(axis)
    t = np.radians(deg)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K
    m = np.eye(4)
    m[:3, :3] = R
    return m


def place(center, scale=1.0, axis=(0.3, 1.0, 0.2), deg=37.0):
    m = rot(axis, deg)
    m[:3, :3] *= scale
    m[:3, 3] = center
    return m.astype(np.float32)


def rays_towards(origins, targets, t_max=np.inf):
    d = targets - origins
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((len(origins), 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7] = origins, t_max, d
    return r


def obj_space(m, rays):
    """Rays into the shape's object space in FP64."""
    inv = np.linalg.inv(m.astype(np.float64))
    o = rays[:, :3].astype(np.float64) @ inv[:3, :3].T + inv[:3, 3]
    d = rays[:, 4:7].astype(np.float64) @ inv[:3, :3].T
    return o, d


def sphere_fp64(m, rays, r, zmin=None, zmax=None, phimax=360.0):
    """Nearest root t > 0 of |o + t d|^2 = r^2 in object space whose point passes the
    clipping (z within [zmin, zmax], phi within [0, phimax]), else the far root."""
    o, d = obj_space(m, rays)
    zmin = -r if zmin is None else zmin
    zmax = r if zmax is None else zmax
    a = (d * d).sum(1)
    b = 2 * (d * o).sum(1)
    c = (o * o).sum(1) - r * r
    disc = b * b - 4 * a * c
    t = np.full(len(rays), np.nan)
    margin = np.zeros(len(rays))
    for i in range(len(rays)):
        if disc[i] < 0:
            margin[i] = -disc[i] / (b[i] * b[i] + 1e-300)
            continue
        sq = np.sqrt(disc[i])
        for root in sorted([(-b[i] - sq) / (2 * a[i]), (-b[i] + sq) / (2 * a[i])]):
            if root <= 0 or root > rays[i, 3]:
                continue
            p = o[i] + root * d[i]
            phi = np.arctan2(p[1], p[0]) % (2 * np.pi)
            if (zmin > -r and p[2] < zmin) or (zmax < r and p[2] > zmax) or np.degrees(phi) > phimax:
                continue
            t[i] = root
            break
        margin[i] = disc[i] / (b[i] * b[i] + 1e-300)
    return t, margin


def test_sphere_hits_match_fp64_quadratic():
    r = 1.3
    m = place((0.4, -0.2, 2.0), scale=0.8)
    n = 3000
    org = RNG.uniform(-4, 4, (n, 3))
    tgt = m[:3, 3] + RNG.uniform(-1.6, 1.6, (n, 3))
    rays = rays_towards(org, tgt)
    hit, t, p, nn = O.shape_intersect("sphere", m, rays, radius=r)
    want, margin = sphere_fp64(m, rays, r)
    clear = margin > 1e-3  # away from tangency
    assert clear.mean() > 0.9
    assert np.array_equal(hit[clear], ~np.isnan(want[clear]))
    h = hit & clear
    assert h.sum() > 500
    assert np.all(np.abs(t[h] - want[h]) <= 2e-5 * (want[h] + SCALE))
    # on the surface (radius r scaled by 0.8 in render space), outward normal
    c = m[:3, 3].astype(np.float64)
    dist = np.linalg.norm(p[h] - c, axis=1)
    assert np.allclose(dist, 0.8 * r, rtol=2e-5)
    assert np.allclose(nn[h], (p[h] - c) / dist[:, None], atol=2e-4)
    # ReverseOrientation flips the normal
    _, _, _, nr = O.shape_intersect("sphere", m, rays, reverse=True, radius=r)
    assert np.allclose(nr[h], -nn[h])


@pytest.mark.parametrize("zmin,zmax,phimax", [(-0.5, 0.9, 360.0), (-1.3, 1.3, 250.0), (0.2, 1.0, 120.0)])
def test_partial_sphere_clipping(zmin, zmax, phimax):
    r = 1.3
    m = place((0.0, 0.5, 1.0), axis=(1, 0.2, 0), deg=-60.0)
    n = 3000
    org = RNG.uniform(-4, 4, (n, 3))
    rays = rays_towards(org, m[:3, 3] + RNG.uniform(-1.4, 1.4, (n, 3)))
    # rays starting inside the sphere see the clipped far side too
    rays[: n // 4, :3] = m[:3, 3] + RNG.uniform(-0.4, 0.4, (n // 4, 3))
    hit, t, p, _ = O.shape_intersect("sphere", m, rays, radius=r, zmin=zmin, zmax=zmax, phimax=phimax)
    want, margin = sphere_fp64(m, rays, r, zmin, zmax, phimax)
    # away from tangency and from the clipping boundaries
    o, d = obj_space(m, rays)
    pw = o + np.nan_to_num(want)[:, None] * d
    phi = np.degrees(np.arctan2(pw[:, 1], pw[:, 0]) % (2 * np.pi))
    far = (np.abs(pw[:, 2] - zmin) > 1e-3) & (np.abs(pw[:, 2] - zmax) > 1e-3) & (np.abs(phi - phimax) > 1e-2)
    clear = (margin > 1e-3) & (far | np.isnan(want))
    assert np.array_equal(hit[clear], ~np.isnan(want[clear]))
    h = hit & clear
    assert h.sum() > 300
    assert np.all(np.abs(t[h] - want[h]) <= 2e-5 * (want[h] + SCALE))


@pytest.mark.parametrize("inner,phimax", [(0.0, 360.0), (0.4, 360.0), (0.3, 200.0)])
def test_disk_hits_match_plane(inner, phimax):
    R, height = 1.5, 0.25
    m = place((0.2, 0.3, 1.5), axis=(0, 1, 0.4), deg=50.0)
    n = 3000
    org = RNG.uniform(-4, 4, (n, 3))
    rays = rays_towards(org, m[:3, 3] + RNG.uniform(-1.6, 1.6, (n, 3)))
    hit, t, p, nn = O.shape_intersect("disk", m, rays, radius=R, height=height, innerradius=inner, phimax=phimax)
    o, d = obj_space(m, rays)
    tw = (height - o[:, 2]) / d[:, 2]
    q = o + tw[:, None] * d
    rr = np.hypot(q[:, 0], q[:, 1])
    phi = np.degrees(np.arctan2(q[:, 1], q[:, 0]) % (2 * np.pi))
    want = (tw > 0) & (rr <= R) & (rr >= inner) & (phi <= phimax)
    clear = (np.abs(rr - R) > 1e-4) & (np.abs(rr - inner) > 1e-4) & (np.abs(phi - phimax) > 1e-3)
    assert np.array_equal(hit[clear], want[clear])
    h = hit & clear
    assert h.sum() > 200
    assert np.all(np.abs(t[h] - tw[h]) <= 2e-5 * (tw[h] + SCALE))
    nz = (m[:3, :3].astype(np.float64) @ np.array([0, 0, 1.0]))
    nz /= np.linalg.norm(nz)
    assert np.allclose(np.abs(nn[h] @ nz), 1.0, atol=1e-5)


# ------------------------------------------------------------ sampling
def _ctx(points):
    c = np.zeros((len(points), 9), np.float32)
    c[:, :3] = points
    return c


def test_sphere_cone_sampling():
    r = 0.7
    m = place((0.0, 0.0, 0.0))
    n = 20000
    for dist in (1.5, 4.0, 60.0):  # 60: below sin^2(1.5 deg), the small-angle branch
        ref = np.tile(np.array([dist, 0.3 * dist / 4, -0.2 * dist / 4], np.float32), (n, 1))
        D = np.linalg.norm(ref[0].astype(np.float64))
        u = RNG.random((n, 2)).astype(np.float32)
        ok, p, nn, pdf = O.shape_sample("sphere", m, _ctx(ref), u, radius=r)
        assert ok.all()
        cos_max = np.sqrt(1 - (r / D) ** 2)
        assert np.allclose(pdf, 1 / (2 * np.pi * (1 - cos_max)), rtol=3e-4)
        assert np.allclose(np.linalg.norm(p, axis=1), r, rtol=1e-5)  # on the surface
        assert np.allclose(nn, p / r, atol=1e-4)  # outward normal
        w = p - ref
        w /= np.linalg.norm(w, axis=1, keepdims=True)
        toc = -ref[0] / D
        cosw = w @ toc
        assert (cosw >= cos_max - 1e-5).all()  # inside the cone
        # the visible cap: facing the reference point
        assert ((p @ (ref[0] / D)) >= -1e-4).all()
        # uniform in the cone: a fraction (1 - cos a) / (1 - cos max) within angle a
        a_cos = 1 - 0.5 * (1 - cos_max)
        frac = (cosw >= a_cos).mean()
        assert abs(frac - 0.5) < 4 * np.sqrt(0.25 / n)
        # pdf() at the sampled directions equals the sampling density
        assert np.allclose(O.shape_pdf("sphere", m, _ctx(ref), w, radius=r), pdf, rtol=1e-5)


def _uniform_dirs(n):
    v = RNG.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def test_sphere_pdf_integrates_to_one():
    """(Sphere sampling and pdf use the object-space radius with render-space
    distances, as pbrt-v4's do: correct for transforms without scale.)"""
    r = 0.9
    m = place((0.1, -0.3, 0.2))
    n = 200000
    w = _uniform_dirs(n)
    # outside: pdf is 1 / solid angle over the cone of directions to the sphere
    ref = np.tile(np.array([2.5, 0.4, -0.7], np.float32), (n, 1))
    pdf = O.shape_pdf("sphere", m, _ctx(ref), w, radius=r).astype(np.float64)
    c = m[:3, 3].astype(np.float64) - ref[0]
    sin_max = r / np.linalg.norm(c)
    in_cone = (w @ (c / np.linalg.norm(c))) >= np.sqrt(1 - sin_max ** 2)
    est = 4 * np.pi * np.mean(pdf * in_cone)
    assert abs(est - 1) < 5 * 4 * np.pi * np.std(pdf * in_cone) / np.sqrt(n)
    # inside: area sampling, the pdf over all directions integrates to 1
    ref = np.tile(m[:3, 3] + np.array([0.2, 0.1, -0.3], np.float32), (n, 1))
    pdf = O.shape_pdf("sphere", m, _ctx(ref), w, radius=r).astype(np.float64)
    assert (pdf > 0).all()
    est = 4 * np.pi * np.mean(pdf)
    assert abs(est - 1) < 5 * 4 * np.pi * np.std(pdf) / np.sqrt(n)
    # and sampling from inside draws points on the surface with the pdf() density
    u = RNG.random((2000, 2)).astype(np.float32)
    ok, p, _, spdf = O.shape_sample("sphere", m, _ctx(ref[:2000]), u, radius=r)
    assert ok.all()
    assert np.allclose(np.linalg.norm(p - m[:3, 3], axis=1), r, rtol=1e-5)
    wi = p - ref[:2000]
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    back = O.shape_pdf("sphere", m, _ctx(ref[:2000]), wi, radius=r)
    assert np.allclose(back, spdf, rtol=2e-3)


def test_disk_pdf_and_sampling():
    R = 1.2
    m = place((0.0, 2.0, 0.0), axis=(1, 0, 0), deg=90.0)  # normal towards -y
    n = 200000
    ref = np.tile(np.array([0.3, 0.0, -0.4], np.float32), (n, 1))
    w = _uniform_dirs(n)
    pdf = O.shape_pdf("disk", m, _ctx(ref), w, radius=R, innerradius=0.3).astype(np.float64)
    est = 4 * np.pi * np.mean(pdf)
    assert abs(est - 1) < 5 * 4 * np.pi * np.std(pdf) / np.sqrt(n)
    u = RNG.random((4000, 2)).astype(np.float32)
    ok, p, nn, spdf = O.shape_sample("disk", m, _ctx(ref[:4000]), u, radius=R, innerradius=0.3)
    assert ok.all()
    assert np.allclose(p[:, 1], 2.0, atol=1e-5)
    rr = np.hypot(p[:, 0], p[:, 2])
    assert (rr <= R + 1e-5).all()
    assert np.allclose(nn, [0, -1, 0], atol=1e-6)
    # area density 1 / (pi R^2) (the sampler covers the whole disk; innerradius only
    # enters the area), in solid angle: d^2 / (A cos)
    d = p - ref[:4000]
    dist2 = (d * d).sum(1)
    cos = np.abs(d[:, 1]) / np.sqrt(dist2)
    area = 0.5 * 2 * np.pi * (R * R - 0.3 * 0.3)
    assert np.allclose(spdf, dist2 / (area * cos), rtol=1e-4)
    # uniform over the disk: half of the samples within R / sqrt(2)
    assert abs((rr <= R / np.sqrt(2)).mean() - 0.5) < 4 * np.sqrt(0.25 / 4000)
    wi = d / np.sqrt(dist2)[:, None]
    back = O.shape_pdf("disk", m, _ctx(ref[:4000]), wi, radius=R, innerradius=0.3)
    inside = rr >= 0.3 + 1e-4  # the pdf intersects the real (annular) disk
    assert np.allclose(back[inside], spdf[inside], rtol=2e-3)
    assert (back[rr < 0.3 - 1e-4] == 0).all()


# ------------------------------------------------------------ whole renders
def _xforms(*ms):
    xf = np.stack([np.asarray(m, np.float32) for m in ms])
    inv = np.stack([np.linalg.inv(m.astype(np.float64)).astype(np.float32) for m in xf])
    return xf, inv


def _plane_and_shape(shape, xform, cam_h, fov, look_up=False, depth=1, spp=64, res=16, albedo=0.5):
    P = [(-50, 0, -50), (50, 0, -50), (50, 0, 50), (-50, 0, 50)]
    tri = np.asarray([P[0], P[1], P[2], P[2], P[3], P[0], (0, 0, 0), (0, 0, 0), (0, 0, 0)],
                     np.float32).reshape(3, 3, 3)
    xf, inv = _xforms(xform)
    look = (0.0, 100.0, 0.0) if look_up else (0.0, 0.0, 0.0)
    return dict(eye=(0.0, cam_h, 0.0), look=look, up=(0, 0, 1), fov=fov, width=res, height=res, spp=spp,
                max_depth=depth, tri=tri, tri_material=np.array([0, 0, 1], np.uint32),
                tri_light=np.array([-1, -1, 0], np.int32), albedo=np.array([[albedo] * 3, [0.0] * 3], np.float32),
                lights=[dict(rgb=(1.0, 0.9, 0.7), scale=2.0, two_sided=False)], xforms=xf, xforms_inv=inv,
                tri_xform=np.array([0xFFFFFFFF, 0xFFFFFFFF, 0], np.uint32), shapes=[dict(shape, xform=0)],
                tri_shape=np.array([0xFFFFFFFF, 0xFFFFFFFF, 0], np.uint32))


def _sensor_mean(sc):
    _, aux, _ = O.render(sc, threads=8)
    aux = aux.astype(np.float64).reshape(-1, 4)
    return aux[:, :3].sum(0) / aux[:, 3].sum(), aux


@pytest.mark.parametrize("kind", ["sphere", "disk"])
def test_plane_under_emitter_shape(kind):
    """A Lambertian plane y = 0 under an emissive sphere (radius r, centre at height
    h: form factor (r / h)^2 at the point below) or a downward disk (radius R at
    height h: R^2 / (R^2 + h^2)). Le is read off a render looking straight at the
    emitter; NEE samples the sphere's cone or the disk's area, BSDF samples hit the
    shape, and the power-heuristic weights sum to one."""
    h = 2.0
    if kind == "sphere":
        r = 0.8
        shape, xform, F = dict(type="sphere", radius=r), place((0, h, 0), axis=(0, 1, 0), deg=0.0), (r / h) ** 2
    else:
        R = 1.1
        shape, xform = dict(type="disk", radius=R), place((0, h, 0), axis=(1, 0, 0), deg=90.0)
        F = R * R / (R * R + h * h)
    Le, _ = _sensor_mean(_plane_and_shape(shape, xform, 0.5, 3.0, look_up=True, spp=16))
    got, aux = _sensor_mean(_plane_and_shape(shape, xform, 0.2, 0.5, spp=256, res=12))
    want = 0.5 * Le * F
    px = aux[:, :3] / aux[:, 3:4]
    se = px.std(0) / np.sqrt(len(px))
    assert np.all(np.abs(got - want) < 5 * se + 5e-3 * want), (kind, got, want, se)


def test_instance_renders_as_the_placed_mesh():
    """Two boxes instanced with ObjectInstance transforms render as the same boxes
    placed by their transforms: equal within the noise of differing sample paths
    (TransformedPrimitive shifts the ray origin by the transform's error bound)."""
    import scenes

    sc = scenes.cornell34(32, 32, spp=16, max_depth=3)
    tri = np.asarray(sc["tri"], np.float32).reshape(-1, 9)
    boxes = np.nonzero(np.asarray(sc["tri_light"]) < 0)[0][-24:]  # the two boxes
    T = place((15.0, 0.0, -10.0), axis=(0, 1, 0), deg=12.0).astype(np.float64)
    moved = tri.copy()
    for i in boxes:
        p = moved[i].reshape(3, 3).astype(np.float64)
        moved[i] = (p @ T[:3, :3].T + T[:3, 3]).astype(np.float32).reshape(9)
    placed = dict(sc, tri=moved)
    inst = dict(sc)
    xf, inv = _xforms(T.astype(np.float32))
    inst.update(xforms=xf, xforms_inv=inv, shapes=[dict(type="instance", xform=0)],
                tri_shape=np.where(np.isin(np.arange(len(tri)), boxes), 0, 0xFFFFFFFF).astype(np.uint32))
    a, _, _ = O.render(placed, threads=8)
    b, _, _ = O.render(inst, threads=8)
    assert np.isfinite(b).all()
    diff = np.abs(a.astype(np.float64) - b).mean()
    assert diff < 0.02 * np.abs(a).mean(), diff
    # an identity instance is the mesh itself, bit for bit
    ident = dict(sc, xforms=np.eye(4, dtype=np.float32)[None], xforms_inv=np.eye(4, dtype=np.float32)[None],
                 shapes=[dict(type="instance", xform=0)],
                 tri_shape=np.where(np.isin(np.arange(len(tri)), boxes), 0, 0xFFFFFFFF).astype(np.uint32))
    c, _, _ = O.render(sc, threads=8)
    d, _, _ = O.render(ident, threads=8)
    assert np.array_equal(c.view(np.uint32), d.view(np.uint32))
