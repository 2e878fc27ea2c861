"""GPU parity of the wavefront path tracer through the C ABI against the CPU oracle.

Both sides render the same scene from the same (pixel, sample) sampler streams;
the comparison is on the float film (RGBFilm::get_pixel_rgb) and on the raw
sensor-RGB and filter-weight sums. The arithmetic is the same FP32 sequence on
both sides (transcendentals in double, rounded once), so films are expected to
be bit-identical; the bar the test enforces is the 1e-4 relative per-pixel
tolerance, and the count of bit-identical pixels is printed.
"""
import numpy as np
import pytest

import oracle as O
import pbrtm
import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def compare(gpu, ref, label):
    a, b = gpu.reshape(-1), ref.reshape(-1)
    exact = np.mean(a.view(np.uint32) == b.view(np.uint32))
    scale = np.maximum(np.abs(a), np.abs(b))
    floor = 1e-6 * max(1e-30, float(np.mean(np.abs(b))))
    rel = np.abs(a - b) / np.maximum(scale, floor)
    log(f"{label}: bit-identical {exact:.6f}, max rel {float(rel.max()):.3g}, nan gpu {np.isnan(a).sum()} ref {np.isnan(b).sum()}")
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    assert float(rel[ok].max(initial=0.0)) <= 1e-4, label
    return exact


def render_both(ctx, sc, spp, depth, pool=0):
    log(f"render {sc['width']}x{sc['height']} spp {spp} depth {depth} pool {pool}")
    s = ctx.scene(sc)
    try:
        rgb, aux, st = s.render(spp=spp, max_depth=depth, pool_size=pool)
    finally:
        s.free()
    osc = dict(sc, spp=spp, max_depth=depth)
    orgb, oaux, ost = O.render(osc, pool_size=pool if pool else 2 * 1024 * 1024, threads=16)
    log(f"  gpu {st['ms_total']:.2f} ms rounds {st['rounds']} rays {st['rays']} shadow {st['shadow_rays']}; "
        f"oracle rounds {ost['rounds']} rays {ost['rays']} shadow {ost['shadow_rays']}")
    assert st["rounds"] == ost["rounds"]
    assert st["rays"] == ost["rays"]
    assert st["shadow_rays"] == ost["shadow_rays"]
    return rgb, aux, orgb, oaux


def test_cornell_small(ctx):
    sc = scenes.cornell34(64, 64)
    rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 5)
    compare(aux, oaux, "aux 64x64")
    compare(rgb, orgb, "film 64x64")
    assert np.isfinite(rgb).all()


def test_cornell_config1(ctx):
    """256x256, 4 spp stratified, max depth 5: the whole image in one pool."""
    sc = scenes.cornell34(256, 256)
    rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 5)
    compare(aux, oaux, "aux 256x256")
    exact = compare(rgb, orgb, "film 256x256")
    assert exact > 0.99


def test_pool_refills_non_square(ctx):
    """A pool far smaller than the path count: slots are recycled over many rounds,
    so film order (round, then sample) and new-path scheduling are exercised."""
    sc = scenes.cornell34(48, 32)
    rgb, aux, orgb, oaux = render_both(ctx, sc, 9, 3, pool=1000)
    compare(aux, oaux, "aux pool 1000")
    compare(rgb, orgb, "film pool 1000")


def test_russian_roulette_depth(ctx):
    """maxdepth 12 reaches the Russian-roulette branch (path_length > 8)."""
    sc = scenes.cornell34(40, 40)
    rgb, aux, orgb, oaux = render_both(ctx, sc, 1, 12)
    compare(aux, oaux, "aux depth 12")
    compare(rgb, orgb, "film depth 12")


def test_depth_zero_and_one(ctx):
    sc = scenes.cornell34(32, 24)
    for depth in (0, 1):
        rgb, aux, orgb, oaux = render_both(ctx, sc, 4, depth)
        compare(rgb, orgb, f"film depth {depth}")


def test_two_sided_reversed_light(ctx):
    """A two-sided emitter and a reversed-orientation wall."""
    sc = scenes.cornell34(32, 32)
    sc["lights"] = [{"rgb": (0.9, 0.6, 0.3), "scale": 15.0, "two_sided": True}]
    rev = np.zeros(len(sc["tri"]), np.uint8)
    rev[2:4] = 1
    sc["tri_reverse"] = rev
    rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 4)
    compare(rgb, orgb, "film two-sided")


def test_render_errors(ctx):
    sc = scenes.cornell34(16, 16)
    s = ctx.scene(sc)
    try:
        with pytest.raises(pbrtm.PbrtmError):
            s.render(spp=3)
    finally:
        s.free()
    bad = dict(sc, materials=[{"type": "diffuse", "rgb": (1.5, 0.2, 0.2)}] + sc["materials"][1:])
    with pytest.raises(pbrtm.PbrtmError):
        ctx.scene(bad)


def cornell_materials(variant):
    """The Cornell scene with the boxes' materials swapped for rough and refractive ones."""
    sc = scenes.cornell34(48, 48)
    mats = list(sc["materials"])
    if variant == "rough":
        mats[3] = {"type": "conductor", "metal": "Cu", "roughness": 0.2, "remap": True}
        mats[0] = {"type": "dielectric", "eta": 1.5, "roughness": 0.1, "remap": True}
    elif variant == "glass":
        mats[3] = {"type": "dielectric", "eta": "glass-BK7", "roughness": 0.0}
        mats[0] = {"type": "conductor", "metal": "Au", "uroughness": 0.05, "vroughness": 0.3, "remap": False}
    elif variant == "smooth_glass":
        mats[3] = {"type": "dielectric", "eta": 1.5, "roughness": 0.0}
    sc["materials"] = mats
    return sc


@pytest.mark.parametrize("variant", ["rough", "glass", "smooth_glass"])
def test_rough_and_dielectric_materials(ctx, variant):
    sc = cornell_materials(variant)
    rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 6)
    compare(aux, oaux, f"aux {variant}")
    compare(rgb, orgb, f"film {variant}")


def test_sphere_grid_small(ctx):
    """The 2M-triangle displaced-sphere grid (diffuse, rough copper, glass) at 64x36, 4 spp, depth 8."""
    sc = scenes.sphere_grid(width=64, height=36, spp=4)
    rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 8)
    compare(aux, oaux, "aux sphere grid")
    compare(rgb, orgb, "film sphere grid")


def test_render_from_pbrt_file(ctx, tmp_path):
    """The fixture file through the .pbrt front end, then the same with a rotated and
    translated copy of the tall box (object-to-world transforms at Shape time)."""
    import os
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures", "cornell34.pbrt")
    text = open(fixture).read().replace("256", "64")
    extra = ('AttributeBegin\nTranslate 40 0 -30\nRotate 17 0 1 0\nScale 0.5 0.7 0.5\nMaterial "dielectric" '
             '"float eta" 1.33\nShape "trianglemesh" "integer indices" [0 1 2 2 3 0] "point3 P" '
             '[100 10 100  200 10 100  200 120 150  100 120 150]\nAttributeEnd\n')
    for label, body in (("fixture", text), ("transformed", text + extra)):
        p = tmp_path / f"{label}.pbrt"
        p.write_text(body)
        f = pbrtm.PbrtFile(str(p))
        s = ctx.scene_from_file(f)
        try:
            rgb, aux, st = s.render(spp=f.spp, max_depth=f.max_depth)
        finally:
            s.free()
        d = f.to_dict()
        orgb, oaux, ost = O.render(d, threads=16)
        assert st["rays"] == ost["rays"]
        compare(rgb, orgb, f"film from file ({label})")


def test_cli_renders_fixture_png(tmp_path):
    """pbrtm-render (the reference's main.cu over the C ABI) on the fixture: the PNG
    decodes to the sRGB encoding of the oracle's film."""
    import os
    import subprocess
    from PIL import Image
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "pbr-rust_amd", "pbrtm-render")
    fixture = os.path.join(root, "tests", "fixtures", "cornell34.pbrt")
    out = tmp_path / "c.png"
    r = subprocess.run([exe, fixture, "--spp", "4", "--outfile", str(out)], capture_output=True, text=True,
                       timeout=120)
    log(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    img = np.asarray(Image.open(out).convert("RGBA"))
    orgb, _, _ = O.render(scenes.cornell34(256, 256), threads=16)
    assert np.array_equal(img, O.srgb8(orgb))


def test_edge_shapes(ctx):
    """One row, one column and one pixel at the largest spp; a scene without emitters."""
    base = scenes.cornell34(8, 8)
    for w, h, spp, depth in ((37, 1, 4, 5), (1, 23, 9, 3), (1, 1, 1024, 6)):
        sc = dict(base, width=w, height=h)
        rgb, aux, orgb, oaux = render_both(ctx, sc, spp, depth)
        compare(aux, oaux, f"aux {w}x{h}x{spp}")
        compare(rgb, orgb, f"film {w}x{h}x{spp}")
    dark = scenes.cornell34(16, 16)
    dark.update(lights=[], tri_light=np.full(len(dark["tri"]), -1, np.int32))
    rgb, aux, orgb, oaux = render_both(ctx, dark, 4, 5)
    assert not np.any(rgb) and not np.any(orgb)


def test_pool_larger_than_image_and_tiny_pool(ctx):
    sc = scenes.cornell34(20, 12)
    for pool in (1, 7, 1 << 22):
        rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 4, pool=pool)
        compare(rgb, orgb, f"film pool {pool}")


def test_counting_render_is_the_same_render(ctx):
    """The counting traversal (used to price a render's traversal bytes) changes no
    film bit and no ray count, and its totals are those of a real traversal."""
    sc = scenes.sphere_grid(width=48, height=32, spp=4)
    s = ctx.scene(sc)
    try:
        rgb, aux, st = s.render(spp=4, max_depth=6)
        crgb, caux, cst = s.render(spp=4, max_depth=6, count=True)
    finally:
        s.free()
    log(f"counts closest {cst['closest_counts']} shadow {cst['shadow_counts']} "
        f"rays {cst['rays']} shadow rays {cst['shadow_rays']}")
    assert np.array_equal(rgb.view(np.uint32), crgb.view(np.uint32))
    assert np.array_equal(aux.view(np.uint32), caux.view(np.uint32))
    assert (st["rays"], st["shadow_rays"]) == (cst["rays"], cst["shadow_rays"])
    assert st["closest_counts"] == [0, 0, 0] and st["shadow_counts"] == [0, 0, 0]
    assert st["grid_quarters"] in (2, 3, 4) and cst["grid_quarters"] in (2, 3, 4)
    for key, n in (("closest_counts", cst["rays"]), ("shadow_counts", cst["shadow_rays"])):
        visits, boxes, tris = cst[key]
        # eight child boxes per visit, plus the re-tests of boxes popped after a
        # leaf of the same node shortened the ray
        assert n > 0 and visits >= n // 2 and tris > 0, key
        assert 8 * visits <= boxes <= 16 * visits, key
