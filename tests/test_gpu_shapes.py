"""GPU parity for analytic shapes (spheres, disks) and object instancing.

The reference's own example scene (tests/fixtures/cornell-box-specular.pbrt: a
glass Shape "sphere" in the Cornell box) through the .pbrt front end, and
shapes_room (partial and reversed spheres, emissive sphere and disk, a mirrored
annular disk, ObjectInstance copies of a box) through the scene description, are
rendered on the GPU and by the oracle from the same sampler streams: films,
sums and ray counts are held to bit-identity, for the wavefront path integrator,
the megakernel path and ambient-occlusion integrators, a 1000-slot pool and a
batch. (shapes/sphere.cu, shapes/disk.cu, primitives/transformed_primitive.cu,
scene_builder.cu:742-810.)
"""
import os

import numpy as np
import pytest

import oracle as O
import pbrtm
import scenes

pytestmark = pytest.mark.gpu

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures")


def log(msg):
    print(msg, flush=True)


def same(a, b, label):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    exact = np.mean(a.view(np.uint32) == b.view(np.uint32))
    log(f"{label}: bit-identical {exact:.6f}")
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), label


def test_reference_example_scene(ctx, tmp_path):
    """example/cornell-box-specular.pbrt at 96x96, 16 spp (its Film resolution reduced)."""
    text = open(os.path.join(FIXTURES, "cornell-box-specular.pbrt")).read().replace("1080", "96")
    text = text.replace('Integrator "path"', 'Sampler "stratified" "integer pixelsamples" [16]\nIntegrator "path"')
    p = tmp_path / "example.pbrt"
    p.write_text(text)
    f = pbrtm.PbrtFile(str(p))
    d = f.to_dict()
    assert [q["type"] for q in d["shapes"]] == ["sphere"]
    log("example scene: GPU")
    s = ctx.scene_from_file(f)
    try:
        rgb, aux, st = s.render(spp=f.spp, max_depth=f.max_depth)
    finally:
        s.free()
    log("example scene: oracle")
    orgb, oaux, ost = O.render(d, threads=16)
    assert st["rays"] == ost["rays"] and st["shadow_rays"] == ost["shadow_rays"]
    assert np.isfinite(rgb).all()
    same(aux, oaux, "example aux")
    same(rgb, orgb, "example film")


def render_pair(ctx, sc, spp, depth, pool=0, integrator="path"):
    log(f"shapes_room {sc['width']}x{sc['height']} spp {spp} depth {depth} pool {pool} {integrator}")
    s = ctx.scene(sc)
    try:
        rgb, aux, st = s.render(spp=spp, max_depth=depth, pool_size=pool, integrator=integrator)
    finally:
        s.free()
    osc = dict(sc, spp=spp, max_depth=depth)
    if integrator == "path":
        orgb, oaux, ost = O.render(osc, pool_size=pool if pool else 2 * 1024 * 1024, threads=16)
        assert st["rays"] == ost["rays"] and st["shadow_rays"] == ost["shadow_rays"]
    else:
        orgb, oaux, _ = O.render_megakernel(osc, integrator, threads=16)
    return rgb, aux, orgb, oaux


@pytest.mark.parametrize("integrator", ["path", "megakernelpath", "ambientocclusion"])
def test_shapes_room(ctx, integrator):
    sc = scenes.shapes_room(64, 64)
    rgb, aux, orgb, oaux = render_pair(ctx, sc, 16, 6, integrator=integrator)
    assert np.isfinite(rgb).all() and rgb.mean() > 0
    same(aux, oaux, f"shapes_room aux ({integrator})")
    same(rgb, orgb, f"shapes_room film ({integrator})")


def test_shapes_room_small_pool_and_depth(ctx):
    sc = scenes.shapes_room(40, 24)
    for depth, pool in ((1, 1000), (12, 0)):
        rgb, aux, orgb, oaux = render_pair(ctx, sc, 9, depth, pool=pool)
        same(rgb, orgb, f"shapes_room depth {depth} pool {pool}")


def test_shapes_in_a_batch(ctx):
    """A batch holding a shapes scene and a triangle-only scene equals their one-by-one films."""
    jobs = [(scenes.shapes_room(48, 40), 4, 5), (scenes.cornell34(32, 32), 4, 3)]
    handles = [ctx.scene(sc) for sc, _, _ in jobs]
    try:
        ref = [h.render(spp=spp, max_depth=depth)[0] for h, (_, spp, depth) in zip(handles, jobs)]
        films, _, _, _ = ctx.render_batch([(h, spp, depth, 0) for h, (_, spp, depth) in zip(handles, jobs)],
                                          pool_size=3000)
    finally:
        for h in handles:
            h.free()
    for j in range(len(jobs)):
        same(films[j], ref[j], f"batch job {j}")


def test_instancing_from_pbrt(ctx, tmp_path):
    """ObjectBegin / ObjectInstance through the front end: one identity instance
    (the definition's triangles themselves) and two transformed ones, plus a disk
    and a partial sphere."""
    body = open(os.path.join(FIXTURES, "cornell34.pbrt")).read().replace("256", "48")
    body = body.replace("WorldBegin", "WorldBegin\n" + (
        'ObjectBegin "crate"\nMaterial "diffuse" "rgb reflectance" [0.3 0.5 0.7]\n'
        'Shape "trianglemesh" "integer indices" [0 1 2 2 3 0  4 5 6 6 7 4] "point3 P" '
        '[0 0 0  60 0 0  60 60 0  0 60 0  0 0 60  60 0 60  60 60 60  0 60 60]\nObjectEnd\n'), 1)
    body += ('ObjectInstance "crate"\n'
             'AttributeBegin\nTranslate 380 0 100\nRotate 30 0 1 0\nObjectInstance "crate"\nAttributeEnd\n'
             'AttributeBegin\nTranslate 120 300 250\nScale 1.5 0.5 1.5\nObjectInstance "crate"\nAttributeEnd\n'
             'AttributeBegin\nTranslate 280 1 300\nRotate -90 1 0 0\nMaterial "conductor"\n'
             'Shape "disk" "float radius" [70] "float innerradius" [20]\nAttributeEnd\n'
             'AttributeBegin\nTranslate 200 250 300\nRotate 40 1 1 0\nMaterial "dielectric" "float eta" [1.4]\n'
             'Shape "sphere" "float radius" [50] "float zmin" [-30] "float zmax" [45] "float phimax" [270]\n'
             'AttributeEnd\n')
    p = tmp_path / "instances.pbrt"
    p.write_text(body)
    f = pbrtm.PbrtFile(str(p))
    d = f.to_dict()
    kinds = [q["type"] for q in d["shapes"]]
    assert kinds.count("instance") == 3 and "disk" in kinds and "sphere" in kinds
    s = ctx.scene_from_file(f)
    try:
        rgb, aux, st = s.render(spp=f.spp, max_depth=f.max_depth)
    finally:
        s.free()
    orgb, oaux, ost = O.render(d, threads=16)
    assert st["rays"] == ost["rays"]
    same(rgb, orgb, "instances film")
