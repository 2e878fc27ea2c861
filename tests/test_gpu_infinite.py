"""GPU parity of the uniform infinite light (lights/uniform_infinite_light.cu): its
light samples (a direction uniform on the sphere, the light point 2R away), the
escaped paths of both path integrators, and the wavefront's sky on paths the
depth limit ends. Films are held bit-identical to the oracle's."""
import numpy as np
import pytest

import oracle as O
import pbrtm
import scenes
from test_gpu_integrators import compare
from test_gpu_integrators import render_both as render_mk
from test_gpu_pathtracer import log, render_both
from test_oracle_infinite import closed_box, quad_at, scene

pytestmark = pytest.mark.gpu

SKY = {"rgb": (0.2, 0.4, 0.9), "scale": 2.0}


def cornell_sky(light_index, width=48, height=48):
    """The Cornell box (open towards the camera) under a blue sky, the sky placed
    before, between or after the two emissive triangles in the light list."""
    return dict(scenes.cornell34(width, height), infinite_lights=[dict(SKY, light_index=light_index)])


@pytest.mark.parametrize("light_index", [0, 1, 2])
def test_cornell_sky_wavefront(ctx, light_index):
    rgb, aux, orgb, oaux = render_both(ctx, cornell_sky(light_index), 4, 5)
    compare(aux, oaux, f"aux sky at {light_index}")
    compare(rgb, orgb, f"film sky at {light_index}")


@pytest.mark.parametrize("depth", [1, 5])
def test_cornell_sky_megakernel(ctx, depth):
    rgb, aux, orgb, oaux = render_mk(ctx, cornell_sky(0), "megakernelpath", 4, depth)
    compare(aux, oaux, f"aux megakernel sky depth {depth}")
    compare(rgb, orgb, f"film megakernel sky depth {depth}")


def test_two_skies_and_illuminance(ctx):
    sc = dict(scenes.cornell34(32, 32), infinite_lights=[
        {"rgb": None, "scale": 1.0, "illuminance": 3.0, "light_index": 0}, dict(SKY, light_index=2)])
    rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 3, pool=1000)
    compare(aux, oaux, "aux two skies")
    compare(rgb, orgb, "film two skies")


@pytest.mark.parametrize("depth", [0, 2])
def test_closed_box_depth_limited_sky(ctx, depth):
    rgb, aux, orgb, oaux = render_both(ctx, closed_box(width=32, height=32), 4, depth)
    compare(aux, oaux, f"aux closed box depth {depth}")
    compare(rgb, orgb, f"film closed box depth {depth}")
    assert np.any(rgb)


def test_plane_under_sky(ctx):
    sc = dict(scene(quad_at(0.0), width=64, height=64), infinite_lights=[{"rgb": None, "scale": 1.0}])
    rgb, aux, orgb, oaux = render_both(ctx, sc, 16, 0)
    compare(rgb, orgb, "film plane under sky")
    rgb, aux, orgb, oaux = render_mk(ctx, sc, "megakernelpath", 16, 1)
    compare(rgb, orgb, "film plane under sky (megakernel)")


def test_soup_room_with_sky(ctx):
    sc = dict(scenes.soup_room(50_000, width=48, height=40), infinite_lights=[dict(SKY, light_index=2)])
    rgb, aux, orgb, oaux = render_both(ctx, sc, 4, 4)
    compare(aux, oaux, "aux soup with sky")
    compare(rgb, orgb, "film soup with sky")


def test_batch_with_sky_equals_one_by_one(ctx):
    jobs = [(cornell_sky(0, 40, 32), 4, 5, 0), (scenes.cornell34(24, 24), 4, 3, 0),
            (cornell_sky(2, 24, 16), 9, 8, 300)]
    scs = [ctx.scene(sc) for sc, *_ in jobs]
    try:
        films, _, _, _ = ctx.render_batch([(s, spp, d, p) for s, (_, spp, d, p) in zip(scs, jobs)])
        for k, (s, (_, spp, d, p)) in enumerate(zip(scs, jobs)):
            one, _, _ = s.render(spp=spp, max_depth=d, pool_size=p)
            compare(films[k], one, f"batched job {k}")
    finally:
        for s in scs:
            s.free()


def test_pbrt_file_sky(ctx, tmp_path):
    import os
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures", "cornell34.pbrt")
    text = open(fixture).read().replace("256", "48").replace(
        "WorldBegin", 'WorldBegin\nLightSource "infinite" "rgb L" [0.2 0.4 0.9] "float scale" 2')
    p = tmp_path / "sky.pbrt"
    p.write_text(text)
    f = pbrtm.PbrtFile(str(p))
    s = ctx.scene_from_file(f)
    try:
        rgb, aux, st = s.render(spp=f.spp, max_depth=f.max_depth)
    finally:
        s.free()
    d = f.to_dict()
    assert d["infinite_lights"][0]["light_index"] == 0
    orgb, oaux, _ = O.render(d, threads=16)
    compare(aux, oaux, "aux .pbrt sky")
    compare(rgb, orgb, "film .pbrt sky")


def test_gradient_rejects_infinite_lights(ctx):
    s = ctx.scene(cornell_sky(0, 8, 8))
    try:
        with pytest.raises(pbrtm.PbrtmError, match="infinite"):
            s.render_grad(np.zeros((8, 8, 3), np.float32), spp=4, max_depth=2)
    finally:
        s.free()
    log("gradient rejected the infinite light")
