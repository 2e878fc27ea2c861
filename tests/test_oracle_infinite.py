"""Pins of the oracle's uniform infinite light (lights/uniform_infinite_light.cu).

Per path, with the camera's sampler dimensions (so the wavelengths) shared across
scenes: a ray that escapes at depth 0 carries exactly Le = scale L(lambda) /
Y(illuminant), the same value as ambient occlusion on an unoccluded plane; a
closed grey box under a sky is black for the megakernel path integrator but, for
the wavefront one, every path ended by the depth limit still collects the sky
(wavefront_path.cu:140-160 adds infinite lights to any path ending with a
positive beta): L = R^(d+1) Le. A grey plane under the sky, traced to depth 0,
sums the light sample (MIS-weighted, pdf 1/(4 pi)) and the escaped BSDF sample
at full weight (pdf_li with an incomplete pdf allowed is 0), whose expectation
is R Le (1 + ln(17) / 16).
"""
import math

import numpy as np

import oracle as O
import scenes

GREY = {"type": "diffuse", "rgb": (0.5, 0.5, 0.5)}


def view(width=24, height=24, spp=4):
    """The camera of the plane scenes: (0, 5, 0.001) looking straight down."""
    return {"eye": (0.0, 5.0, 0.001), "look": (0.0, 0.0, 0.0), "up": (0.0, 0.0, 1.0), "fov": 40.0,
            "width": width, "height": height, "spp": spp, "max_depth": 0, "materials": [GREY], "lights": []}


def quad_at(y, size=100.0):
    return np.array(scenes.quad((-size, y, -size), (size, y, -size), (size, y, size), (-size, y, size)),
                    np.float32).reshape(-1, 9)


def scene(tri, **kw):
    sc = view(**kw)
    sc.update(tri=tri, tri_material=np.zeros(len(tri), np.uint32), tri_light=np.full(len(tri), -1, np.int32))
    return sc


def sky(**kw):
    """Only sky in view: the one quad hangs above the camera."""
    return dict(scene(quad_at(10.0), **kw), infinite_lights=[{"rgb": None, "scale": 1.0}])


def test_sky_seen_directly_is_the_normalised_illuminant():
    sc = sky()
    wf, _ = O.render_path_L(sc)
    _, _, st, mk = O.render_megakernel(sc, "megakernelpath", path_L=True)
    assert np.array_equal(wf.view(np.uint32), mk.view(np.uint32)) and np.all(wf > 0)
    # ambient occlusion of an unoccluded plane seen through the same pixels:
    # illuminant / Y * cos / (pi pdf), the same wavelengths per path
    _, _, _, ao = O.render_megakernel(scene(quad_at(0.0)), "ambientocclusion", path_L=True)
    np.testing.assert_allclose(wf, ao, rtol=1e-5)
    assert st["rays"] == sc["width"] * sc["height"] * sc["spp"] and st["shadow_rays"] == 0


def test_rgb_sky_film():
    """rgb L through the RGB-to-spectrum table and the sensor: the film reproduces
    rgb x scale (the table's D65 round trip, spectral noise at 16 spp)."""
    sc = dict(sky(width=32, height=32, spp=16), infinite_lights=[{"rgb": (0.2, 0.4, 0.9), "scale": 2.0}])
    _, aux, _ = O.render(sc)
    mean = aux[..., :3].sum(axis=(0, 1)) / aux[..., 3].sum()
    rgb = np.array(O.srgb_matrices()[1]).reshape(3, 3) @ mean  # sensor XYZ -> output RGB
    np.testing.assert_allclose(rgb, [0.4, 0.8, 1.8], rtol=0.03)


def closed_box(**kw):
    tri = np.array(scenes.box(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0), np.float32).reshape(-1, 9)
    sc = scene(tri, **kw)
    sc.update(eye=(0.0, 0.0, -0.5), look=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), fov=60.0,
              infinite_lights=[{"rgb": None, "scale": 1.0}])
    return sc


def test_closed_box_depth_limited_paths_collect_the_sky():
    sc = closed_box()
    # the sky through the same camera, for Le at each path's wavelengths
    le, _ = O.render_path_L(dict(sky(), eye=sc["eye"], look=sc["look"], up=sc["up"], fov=sc["fov"],
                                 tri=np.array(scenes.box(5.0, 6.0, 5.0, 6.0, 5.0, 6.0), np.float32).reshape(-1, 9),
                                 tri_material=np.zeros(12, np.uint32), tri_light=np.full(12, -1, np.int32)))
    assert np.all(le > 0)
    for d in (0, 1, 2):
        wf, _ = O.render_path_L(dict(sc, max_depth=d))
        ratio = wf / le
        # beta after d + 1 grey bounces (f cos / pdf = R for cosine sampling); every
        # light sample is blocked by the box
        np.testing.assert_allclose(ratio, 0.5 ** (d + 1), rtol=2e-3)
        _, _, st, mk = O.render_megakernel(dict(sc, max_depth=d), "megakernelpath", path_L=True)
        assert not np.any(mk), "the megakernel path integrator never escapes a closed box"


def test_plane_under_sky_light_sample_and_escape():
    kw = dict(width=64, height=64, spp=16)
    plane = dict(scene(quad_at(0.0), **kw), infinite_lights=[{"rgb": None, "scale": 1.0}])
    le, _ = O.render_path_L(sky(**kw))
    for integrator, depth in (("path", 0), ("megakernelpath", 1)):
        if integrator == "path":
            L, _ = O.render_path_L(dict(plane, max_depth=depth))
        else:
            L = O.render_megakernel(dict(plane, max_depth=depth), integrator, path_L=True)[3]
        got = float((L / le).mean())
        want = 0.5 * (1 + math.log(17.0) / 16)
        assert abs(got / want - 1) < 5e-3, (integrator, got, want)
