"""Pins of the oracle's megakernel integrators (megakernelpath, ambientocclusion,
surfacenormal), base/megakernel_integrator.cu:14-147.

The megakernel path integrator differs from the wavefront one only at the last
vertex (emission, then the path ends: no light sample, no BSDF sample) and in
when Russian roulette draws its sample. With the same sampler streams that
fixes, path by path: WF(d - 1) <= MK(d) <= WF(d), and MK(d) == WF(d) where no
vertex takes a light sample (an all-specular scene), Russian roulette included.
Ambient occlusion and surface normals are checked against closed forms: an
unoccluded plane reflects the normalised illuminant (the same radiance as a
unit emitter), a closed box is black, and a plane facing +y has the softmax
colour (1, 10, 1) / 12 of three 0.01-albedo spectra.
"""
import numpy as np
import pytest

import oracle as O
import scenes


def all_specular(sc):
    sc = dict(sc)
    sc["materials"] = [{"type": "conductor", "metal": "Ag", "roughness": 0.0, "remap": True}] * len(sc["materials"])
    return sc


@pytest.mark.parametrize("depth", [1, 2, 3, 5])
def test_megakernel_path_between_wavefront_depths(depth):
    sc = scenes.cornell34(24, 24, spp=4, max_depth=depth)
    _, _, _, mk = O.render_megakernel(sc, "megakernelpath", path_L=True)
    lo, _ = O.render_path_L(dict(sc, max_depth=depth - 1))
    hi, _ = O.render_path_L(sc)
    assert np.all(mk >= lo) and np.all(mk <= hi)
    # both bounds are strict somewhere: the last vertex's emission is added, its
    # light sample is not
    assert np.any(mk != lo) and np.any(mk != hi)
    eq_lo = np.all(mk == lo, axis=-1).mean()
    assert eq_lo > 0.9, eq_lo


@pytest.mark.parametrize("depth", [3, 12])
def test_megakernel_path_equals_wavefront_without_light_samples(depth):
    """No light samples on a smooth-silver scene: the last vertex's emission is the
    only difference left, and both add it. At depth 12 Russian roulette draws a
    sample past 8 bounces on both (silver keeps beta below 1)."""
    sc = all_specular(scenes.cornell34(24, 24, spp=4, max_depth=depth))
    _, _, st, mk = O.render_megakernel(sc, "megakernelpath", path_L=True)
    wf, _ = O.render_path_L(sc)
    assert st["shadow_rays"] == 0
    assert np.array_equal(mk.view(np.uint32), wf.view(np.uint32))
    assert np.any(mk > 0)


def plane(width=24, height=24, spp=4, emissive=False):
    """A 200 x 200 quad at y = 0 seen from (0, 5, 0.001) straight down: every camera
    ray hits it and nothing lies above it."""
    tri = np.array(scenes.quad((-100.0, 0.0, -100.0), (100.0, 0.0, -100.0), (100.0, 0.0, 100.0),
                               (-100.0, 0.0, 100.0)), np.float32).reshape(-1, 9)
    return {
        "eye": (0.0, 5.0, 0.001), "look": (0.0, 0.0, 0.0), "up": (0.0, 0.0, 1.0), "fov": 40.0,
        "width": width, "height": height, "spp": spp, "max_depth": 0,
        "tri": tri, "tri_material": np.zeros(2, np.uint32),
        "tri_light": np.zeros(2, np.int32) if emissive else np.full(2, -1, np.int32),
        "materials": [{"type": "diffuse", "rgb": (0.5, 0.5, 0.5)}],
        "lights": [{"rgb": (1.0, 1.0, 1.0), "scale": 1.0, "two_sided": True}] if emissive else [],
    }


def test_ambient_occlusion_open_plane_is_unit_illuminant():
    """Unoccluded: L = illuminant(lambda) / Y(illuminant) * cos / (pi pdf) with pdf =
    cos / pi, i.e. the radiance of a unit-scale rgb (1, 1, 1) emitter seen directly
    (megakernelpath at depth 0 on the same plane made emissive)."""
    sc = plane(emissive=True)
    _, _, st, ao = O.render_megakernel(sc, "ambientocclusion", path_L=True)
    _, _, _, le = O.render_megakernel(sc, "megakernelpath", path_L=True)
    n = sc["width"] * sc["height"] * sc["spp"]
    assert st["rays"] == n and st["shadow_rays"] == n
    assert np.all(le > 0)
    ratio = ao / le
    assert float(np.abs(ratio - 1).max()) < 2e-3, float(np.abs(ratio - 1).max())


def test_ambient_occlusion_closed_box_is_black():
    sc = plane()
    sc.update(tri=np.array(scenes.box(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0), np.float32).reshape(-1, 9),
              eye=(0.0, 0.0, -0.5), look=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), fov=60.0)
    sc["tri_material"] = np.zeros(len(sc["tri"]), np.uint32)
    sc["tri_light"] = np.full(len(sc["tri"]), -1, np.int32)
    rgb, _, st, L = O.render_megakernel(sc, "ambientocclusion", path_L=True)
    n = sc["width"] * sc["height"] * sc["spp"]
    assert st["rays"] == n and st["shadow_rays"] == n
    assert not np.any(L) and not np.any(rgb)


def sigmoid(c, lam):
    x = (c[0] * lam + c[1]) * lam + c[2]
    return 0.5 + x / (2.0 * np.sqrt(1.0 + x * x))


def test_surface_normal_plane_softmax():
    sc = plane(width=8, height=8)
    _, _, st, L = O.render_megakernel(sc, "surfacenormal", path_L=True)
    assert st["shadow_rays"] == 0
    coef = [O.rgb_to_sigmoid(np.eye(3)[c] * 0.01).astype(np.float64) for c in range(3)]
    col = np.array([1.0, 10.0, 1.0]) / 12.0  # softmax of the +y normal, base 10
    spd, wh = 2, sc["width"] * sc["height"]
    for pixel in (0, 13, 63):
        for sample in range(sc["spp"]):
            u = O.sampler_sequence(spd, pixel, sample, [2, 1])  # filter 2D, wavelengths 1D
            lam, _ = O.sample_visible(u[2])
            want = sum(col[c] * sigmoid(coef[c], lam.astype(np.float64)) for c in range(3))
            got = L[sample].reshape(wh, 4)[pixel]
            # FP32 sigmoid: 0.5 + x / (2 sqrt(1 + x^2)) cancels for these dark
            # spectra, an absolute error of a few ulp of 0.5
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=1.5e-7)
