"""Pins for the oracle's path-tracer pieces (CPU only), each against something the
mathematics or a published table fixes, plus host-side parity of the library's
spectral set-up with the oracle.

- CIE 1931 Y integral over 360-830 nm = 106.856895 (spectrum_constants_cie.h:11);
  D65 relative power 100.0 at 560 nm and 109.354 at 500 nm (CIE D65 table);
  the normalised D65 integrates against Y to that constant.
- sRGB matrices against the published D65 sRGB XYZ->RGB matrix.
- Jakob-Hanika round trip: RGB -> sigmoid spectrum -> XYZ under D65 -> RGB
  within 1e-3; uniform RGB gives the closed-form c2 with s(c2) = rgb.
- Mitchell filter sampling: E[f/pdf] = integral of the filter = 1 (B + 2C = 1),
  |weight| constant (importance sampling of |f|), samples inside the radius.
- Visible-wavelength sampling: the returned pdf equals du/dlambda of the
  inverse-CDF map (finite differences); the four wavelengths are the rotations
  u + i/4.
- Stratified sampler: per dimension the strata over a pixel's samples form a
  permutation (every stratum exactly once).
- Camera: the centre ray is the view direction; off-centre rays follow the
  pinhole tangent law.
- Spherical triangle sampling: pdf = 1/solid angle (Van Oosterom-Strackee);
  inverting a sample returns its u.
- Conductor Fresnel at normal incidence: ((n-1)^2 + k^2) / ((n+1)^2 + k^2).
- Whole renders: a full-screen emitter renders as rgb * scale; a scene
  without emitters renders black.
"""
import math

import numpy as np
import pytest

import oracle as O
import scenes

LAM = np.arange(360, 831)


def spectrum(which):
    return np.array([O.spectrum_eval(which, float(l)) for l in LAM], np.float64)


def test_cie_y_integral_and_d65():
    y = spectrum(1)
    assert abs(y.sum() - 106.856895) < 1e-4 * 106.856895
    d65 = spectrum(3)
    assert abs((d65 * y).sum() - 106.856895) < 1e-5 * 106.856895
    assert abs(O.spectrum_eval(3, 560.0) / O.spectrum_eval(3, 500.0) - 100.0 / 109.354) < 1e-6
    # outside the table and at its ends
    assert O.spectrum_eval(1, 360.0) == pytest.approx(0.000003917, rel=1e-3)
    assert O.spectrum_eval(1, 555.0) == 1.0


def test_srgb_matrices_match_published():
    xyz_from_rgb, rgb_from_xyz, white = O.srgb_matrices()
    published = np.array([[3.2404542, -1.5371385, -0.4985314],
                          [-0.9692660, 1.8760108, 0.0415560],
                          [0.0556434, -0.2040259, 1.0572252]])
    assert np.abs(rgb_from_xyz - published).max() < 1e-4
    assert np.abs(white - np.array([0.3127, 0.3290])).max() < 1e-4
    assert np.abs(xyz_from_rgb @ rgb_from_xyz - np.eye(3)).max() < 1e-6


def sigmoid_poly(c, lam):
    t = (c[0] * lam + c[1]) * lam + c[2]
    return 0.5 + t / (2 * np.sqrt(1 + t * t))


@pytest.mark.parametrize("rgb", [(0.65, 0.05, 0.05), (0.12, 0.45, 0.15), (0.2, 0.9, 0.4), (0.9, 0.9, 0.1),
                                 (0.05, 0.05, 0.9), (0.5, 0.25, 0.75), (0.99, 0.01, 0.5)])
def test_rgb_to_spectrum_round_trip(rgb):
    x, y, z, d65 = spectrum(0), spectrum(1), spectrum(2), spectrum(3)
    _, rgb_from_xyz, _ = O.srgb_matrices()
    s = sigmoid_poly(O.rgb_to_sigmoid(rgb).astype(np.float64), LAM.astype(np.float64))
    xyz = np.array([(x * s * d65).sum(), (y * s * d65).sum(), (z * s * d65).sum()]) / (y * d65).sum()
    back = rgb_from_xyz.astype(np.float64) @ xyz
    assert np.abs(back - np.array(rgb)).max() < 1e-3


@pytest.mark.parametrize("v", [0.0, 0.1, 0.5, 0.73, 0.95])
def test_uniform_rgb_closed_form(v):
    c = O.rgb_to_sigmoid((v, v, v))
    assert c[0] == 0 and c[1] == 0
    if 0 < v < 1:
        t = float(c[2])
        assert 0.5 + t / (2 * math.sqrt(1 + t * t)) == pytest.approx(v, abs=1e-6)


def test_library_spectral_setup_matches_oracle():
    pbrtm = pytest.importorskip("pbrtm")
    for rgb in [(0.65, 0.05, 0.05), (0.12, 0.45, 0.15), (0.73, 0.73, 0.73), (0.3, 0.6, 0.9), (1.0, 0.0, 0.0),
                (0.5, 0.3333333, 0.16666667), (0.0, 0.0, 0.0)]:
        a, b = pbrtm.rgb_to_sigmoid(rgb), O.rgb_to_sigmoid(rgb)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), rgb
    for m, n in zip(pbrtm.srgb_matrices(), O.srgb_matrices()):
        assert np.array_equal(m.view(np.uint32), n.view(np.uint32))


def test_filter_sampling():
    n = 200
    ws, ps = [], []
    for i in range(n):
        for j in range(n):
            px, py, w = O.filter_sample((i + 0.5) / n, (j + 0.5) / n)
            ws.append(w)
            ps.append((px, py))
    ws, ps = np.array(ws), np.array(ps)
    assert np.abs(ps).max() < 2.0
    assert np.ptp(np.abs(ws)) < 1e-6 * np.abs(ws).max()
    # The filter integrates to 1 (B + 2C = 1): E[f/pdf] = K (P+ - P-)_x (P+ - P-)_y for the
    # separable table, with K = |weight|; each factor from a fine 1D sweep of u.
    K = float(np.abs(ws).max())
    m = 4000
    ey = np.mean([np.sign(O.filter_sample(0.5, (j + 0.5) / m)[2]) for j in range(m)])
    ex = np.mean([np.sign(O.filter_sample((i + 0.5) / m, 0.5)[2]) for i in range(m)])
    assert abs(K * ex * ey - 1.0) < 1e-3
    # the sign of the weight is the sign of the filter at the sample
    B = C = 1.0 / 3.0

    def m1(x):
        x = abs(x)
        if x <= 1:
            return ((12 - 9 * B - 6 * C) * x ** 3 + (-18 + 12 * B + 6 * C) * x ** 2 + (6 - 2 * B)) / 6
        if x <= 2:
            return ((-B - 6 * C) * x ** 3 + (6 * B + 30 * C) * x ** 2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C)) / 6
        return 0.0

    f = np.array([m1(px) * m1(py) for px, py in ps])
    decided = np.abs(f) > 1e-3
    assert np.all(np.sign(f[decided]) == np.sign(ws[decided]))


def test_visible_wavelengths():
    us = np.linspace(0.001, 0.999, 999)
    lam0 = np.array([O.sample_visible(u)[0][0] for u in us], np.float64)
    pdf0 = np.array([O.sample_visible(u)[1][0] for u in us], np.float64)
    assert np.all(np.diff(lam0) > 0)
    assert lam0.min() >= 359.9 and lam0.max() <= 830.1
    dudl = np.diff(us) / np.diff(lam0)
    mid = 0.5 * (pdf0[1:] + pdf0[:-1])
    assert np.abs(dudl / mid - 1).max() < 2e-3
    lam, pdf = O.sample_visible(0.3)
    for i in range(4):
        u = (0.3 + i / 4) % 1.0
        assert lam[i] == pytest.approx(O.sample_visible(u)[0][0], rel=1e-6)


@pytest.mark.parametrize("spd", [1, 2, 3, 8])
def test_stratified_strata_are_permutations(spd):
    spp = spd * spd
    kinds = [2, 1, 1, 2, 1, 2]  # the order a path draws them: filter, lambda, NEE, BSDF
    for pixel in (0, 17, 65535):
        seqs = np.array([O.sampler_sequence(spd, pixel, s, kinds) for s in range(spp)])
        assert np.all((seqs >= 0) & (seqs < 1))
        col = 0
        for k in kinds:
            if k == 1:
                strata = np.floor(seqs[:, col] * spp).astype(int)
                assert sorted(strata) == list(range(spp))
                col += 1
            else:
                sx = np.floor(seqs[:, col] * spd).astype(int)
                sy = np.floor(seqs[:, col + 1] * spd).astype(int)
                assert sorted(sy * spd + sx) == list(range(spp))
                col += 2


def test_camera_rays():
    sc = scenes.cornell34(256, 192)
    o, d = O.camera_ray(sc, 128.0, 96.0)
    assert np.allclose(o, sc["eye"], atol=1e-3)
    assert np.allclose(d, [0, 0, 1], atol=1e-6)
    t = math.tan(math.radians(37.5) / 2)  # fov spans the shorter (vertical) side
    for px, py in [(0.0, 0.0), (256.0, 192.0), (40.5, 150.25)]:
        _, d = O.camera_ray(sc, px, py)
        sx = (px / 256 * 2 - 1) * t * 256 / 192
        sy = (1 - py / 192 * 2) * t
        e = np.array([sx, sy, 1.0])  # LookAt right = up x dir = +x (pbrt's left-handed camera)
        e /= np.linalg.norm(e)
        assert np.allclose(d, e, atol=2e-6)


def solid_angle(v, p):
    a, b, c = (np.asarray(x, np.float64) - p for x in v)
    la, lb, lc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)
    num = np.dot(a, np.cross(b, c))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return abs(2 * math.atan2(num, den))


def test_spherical_triangle_sampling():
    rng = np.random.default_rng(7)
    for _ in range(40):
        v = rng.normal(0, 1, (3, 3)).astype(np.float32) + np.float32([0, 0, 3])
        p = rng.normal(0, 0.2, 3).astype(np.float32)
        omega = solid_angle(v, p.astype(np.float64))
        for _ in range(5):
            u = rng.random(2).astype(np.float32) * 0.98 + 0.01
            b, pdf = O.spherical_triangle(v, p, u)
            assert pdf == pytest.approx(1 / omega, rel=1e-3)
            assert abs(b.sum() - 1) < 1e-5 and np.all(b >= -1e-6)
            q = (b[:, None].astype(np.float64) * v).sum(0) - p
            w = (q / np.linalg.norm(q)).astype(np.float32)
            back = O.invert_spherical_triangle(v, p, w)
            assert np.abs(back - u).max() < 5e-3  # FP32 angle differences near u0 = 0


@pytest.mark.parametrize("eta,k", [(0.15, 3.6), (1.5, 0.0), (0.2, 2.0), (2.5, 1.0)])
def test_conductor_fresnel_normal_incidence(eta, k):
    expect = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)
    assert O.fr_complex(1.0, eta, k) == pytest.approx(expect, rel=1e-5)
    assert O.fr_complex(0.0, eta, k) == pytest.approx(1.0, abs=1e-5)


def full_screen_emitter(rgb, scale, w=48, h=48):
    """A camera looking at one big one-sided emitter that fills the view."""
    q = [(-500.0, -500.0, 10.0), (500.0, -500.0, 10.0), (500.0, 500.0, 10.0), (-500.0, 500.0, 10.0)]
    # winding so the geometric normal faces the camera (-z)
    tris = [q[0] + q[2] + q[1], q[0] + q[3] + q[2]]
    return {
        "eye": (0.0, 0.0, 0.0), "look": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0), "fov": 30.0,
        "width": w, "height": h, "spp": 16, "max_depth": 0,
        "tri": np.array(tris, np.float32).reshape(-1, 9), "tri_material": np.zeros(2, np.uint32),
        "tri_light": np.zeros(2, np.int32), "materials": [{"type": "diffuse", "rgb": (0.5, 0.5, 0.5)}],
        "lights": [{"rgb": rgb, "scale": scale}],
    }


@pytest.mark.parametrize("rgb,scale", [((1.0, 1.0, 1.0), 3.0), ((0.8, 0.5, 0.2), 5.0)])
def test_emitter_renders_as_its_rgb(rgb, scale):
    sc = full_screen_emitter(rgb, scale)
    img, aux, st = O.render(sc, threads=8)
    assert st["paths"] == 48 * 48 * 16
    mean = img.reshape(-1, 3).astype(np.float64).mean(0)
    assert np.abs(mean / (np.array(rgb) * scale) - 1).max() < 0.02


def test_no_emitters_is_black():
    sc = scenes.cornell34(24, 24)
    sc["tri_light"] = np.full(len(sc["tri"]), -1, np.int32)
    sc["lights"] = []
    img, _, _ = O.render(sc, threads=8)
    assert np.all(img == 0)


def test_srgb8_encoding_and_png(tmp_path):
    """Film::write_to_png's 8-bit encoding: the oracle against the exact sRGB curve
    (the reference's rational fit may differ by one code at rounding boundaries),
    the library against the oracle bit for bit, and the PNG decodes to those bytes."""
    pbrtm = pytest.importorskip("pbrtm")
    rng = np.random.default_rng(3)
    v = np.concatenate([np.linspace(-0.1, 1.1, 4001), rng.random(20000) ** 3, [np.nan, np.inf, -np.inf, 0.0031308]])
    rgb = np.stack([v, v[::-1], np.roll(v, 7)], axis=1).astype(np.float32)
    ours = O.srgb8(rgb)
    x = np.clip(np.nan_to_num(rgb.astype(np.float64), nan=0.0, posinf=1.0, neginf=0.0), 0, 1)
    exact = np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)
    assert np.abs(ours[:, :3].astype(int) - np.round(255 * exact).astype(int)).max() <= 1
    assert np.all(ours[:, 3] == 255)
    lib = pbrtm.srgb8(rgb)
    assert np.array_equal(lib, ours)
    img = rgb[:24000].reshape(120, 200, 3)
    path = tmp_path / "film.png"
    pbrtm.write_png(img, str(path))
    from PIL import Image
    dec = np.asarray(Image.open(path).convert("RGBA"))
    assert np.array_equal(dec, O.srgb8(img))
