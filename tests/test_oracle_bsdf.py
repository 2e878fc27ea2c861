"""Pins of the oracle's shading functions to the mathematics they implement.

The GPU kernels are compared with the oracle bit for bit elsewhere; these tests
check the oracle itself against facts that do not depend on reading the
reference a second time: textbook Fresnel and Snell's law, the normalisation and
Smith identities of the Trowbridge-Reitz distribution, the agreement of every
lobe's sampling routine with its own density and value, reciprocity, energy
conservation of the smooth dielectric, the power light sampler's closed form,
an area light's irradiance closed form through NEE + MIS, and the measured
spectral tables against published values.

Oracle functions covered (pt_oracle.cpp): fr_dielectric (scattering.h:82-102),
refract (scattering.h:15-44), TR D / Lambda / G1 / D_w / sample_wm
(scattering.h:105-229), the diffuse, conductor and dielectric BxDFs' f / pdf /
sample_f (diffuse_bxdf.h, conductor_bxdf.cu:3-106, dielectric_bxdf.cu:4-230),
the power light sampler's pmf (power_light_sampler.cu:7-45,
diffuse_area_light.cu:80-85, uniform_infinite_light.cu phi), and NEE with MIS
against area lights (wavefront_path.cu:574-624).
"""
import numpy as np
import pytest

import oracle as O


# ----------------------------------------------------------------- helpers
def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def sph(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def hemisphere_grid(a, nx=384, nphi=384):
    """Directions over the upper hemisphere and their solid-angle weights, with
    polar nodes concentrated where a Trowbridge-Reitz lobe of width a lives:
    tan(theta) = a tan(x), x Gauss-Legendre on [0, pi/2]."""
    x, wx = np.polynomial.legendre.leggauss(nx)
    x = (x + 1) * np.pi / 4
    wx = wx * np.pi / 4
    th = np.arctan(a * np.tan(x))
    dth = a / np.cos(x) ** 2 / (1 + (a * np.tan(x)) ** 2)
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, P = np.meshgrid(th, phi, indexing="ij")
    w = (wx * dth)[:, None] * (2 * np.pi / nphi) * np.sin(T)
    return sph(T, P).reshape(-1, 3), w.reshape(-1)


def sphere_grid(n=1600):
    """The whole sphere: Gauss-Legendre in cos(theta) x uniform phi."""
    c, wc = np.polynomial.legendre.leggauss(n)
    phi = (np.arange(2 * n) + 0.5) * np.pi / n
    C, P = np.meshgrid(c, phi, indexing="ij")
    s = np.sqrt(1 - C * C)
    d = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
    w = (wc[:, None] * np.full(2 * n, np.pi / n)[None, :]).reshape(-1)
    return d, w


def textbook_fresnel(cos_i, n1, n2):
    """Unpolarised reflectance of a dielectric interface (Fresnel's equations)."""
    sin_t = n1 / n2 * np.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t >= 1:
        return 1.0
    cos_t = np.sqrt(1 - sin_t * sin_t)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return (rs * rs + rp * rp) / 2


# ------------------------------------------------------- Fresnel, refraction
@pytest.mark.parametrize("eta", [1.33, 1.5, 2.4, 0.7])
def test_fr_dielectric_is_textbook_fresnel(eta):
    for c in np.linspace(-1, 1, 201):
        n1, n2 = (1.0, eta) if c >= 0 else (eta, 1.0)
        want = textbook_fresnel(abs(c), n1, n2)
        assert O.fr_dielectric(c, eta) == pytest.approx(want, abs=3e-6), (c, eta)
    # normal incidence from either side: ((eta - 1) / (eta + 1))^2
    assert O.fr_dielectric(1.0, eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-6)
    assert O.fr_dielectric(-1.0, eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-6)
    # past the critical angle (leaving the denser side) everything is reflected
    dense_side = -1.0 if eta > 1 else 1.0
    n_rel = eta if eta > 1 else 1 / eta
    crit = np.sqrt(1 - 1 / n_rel ** 2)  # cos of the critical angle
    for c in np.linspace(0.0, crit * 0.999, 20):
        assert O.fr_dielectric(dense_side * c, eta) == 1.0
    assert O.fr_dielectric(1.0, 1.0) == 0.0


def test_refract_follows_snell_and_total_internal_reflection():
    rng = np.random.default_rng(1)
    n = np.array([0, 0, 1], np.float32)
    for eta in (1.5, 1.33, 0.8):
        for _ in range(400):
            wi = unit(rng.normal(size=3)).astype(np.float32)
            if abs(wi[2]) < 1e-3:
                continue
            cos_i = float(wi[2])
            n1, n2 = (1.0, eta) if cos_i > 0 else (eta, 1.0)
            sin_i = np.sqrt(max(0.0, 1 - cos_i * cos_i))
            r = O.refract(wi, n, eta)
            if n1 / n2 * sin_i > 1 + 1e-6:
                assert r is None
                continue
            if n1 / n2 * sin_i > 1 - 1e-6:
                continue  # at the critical angle either answer is within rounding
            assert r is not None
            wt, etap = r
            assert etap == pytest.approx(n2 / n1, rel=1e-6)
            assert np.linalg.norm(wt) == pytest.approx(1.0, abs=2e-6)
            assert wt[2] * cos_i < 0  # crosses the interface
            sin_t = np.hypot(wt[0], wt[1])
            assert n2 * sin_t == pytest.approx(n1 * sin_i, abs=3e-6)  # Snell
            # tangential parts anti-parallel (wi points away from the surface)
            t_i, t_t = wi[:2].astype(np.float64), wt[:2].astype(np.float64)
            if sin_i > 1e-3:
                assert np.dot(t_i, t_t) == pytest.approx(-np.linalg.norm(t_i) * np.linalg.norm(t_t), rel=1e-5)


# ------------------------------------------------------- Trowbridge-Reitz
ALPHAS = [(0.05, 0.05), (0.3, 0.3), (0.8, 0.8), (0.1, 0.5), (0.6, 0.15)]


@pytest.mark.parametrize("ax,ay", ALPHAS)
def test_tr_normalisation(ax, ay):
    """Projected microfacet area equals the macro surface: int D(wm) cos(theta_m) dwm = 1."""
    d, w = hemisphere_grid(np.sqrt(ax * ay), nx=512, nphi=512)
    D = O.tr_eval(ax, ay, d, d)[:, 0].astype(np.float64)
    assert np.sum(D * d[:, 2] * w) == pytest.approx(1.0, abs=2e-4)


@pytest.mark.parametrize("ax,ay", ALPHAS)
def test_tr_smith_projected_area(ax, ay):
    """Smith: int max(0, w.wm) D(wm) dwm = cos(theta) (1 + Lambda(w)), and G1 = 1 / (1 + Lambda);
    the visible normal density D_w integrates to 1 over the normals facing w (its
    formula takes |w.wm|, but sampling never returns a back-facing normal)."""
    d, w = hemisphere_grid(np.sqrt(ax * ay), nx=512, nphi=512)
    for th, ph in [(0.0, 0.0), (0.5, 0.3), (1.0, 1.2), (1.35, 2.5), (1.5, 4.0)]:
        wo = sph(th, ph)
        out = O.tr_eval(ax, ay, np.broadcast_to(wo, d.shape), d).astype(np.float64)
        D, lam, g1, dw = out[:, 0], out[0, 1], out[0, 2], out[:, 3]
        proj = np.sum(np.maximum(0.0, d @ wo) * D * w)
        assert proj == pytest.approx(np.cos(th) * (1 + lam), rel=1e-3), (th, ph)
        assert g1 == pytest.approx(1 / (1 + lam), rel=1e-6)
        assert np.sum(dw * (d @ wo > 0) * w) == pytest.approx(1.0, abs=2e-3), (th, ph)


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.1, 0.5), (0.8, 0.4)])
def test_tr_sample_wm_draws_the_visible_normal_density(ax, ay):
    """Moments of Sample_wm's normals equal the moments of D_w by quadrature."""
    rng = np.random.default_rng(7)
    d, w = hemisphere_grid(np.sqrt(ax * ay), nx=384, nphi=384)
    for th, ph in [(0.2, 0.0), (0.9, 0.7), (1.4, 2.0)]:
        wo = sph(th, ph)
        dw = O.tr_eval(ax, ay, np.broadcast_to(wo, d.shape), d)[:, 3].astype(np.float64) * (d @ wo > 0)
        n = 200_000
        wm = O.tr_sample(ax, ay, np.broadcast_to(wo, (n, 3)), rng.random((n, 2))).astype(np.float64)
        assert np.allclose(np.linalg.norm(wm, axis=1), 1, atol=2e-6)
        for f in (lambda v: v[:, 2], lambda v: v[:, 0], lambda v: v[:, 1], lambda v: v[:, 0] * v[:, 1],
                  lambda v: v[:, 0] ** 2, lambda v: np.maximum(0, v @ wo)):
            want = np.sum(f(d) * dw * w)
            got, sd = f(wm).mean(), f(wm).std() / np.sqrt(n)
            assert abs(got - want) < 5 * sd + 1e-4, (th, ph, got, want)


# ----------------------------------------------------------------- BxDF lobes
CU = dict(eta4=(0.27, 0.6, 1.1, 1.2), k4=(3.4, 2.6, 2.5, 2.4))
LOBES = {
    "rough_conductor": O.bxdf("conductor", ax=0.3, **CU),
    "aniso_conductor": O.bxdf("conductor", ax=0.15, ay=0.45, **CU),
    "rough_glass": O.bxdf("dielectric", eta=1.5, ax=0.3),
    "rough_glass_aniso": O.bxdf("dielectric", eta=1.33, ax=0.2, ay=0.5),
    "diffuse": O.bxdf("diffuse", r=(0.2, 0.5, 0.7, 0.9)),
}
WO = [sph(0.1, 0.3), sph(0.8, 1.9), sph(1.3, 4.0), sph(np.pi - 0.6, 0.5), sph(np.pi - 1.2, 2.8)]


def _lobe_sides(name, wo):
    return name.startswith("rough_glass") or wo[2] > 0


@pytest.mark.parametrize("name", sorted(LOBES))
def test_sample_f_agrees_with_f_and_pdf(name):
    """sample_f's (f, pdf) equal f() and pdf() at the direction it returns."""
    b = LOBES[name]
    rng = np.random.default_rng(11)
    for wo in WO:
        if not _lobe_sides(name, wo) and name != "diffuse":
            continue
        n = 20000
        f, wi, pdf, eta, fl = O.bxdf_sample(b, np.broadcast_to(wo, (n, 3)), rng.random((n, 3)))
        ok = (fl & 1) != 0
        assert ok.mean() > 0.5
        assert not np.any(fl[ok] & 2)  # rough lobes are not specular
        fe, pe = O.bxdf_eval(b, np.broadcast_to(wo, (int(ok.sum()), 3)), wi[ok])
        good = np.isclose(fe, f[ok], rtol=2e-3, atol=1e-7).all(axis=1) & np.isclose(pe, pdf[ok], rtol=2e-3)
        assert good.mean() > 0.999, (name, wo, good.mean())
        # transmissions report the relative IOR of the side they enter, reflections 1
        tr = ok & ((fl & 4) != 0)
        assert np.all((wi[tr, 2] * wo[2]) < 0)
        if name.startswith("rough_glass"):
            e = b.eta
            assert np.allclose(eta[tr], e if wo[2] > 0 else 1 / e, rtol=1e-6)
        assert np.all(eta[ok & ~tr] == 1.0)


@pytest.mark.parametrize("name", sorted(LOBES))
def test_pdf_integrates_to_the_sampling_success_rate(name):
    """int pdf(wo, wi) dwi over the sphere (transmission included) equals the
    fraction of sample_f calls that return a direction (the rest fall below the
    horizon or into total internal reflection), and <= 1."""
    b = LOBES[name]
    d, w = sphere_grid(1600)
    rng = np.random.default_rng(3)
    for wo in WO:
        if not _lobe_sides(name, wo) and name != "diffuse":
            continue
        _, pdf = O.bxdf_eval(b, np.broadcast_to(wo, d.shape), d)
        integral = np.sum(pdf.astype(np.float64) * w)
        n = 400_000
        fl = O.bxdf_sample(b, np.broadcast_to(wo, (n, 3)), rng.random((n, 3)))[4]
        rate = ((fl & 1) != 0).mean()
        assert integral <= 1 + 2e-3
        assert integral == pytest.approx(rate, abs=3e-3), (name, wo, integral, rate)


@pytest.mark.parametrize("name", sorted(LOBES))
def test_albedo_by_quadrature_equals_sampling_estimate(name):
    """int f(wo, wi) |cos wi| dwi by quadrature equals E[f |cos wi| / pdf] over sample_f."""
    b = LOBES[name]
    d, w = sphere_grid(1600)
    rng = np.random.default_rng(5)
    for wo in WO[:3] + WO[3:] * name.startswith("rough_glass"):
        f, _ = O.bxdf_eval(b, np.broadcast_to(wo, d.shape), d)
        quad = (f.astype(np.float64) * np.abs(d[:, 2:3]) * w[:, None]).sum(0)
        n = 400_000
        fs, wi, pdf, _, fl = O.bxdf_sample(b, np.broadcast_to(wo, (n, 3)), rng.random((n, 3)))
        ok = (fl & 1) != 0
        est = np.zeros((n, 4))
        est[ok] = fs[ok] * np.abs(wi[ok, 2:3]) / pdf[ok, None]
        mean, sd = est.mean(0), est.std(0) / np.sqrt(n)
        assert np.all(np.abs(mean - quad) < 5 * sd + 2e-3 * np.abs(quad) + 1e-5), (name, wo, mean, quad)


def test_reciprocity():
    """Reflection: f(wo, wi) = f(wi, wo). Transmission (radiance, pbrt-v4's
    f(wo, wi) for light arriving along wi): f(wo, wi) / eta_o^2 = f(wi, wo) / eta_i^2."""
    rng = np.random.default_rng(9)
    n = 20000
    a = unit(rng.normal(size=(n, 3)))
    c = unit(rng.normal(size=(n, 3)))
    for name in ("rough_conductor", "aniso_conductor", "rough_glass", "rough_glass_aniso"):
        b = LOBES[name]
        f1, _ = O.bxdf_eval(b, a, c)
        f2, _ = O.bxdf_eval(b, c, a)
        refl = a[:, 2] * c[:, 2] > 0
        m = refl & (np.abs(a[:, 2]) > 0.05) & (np.abs(c[:, 2]) > 0.05)
        assert np.isclose(f1[m], f2[m], rtol=5e-4, atol=1e-6).all(axis=1).mean() > 0.999, name
        if name.startswith("rough_glass"):
            e = b.eta
            eo = np.where(a[:, 2] > 0, 1.0, e)[:, None]  # index on wo's side
            ei = np.where(c[:, 2] > 0, 1.0, e)[:, None]
            t = ~refl & (np.abs(a[:, 2]) > 0.05) & (np.abs(c[:, 2]) > 0.05)
            assert t.sum() > 1000 and np.any(f1[t] > 0)
            close = np.isclose(f1[t] / eo[t] ** 2, f2[t] / ei[t] ** 2, rtol=5e-4, atol=1e-6).all(axis=1)
            assert close.mean() > 0.999, name  # the rest: half vectors near-degenerate in FP32
        if name.endswith("conductor"):
            assert np.all(f1[~refl] == 0)


@pytest.mark.parametrize("eta", [1.5, 1.33, 2.4])
def test_smooth_dielectric_conserves_energy(eta):
    """Smooth dielectric: the reflected (uc below R) and transmitted samples carry
    Fresnel's R and T = 1 - R (transmission scaled by 1 / eta'^2 for radiance),
    their probabilities are R and T, and past the critical angle only reflection
    remains."""
    b = O.bxdf("dielectric", eta=eta, ax=0.0)
    assert O.bxdf_flags(b) == (1 | 2 | 16)
    for th in np.linspace(0.01, np.pi - 0.01, 41):
        if abs(np.cos(th)) < 0.02:
            continue
        wo = sph(th, 0.7)
        n1, n2 = (1.0, eta) if wo[2] > 0 else (eta, 1.0)
        R = textbook_fresnel(abs(wo[2]), n1, n2)
        f, wi, pdf, e, fl = O.bxdf_sample(b, np.array([wo, wo]), np.array([[0.0, 0.5, 0.5], [0.999999, 0.5, 0.5]]))
        # the reflected sample: the mirror direction, f |cos| = R, pdf = R
        assert fl[0] & 1 and fl[0] & 2 and not fl[0] & 4
        assert np.allclose(wi[0], [-wo[0], -wo[1], wo[2]], atol=1e-6)
        assert f[0, 0] * abs(wi[0, 2]) == pytest.approx(R, abs=2e-6)
        assert pdf[0] == pytest.approx(R, abs=2e-6)
        if R >= 1.0:  # total internal reflection: no transmitted sample
            assert not (fl[1] & 4)
            continue
        assert fl[1] & 4 and fl[1] & 2
        T = f[1, 0] * abs(wi[1, 2]) * e[1] ** 2
        assert R + T == pytest.approx(1.0, abs=3e-6)
        assert pdf[0] + pdf[1] == pytest.approx(1.0, abs=3e-6)
        assert e[1] == pytest.approx(n2 / n1, rel=1e-6)
        sin_t = np.hypot(wi[1, 0], wi[1, 1])
        assert n2 * sin_t == pytest.approx(n1 * np.sin(th), abs=3e-6)
    # eta = 1: a pass-through (transmission only)
    assert O.bxdf_flags(O.bxdf("dielectric", eta=1.0)) == (2 | 16)


def test_diffuse_white_furnace():
    """Lambertian albedo 1: pdf = cos / pi and every sample weighs f cos / pdf = 1."""
    b = O.bxdf("diffuse", r=(1, 1, 1, 1))
    rng = np.random.default_rng(2)
    n = 50000
    for wo in (sph(0.3, 1.0), sph(2.5, 0.2)):
        f, wi, pdf, _, fl = O.bxdf_sample(b, np.broadcast_to(wo, (n, 3)), rng.random((n, 3)))
        ok = (fl & 1) != 0
        assert ok.mean() > 0.999
        assert np.all(wi[ok, 2] * wo[2] > 0)
        assert np.allclose(pdf[ok], np.abs(wi[ok, 2]) / np.pi, rtol=1e-5)
        assert np.allclose(f[ok] * np.abs(wi[ok, 2:3]) / pdf[ok, None], 1.0, rtol=1e-5)


def test_conductor_with_unit_fresnel():
    """A conductor with |F| = 1 (k large): the smooth lobe returns all the light,
    f cos / pdf = F; a rough lobe returns E[f cos / pdf] = E[G / G1] <= 1 over
    visible normals (single scattering loses the light that would bounce again),
    equal to the Smith closed form integrated by quadrature."""
    big = dict(eta4=(1.0,) * 4, k4=(1e4,) * 4)
    rng = np.random.default_rng(4)
    smooth = O.bxdf("conductor", ax=0.0, **big)
    assert O.bxdf_flags(smooth) == (1 | 16)
    for th in (0.0, 0.7, 1.3):
        wo = sph(th, 0.4)
        f, wi, pdf, _, fl = O.bxdf_sample(smooth, wo[None], np.array([[0.3, 0.3, 0.3]]))
        assert fl[0] & 2 and pdf[0] == 1.0
        assert f[0, 0] * abs(wi[0, 2]) == pytest.approx(1.0, abs=1e-6)
        assert f[0, 0] * abs(wi[0, 2]) == pytest.approx(O.fr_complex(np.cos(th), 1.0, 1e4), rel=1e-6)
    for a in (0.05, 0.2, 0.6):
        b = O.bxdf("conductor", ax=a, **big)
        wo = sph(0.6, 0.3)
        n = 200_000
        f, wi, pdf, _, fl = O.bxdf_sample(b, np.broadcast_to(wo, (n, 3)), rng.random((n, 3)))
        ok = (fl & 1) != 0
        est = np.zeros(n)
        est[ok] = f[ok, 0] * np.abs(wi[ok, 2]) / pdf[ok]
        # E over visible normals of G(wo, wi) / G1(wo) = (1 + Lo) / (1 + Lo + Li), wi = reflect(wo, wm) above
        d, w = hemisphere_grid(a, nx=384, nphi=384)
        vis = O.tr_eval(a, a, np.broadcast_to(wo, d.shape), d).astype(np.float64)
        ri = 2 * (d @ wo)[:, None] * d - wo
        lam_o, lam_i = vis[0, 1], O.tr_eval(a, a, ri, d)[:, 1].astype(np.float64)
        g = np.where((ri[:, 2] > 0) & (d @ wo > 0), (1 + lam_o) / (1 + lam_o + lam_i), 0.0)
        want = np.sum(vis[:, 3] * g * w)
        assert est.mean() <= 1.0 + 1e-4
        assert est.mean() == pytest.approx(want, abs=5 * est.std() / np.sqrt(n) + 2e-4), (a, est.mean(), want)


# ----------------------------------------------------------- light sampling
def _emitter_scene(lights, tris, tri_light, infinite=()):
    n = len(tris)
    return dict(eye=(0, 0, -5), look=(0, 0, 0), up=(0, 1, 0), fov=40.0, width=8, height=8, spp=1, max_depth=1,
                tri=np.asarray(tris, np.float32).reshape(n, 3, 3), tri_material=np.zeros(n, np.uint32),
                tri_light=np.asarray(tri_light, np.int32), albedo=np.array([[0.5, 0.5, 0.5]], np.float32),
                lights=lights, infinite_lights=list(infinite))


def _tri(ox, size):
    return [(ox, 0, 0), (ox + size, 0, 0), (ox, size, 0)]


def test_power_sampler_pmf_closed_form():
    """pmf ∝ phi = pi (2 if two-sided) area scale L: for lights of one colour the
    spectral factor cancels, so pmf_i = sides_i area_i scale_i / sum."""
    sizes = [1.0, 2.0, 0.5, 3.0, 1.0]
    scales = [1.0, 1.0, 4.0, 0.25, 2.0]
    sides = [False, True, False, True, True]
    lights = [dict(rgb=(0.8, 0.6, 0.3), scale=s, two_sided=t) for s, t in zip(scales, sides)]
    tris = [_tri(3 * i, sz) for i, sz in enumerate(sizes)] + [_tri(20, 1.0)]
    sc = _emitter_scene(lights, tris, list(range(5)) + [-1])
    pmf, tri = O.light_pmf(sc)
    assert list(tri) == [0, 1, 2, 3, 4]
    w = np.array([(2 if t else 1) * (0.5 * sz * sz) * s for sz, s, t in zip(sizes, scales, sides)])
    assert np.allclose(pmf, w / w.sum(), rtol=2e-6)
    assert float(np.sum(pmf, dtype=np.float64)) == pytest.approx(1.0, abs=1e-6)


def _sigmoid_spectrum(c, lam):
    x = np.polyval(c.astype(np.float64), lam)
    return 0.5 + x / (2 * np.sqrt(1 + x * x))


def test_power_sampler_pmf_colour_and_infinite_light():
    """Different colours weigh by the average over the four wavelengths of
    sample_visible(0.5) of L(lambda) / pdf(lambda), L = RGBIlluminantSpectrum
    (scale 2 max(rgb), sigmoid of rgb / scale, times D65); the light scale's
    photometric normalisation uses the illuminant alone
    (rgb_illuminant_spectrum.cu:18-20), the same for every colour. A uniform
    infinite light weighs 4 pi^2 R^2 scale L, R the radius of the scene bounds'
    bounding sphere."""
    lam, pdf = O.sample_visible(0.5)
    lam = lam.astype(np.float64)

    def spectral_weight(rgb):
        rgb = np.asarray(rgb, np.float64)
        s = 2 * rgb.max()
        c = O.rgb_to_sigmoid(rgb / s)
        d65 = np.array([O.spectrum_eval(3, l) for l in lam])
        L = s * _sigmoid_spectrum(c, lam) * d65
        return np.mean(L / pdf)

    cols = [(0.9, 0.2, 0.1), (0.1, 0.3, 0.9), (0.5, 0.5, 0.5)]
    lights = [dict(rgb=c, scale=1.0, two_sided=False) for c in cols]
    tris = [_tri(3 * i, 1.0) for i in range(3)]
    pmf, _ = O.light_pmf(_emitter_scene(lights, tris, [0, 1, 2]))
    w = np.array([spectral_weight(c) for c in cols])
    assert np.allclose(pmf, w / w.sum(), rtol=2e-4)
    # one triangle light and one uniform infinite light of the same colour
    lights = [dict(rgb=(1, 1, 1), scale=3.0, two_sided=False)]
    tris = [_tri(0, 2.0)]
    sky = dict(rgb=(1, 1, 1), scale=0.5, light_index=1)
    pmf, tri = O.light_pmf(_emitter_scene(lights, tris, [0], infinite=[sky]))
    assert list(tri) == [0, -1]
    p = np.asarray(tris[0], np.float64)
    lo, hi = p.min(0), p.max(0)
    R = np.linalg.norm(hi - (lo + hi) / 2)
    w_tri = np.pi * 2.0 * 3.0  # pi * area (2) * scale
    w_sky = 4 * np.pi ** 2 * R * R * 0.5
    assert pmf[0] == pytest.approx(w_tri / (w_tri + w_sky), rel=1e-4)


def _plane_under_emitter(a, h, cam_h, fov, eye_x=0.0, albedo=0.5, depth=1, look_up=False, spp=64, res=16):
    """A grey plane y = 0 (albedo `albedo`) under a two-sided emitter quad
    [-a, a]^2 at y = h with a black material; the camera at height cam_h looks
    straight down at (eye_x, 0, 0) (or straight up at the emitter)."""
    P = [(-50, 0, -50), (50, 0, -50), (50, 0, 50), (-50, 0, 50)]
    E = [(-a, h, -a), (a, h, -a), (a, h, a), (-a, h, a)]
    q = lambda v: [v[0] + v[1] + v[2], v[2] + v[3] + v[0]]
    tri = np.asarray(q(P) + q(E), np.float32).reshape(4, 3, 3)
    eye = (eye_x, cam_h, 0.0)
    look = (eye_x, h + 1, 0.0) if look_up else (eye_x, 0.0, 0.0)
    return dict(eye=eye, look=look, up=(0, 0, 1), fov=fov, width=res, height=res, spp=spp, max_depth=depth,
                tri=tri, tri_material=np.array([0, 0, 1, 1], np.uint32), tri_light=np.array([-1, -1, 0, 0], np.int32),
                albedo=np.array([[albedo] * 3, [0.0] * 3], np.float32),
                lights=[dict(rgb=(1.0, 0.9, 0.7), scale=2.0, two_sided=True)])


def _form_factor_to_centred_square(x, z, a, h):
    """Differential area (normal +y) at (x, 0, z) to the parallel square [-a, a]^2 at height h."""
    def corner(X, Y):
        s = np.sign(X) * np.sign(Y)
        X, Y = abs(X), abs(Y)
        return s / (2 * np.pi) * (X / np.sqrt(1 + X * X) * np.arctan(Y / np.sqrt(1 + X * X)) +
                                  Y / np.sqrt(1 + Y * Y) * np.arctan(X / np.sqrt(1 + Y * Y)))
    xs, zs = ((a - x) / h, (-a - x) / h), ((a - z) / h, (-a - z) / h)
    return corner(xs[0], zs[0]) - corner(xs[0], zs[1]) - corner(xs[1], zs[0]) + corner(xs[1], zs[1])


def _sensor_mean(sc):
    """The image's filter-weighted mean sensor RGB: sum of rgb_sum / sum of weight_sum."""
    _, aux, _ = O.render(sc, threads=8)
    aux = aux.astype(np.float64).reshape(-1, 4)
    return aux[:, :3].sum(0) / aux[:, 3].sum(), aux


@pytest.mark.parametrize("a,eye_x", [(1.0, 0.0), (0.6, 0.5), (4.0, 0.0)])
def test_area_light_nee_mis_closed_form(a, eye_x):
    """One bounce off a Lambertian plane under an emitter: L = albedo * Le * F
    with F the form factor to the emitter, for any mix of light samples and BSDF
    samples with power-heuristic weights. The emitter's own radiance Le is read
    off a render that looks straight at it, so the spectral factors cancel (a
    grey albedo's spectrum is flat); the emitter's material is black, so depth 3
    adds nothing to depth 1."""
    h = 1.0
    Le, _ = _sensor_mean(_plane_under_emitter(a, h, 0.5, 5.0, eye_x=0.0, look_up=True, spp=16))
    F = _form_factor_to_centred_square(eye_x, 0.0, a, h)
    assert 0.1 < F < 0.99
    for depth in (1, 3):
        got, aux = _sensor_mean(_plane_under_emitter(a, h, 0.5, 2.0, eye_x=eye_x, depth=depth, spp=64, res=32))
        want = 0.5 * Le * F
        # noise: the spread of per-pixel estimates (64 samples each)
        px = aux[:, :3] / aux[:, 3:4]
        se = px.std(0) / np.sqrt(len(px))
        assert np.all(np.abs(got - want) < 5 * se + 4e-3 * want), (a, eye_x, depth, got, want, se)


# ------------------------------------------------------------ spectral data
def test_glass_bk7_matches_schott_sellmeier():
    """Schott's published Sellmeier coefficients for N-BK7 (B1..B3, C1..C3 in um^2)."""
    B = (1.03961212, 0.231792344, 1.01046945)
    C = (0.00600069867, 0.0200179144, 103.560653)
    for lam in (380, 400, 450, 486.1, 546.1, 587.6, 656.3, 700, 800):
        l2 = (lam / 1000) ** 2
        n = np.sqrt(1 + sum(b * l2 / (l2 - c) for b, c in zip(B, C)))
        assert O.spectrum_eval(12, lam) == pytest.approx(n, abs=2e-4), lam
    assert O.spectrum_eval(12, 587.6) == pytest.approx(1.5168, abs=1e-4)  # n_d


def test_metal_tables_against_published_optical_constants():
    """Normal-incidence reflectance ((n-1)^2 + k^2) / ((n+1)^2 + k^2) of the named
    metal spectra against handbook values (silver and aluminium bright and flat,
    gold and copper dark in the blue and bright in the red), and a few optical
    constants (Al at 550 nm, Cu at 650 nm)."""
    def R(m, lam):
        n, k = O.spectrum_eval(4 + 2 * m, lam), O.spectrum_eval(5 + 2 * m, lam)
        return ((n - 1) ** 2 + k * k) / ((n + 1) ** 2 + k * k)

    AG, AL, AU, CU = range(4)
    for lam in (450, 550, 650):
        assert 0.90 < R(AG, lam) < 0.99
        assert 0.89 < R(AL, lam) < 0.94
    assert 0.30 < R(AU, 450) < 0.45 and 0.92 < R(AU, 650) < 0.97
    assert 0.50 < R(CU, 450) < 0.62 and 0.92 < R(CU, 650) < 0.97
    assert O.spectrum_eval(4 + 2 * AL, 550) == pytest.approx(0.96, abs=0.05)
    assert O.spectrum_eval(5 + 2 * AL, 550) == pytest.approx(6.69, abs=0.2)
    assert O.spectrum_eval(4 + 2 * CU, 650) == pytest.approx(0.21, abs=0.03)
    assert O.spectrum_eval(5 + 2 * CU, 650) == pytest.approx(3.67, abs=0.15)


def test_cie_1931_published_values():
    """CIE 1931 2-degree observer at tabulated points, and D65 = 100 at 560 nm
    before normalisation (its normalised value is 100 / its Y integral ratio)."""
    assert O.spectrum_eval(0, 600) == pytest.approx(1.0622, abs=1e-4)
    assert O.spectrum_eval(1, 555) == pytest.approx(1.0000, abs=1e-4)
    assert O.spectrum_eval(2, 450) == pytest.approx(1.7721, abs=1e-4)
    assert O.spectrum_eval(0, 450) == pytest.approx(0.3362, abs=1e-4)
    assert O.spectrum_eval(1, 500) == pytest.approx(0.3230, abs=1e-4)
