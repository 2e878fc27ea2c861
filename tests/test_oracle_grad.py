"""The oracle's FP64 path-replay gradient against central differences (CPU).

Loss = sum over pixels and channels of (film - target)^2, target rendered from
parameters perturbed by up to 10 %. Albedo derivatives pass through the
RGB-to-spectrum table (piecewise trilinear) and every later path contribution;
emitter-scale derivatives through the emitted radiance and the power light
sampler's pmf inside the MIS weights and the light choice's 1/pmf, checked with
the light choice frozen (the select_scale knob) so that no sample's discrete
choice flips between the two renders of a difference. A dropped term (the MIS-weight
derivative, the pmf's dependence on the other emitters, a vertex's own NEE)
moves the gradient by far more than the tolerances below.
"""
import numpy as np
import pytest

import oracle as O
import scenes

H = 1e-3


def loss_of(sc, target):
    f, _, _ = O.render(sc, threads=8)
    return float(np.sum((f.astype(np.float64) - target) ** 2))


def target_for(sc):
    from pbrtm.scenes import perturbed

    t, _, _ = O.render(perturbed(sc), threads=8)
    return t


def test_table_jacobian_matches_differences():
    rng = np.random.default_rng(3)
    for _ in range(20):
        rgb = rng.uniform(0.05, 0.95, 3).astype(np.float32)
        J = O.rgb_to_sigmoid_jacobian(rgb)
        for k in range(3):
            a = rgb.astype(np.float64)
            b = a.copy()
            a[k] += 1e-4
            b[k] -= 1e-4
            fd = (O.rgb_to_sigmoid(a.astype(np.float32)).astype(np.float64)
                  - O.rgb_to_sigmoid(b.astype(np.float32)).astype(np.float64)) / 2e-4
            assert np.abs(fd - J[:, k]).max() <= 5e-3 * (np.abs(J[:, k]).max() + 1e-3)


def test_loss_and_film():
    sc = scenes.cornell_emitters(16, 16, spp=4)
    target = target_for(sc)
    loss, d_rgb, d_scale, film = O.render_grad(sc, target)
    ref, _, _ = O.render(sc, threads=8)
    assert np.array_equal(film, ref)
    assert abs(loss - float(np.sum((ref.astype(np.float64) - target) ** 2))) <= 1e-9 * loss
    assert d_rgb.shape == (len(sc["albedo"]), 3) and d_scale.shape == (3,)


def test_albedo_gradient_matches_central_differences():
    sc = scenes.cornell_emitters(24, 24, spp=16)
    target = target_for(sc)
    loss, d_rgb, _, _ = O.render_grad(sc, target)
    scale = float(np.abs(d_rgb).max())
    checked = 0
    for m in (2, 4, 6, 8, 11, 13, 25, 30):
        for c in range(3):
            hi, lo = dict(sc), dict(sc)
            hi["albedo"], lo["albedo"] = sc["albedo"].copy(), sc["albedo"].copy()
            hi["albedo"][m, c] += H
            lo["albedo"][m, c] -= H
            fd = (loss_of(hi, target) - loss_of(lo, target)) / (2 * H)
            assert abs(fd - d_rgb[m, c]) <= 0.03 * abs(fd) + 0.01 * scale, (m, c, fd, d_rgb[m, c])
            checked += 1
    assert checked == 24


def test_emitter_gradient_with_frozen_light_choice():
    """With the light choice frozen (no sample's choice flips), central differences
    measure the fixed-seed estimator's pathwise derivative: the light pmf attached
    in the MIS weights and in the choice's 1/pmf."""
    sc = scenes.cornell_emitters(32, 32, spp=16)
    sc["select_scale"] = np.array([l["scale"] for l in sc["lights"]], np.float32)
    target = target_for(dict(sc, select_scale=None))
    _, _, d_att, _ = O.render_grad(sc, target)
    _, _, d_held, _ = O.render_grad(sc, target, hold_selection=True)
    for li in range(3):
        h = H * sc["lights"][li]["scale"]
        hi = dict(sc, lights=[dict(l) for l in sc["lights"]])
        lo = dict(sc, lights=[dict(l) for l in sc["lights"]])
        hi["lights"][li]["scale"] += h
        lo["lights"][li]["scale"] -= h
        fd = (loss_of(hi, target) - loss_of(lo, target)) / (2 * h)
        assert abs(fd - d_att[li]) <= 2e-3 * abs(fd) + 1e-3 * float(np.abs(d_att).max()), (li, fd, d_att[li])
    # holding the choice's 1/pmf drops a term the differences see
    assert np.abs(d_att - d_held).max() > 1e-2 * np.abs(d_att).max()


def test_single_emitter_scale_is_linear():
    """One emitter: the pmf is 1 whatever the scale, so the loss is a quadratic in
    the scale and the gradient is exact."""
    sc = scenes.cornell34(24, 24, spp=16, max_depth=4)
    sc["materials"] = [dict(m) if m["type"] == "diffuse" else {"type": "diffuse", "rgb": (0.5, 0.6, 0.7)}
                       for m in sc["materials"]]
    target = target_for_lights(sc)
    _, _, d_scale, _ = O.render_grad(sc, target)
    s = sc["lights"][0]["scale"]
    vals = []
    for f in (1 - 1e-2, 1 + 1e-2):
        vals.append(loss_of(dict(sc, lights=[dict(sc["lights"][0], scale=s * f)]), target))
    fd = (vals[1] - vals[0]) / (2e-2 * s)
    assert abs(fd - d_scale[0]) <= 1e-3 * abs(fd)


def target_for_lights(sc):
    t, _, _ = O.render(dict(sc, lights=[dict(sc["lights"][0], scale=sc["lights"][0]["scale"] * 1.1)]), threads=8)
    return t
