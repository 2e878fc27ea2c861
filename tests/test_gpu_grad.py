"""pbrtm_render_grad (forward render + path replay on the GPU) against the
oracle's FP64 path replay, and against central differences of the GPU's own
loss at a larger size."""
import numpy as np
import pytest

import oracle as O
import pbrtm
import scenes
from pbrtm.scenes import grad_room, perturbed

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def close(gpu, ref, rel, label):
    scale = float(np.abs(ref).max())
    err = np.abs(gpu.astype(np.float64) - ref)
    worst = float((err / (rel * np.abs(ref) + rel * scale)).max())
    log(f"{label}: max |gpu - oracle| {float(err.max()):.3g} (scale {scale:.3g}), worst/tolerance {worst:.3g}")
    assert worst <= 1.0, label


@pytest.mark.parametrize("which", ["cornell", "room"])
def test_gradient_matches_oracle(ctx, which):
    if which == "cornell":
        sc = scenes.cornell_emitters(32, 32, spp=16)
    else:
        sc = grad_room(300, 40, 40, spp=16, max_depth=4)
    target, _, _ = O.render(perturbed(sc), threads=16)
    loss_o, d_rgb_o, d_scale_o, film_o = O.render_grad(sc, target, threads=16)
    s = ctx.scene(sc)
    try:
        loss, d_alb, d_sc, film, st = s.render_grad(target, spp=sc["spp"], max_depth=sc["max_depth"])
    finally:
        s.free()
    log(f"{which}: loss gpu {loss:.9g} oracle {loss_o:.9g}; forward {st['ms_forward']:.2f} ms, "
        f"replay {st['ms_replay']:.2f} ms")
    assert np.array_equal(film.view(np.uint32), film_o.view(np.uint32))
    assert abs(loss - loss_o) <= 1e-9 * loss_o
    close(d_alb, d_rgb_o, 2e-3, f"{which} d_albedo")
    close(d_sc, d_scale_o, 2e-3, f"{which} d_scale")


def test_gradient_central_differences_on_gpu(ctx):
    """64k-triangle room at 256x256, 16 spp, depth 4: the 12 wall albedos' 36
    derivatives against central differences (h = 1e-3) of the GPU loss. Albedos
    move no discrete choice, so the differences measure the replayed derivative
    itself. The 8 emitter scales' (h = 5 % of the scale) are only bounded loosely:
    a scale moves the light choice's cdf, and the samples whose choice flips add
    boundary terms (a jump in a pixel enters the L2 loss squared) that the
    fixed-seed derivative leaves out by definition; the emitter derivative is
    pinned against the oracle, and the oracle's against differences with the
    choice frozen."""
    sc = grad_room(65536, 256, 256, spp=16, max_depth=4)
    target_sc = perturbed(sc)
    t = ctx.scene(target_sc)
    target, _, _ = t.render(spp=16, max_depth=4)
    t.free()

    def loss_of(alb=None, lights=None):
        s2 = ctx.scene(dict(sc, albedo=sc["albedo"] if alb is None else alb,
                            lights=sc["lights"] if lights is None else lights))
        f, _, _ = s2.render(spp=16, max_depth=4)
        s2.free()
        return float(np.sum((f.astype(np.float64) - target) ** 2))

    s = ctx.scene(sc)
    loss, d_alb, d_sc, _, st = s.render_grad(target, spp=16, max_depth=4)
    s.free()
    log(f"loss {loss:.6g}, forward {st['ms_forward']:.1f} ms, replay {st['ms_replay']:.1f} ms")
    walls = range(65536, 65536 + 12)
    gmax = float(np.abs(d_alb[65536:65548]).max())
    for m in walls:
        for c in range(3):
            hi, lo = sc["albedo"].copy(), sc["albedo"].copy()
            hi[m, c] += 1e-3
            lo[m, c] -= 1e-3
            fd = (loss_of(alb=hi) - loss_of(alb=lo)) / 2e-3
            log(f"wall {m - 65536} c {c}: fd {fd:.6g} ad {d_alb[m, c]:.6g}")
            assert abs(fd - d_alb[m, c]) <= 0.02 * abs(fd) + 2e-3 * gmax, (m, c)
    smax = float(np.abs(d_sc).max())
    fds = []
    for li in range(8):
        h = 5e-2 * sc["lights"][li]["scale"]
        hi = [dict(l) for l in sc["lights"]]
        lo = [dict(l) for l in sc["lights"]]
        hi[li]["scale"] += h
        lo[li]["scale"] -= h
        fd = (loss_of(lights=hi) - loss_of(lights=lo)) / (2 * h)
        fds.append(fd)
        log(f"emitter {li}: fd {fd:.6g} ad {d_sc[li]:.6g}")
        assert abs(fd - d_sc[li]) <= 0.1 * smax, li
    fds = np.array(fds)
    assert np.linalg.norm(fds - d_sc) <= 0.1 * np.linalg.norm(fds)


def test_gradient_central_differences_full_config(ctx):
    """The gradient benchmark's own configuration: 1M-triangle room (seed 0xD1FF),
    512x512, 16 spp, depth 4, target from the seeded +-10 % parameters. 32
    albedo parameters chosen by seed (PCG32, sequence 0xD1FF, uniform over the
    3M per-triangle albedo channels, skipping those no path reaches: their
    derivative is exactly zero, which the first such draws also check against
    their differences) against central differences with h = 1e-3 on identical
    sampler streams.

    An albedo reaches the renderer through the RGB-to-spectrum table, which is
    trilinear per cell (rgb_to_spectrum_table.cu): across a cell boundary the
    map has a kink, and a central difference over it measures a secant a few
    per cent off the in-cell derivative the gradient holds. Draws whose +-h
    interval crosses a boundary (the table's coefficients are not affine on it)
    are skipped.

    A difference also carries the films' own FP32 rounding: every film value
    the step changes is off by a few units in the last place (the sample's
    radiance, its sensor-RGB conversion, the pixel's weighted sums and the final
    division each round once), which moves the loss difference by about
    sum 2 |f - t| ulp(f) / 2h over those values. Each comparison allows 16 such
    units in quadrature besides 2 % of the difference."""
    from pbrtm.scenes import Pcg32

    sc = grad_room(1_000_000 - 28, 512, 512, spp=16, max_depth=4)
    t = ctx.scene(perturbed(sc))
    target, _, _ = t.render(spp=16, max_depth=4)
    t.free()
    s = ctx.scene(sc)
    loss, d_alb, _, _, st = s.render_grad(target, spp=16, max_depth=4)
    s.free()
    log(f"loss {loss:.6g}, forward {st['ms_forward']:.1f} ms, replay {st['ms_replay']:.1f} ms, "
        f"total {st['ms_total']:.1f} ms")
    n = len(sc["albedo"])

    def film_of(m, c, h):
        alb = sc["albedo"].copy()
        alb[m, c] += h
        s2 = ctx.scene(dict(sc, albedo=alb))
        f, _, _ = s2.render(spp=16, max_depth=4)
        s2.free()
        return f

    def difference(m, c, h=1e-3):
        """Central difference of the loss and its FP32 rounding allowance."""
        fh, fl = film_of(m, c, h), film_of(m, c, -h)
        lh = float(np.sum((fh.astype(np.float64) - target) ** 2))
        ll = float(np.sum((fl.astype(np.float64) - target) ** 2))
        ch = fh != fl
        w = 2 * np.abs(0.5 * (fh[ch].astype(np.float64) + fl[ch]) - target[ch])
        ulp = np.spacing(np.abs(fh[ch])).astype(np.float64) + np.spacing(np.abs(fl[ch]))
        return (lh - ll) / (2 * h), 16 * float(np.sqrt(np.sum((w * ulp) ** 2))) / (2 * h)

    def affine_on_interval(m, c, h=1e-3):
        co = []
        for d in (-h, 0.0, h):
            a = sc["albedo"][m].astype(np.float32).copy()
            a[c] += np.float32(d)
            co.append(pbrtm.rgb_to_sigmoid(a).astype(np.float64))
        d1, d2 = co[1] - co[0], co[2] - co[1]
        return bool(np.all(np.abs(d2 - d1) <= 2e-3 * np.maximum(np.abs(d1), np.abs(d2)) + 1e-9))

    rng = Pcg32(0xD1FF)
    picked, zeros, draws, kinked = [], 0, 0, 0
    while len(picked) < 32:
        draws += 1
        assert draws < 20000, "too few reachable parameters"
        m, c = rng.next_u32() % n, rng.next_u32() % 3
        if d_alb[m, c] != 0.0 and not affine_on_interval(m, c):
            kinked += 1
            continue
        if d_alb[m, c] == 0.0:
            if zeros < 4:  # unreached: the differences must vanish too
                fd, _ = difference(m, c)
                log(f"unreached {m} c {c}: fd {fd:.3g}")
                assert fd == 0.0, (m, c)
                zeros += 1
            continue
        picked.append((m, c))
    fds, tols = [], []
    for m, c in picked:
        fd, tol = difference(m, c)
        fds.append(fd)
        tols.append(tol)
        log(f"albedo {m} c {c}: fd {fd:.6g} ad {d_alb[m, c]:.6g} rounding allowance {tol:.3g}")
    log(f"{draws} draws: {len(picked)} checked, {kinked} across a table cell boundary skipped")
    ad = np.array([d_alb[m, c] for m, c in picked], np.float64)
    fds, tols = np.array(fds), np.array(tols)
    bad = np.abs(fds - ad) > 0.02 * np.abs(fds) + tols
    assert not bad.any(), [picked[i] for i in np.nonzero(bad)[0]]
    assert np.linalg.norm(fds - ad) <= 0.01 * np.linalg.norm(fds) + np.linalg.norm(tols)


def test_gradient_errors(ctx):
    sc = scenes.cornell_emitters(16, 16, spp=4)
    s = ctx.scene(sc)
    try:
        with pytest.raises(pbrtm.PbrtmError):
            s.render_grad(np.zeros((16, 16, 3), np.float32), spp=3)
    finally:
        s.free()
