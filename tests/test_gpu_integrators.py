"""GPU parity of the megakernel integrators (megakernelpath, ambientocclusion,
surfacenormal; base/megakernel_integrator.cu:14-147) through pbrtm_render against
the oracle's per-pixel, sample-by-sample restatement.

The library runs them on the wavefront's path pool, one round per vertex, and
adds each pixel's samples to the film in sample order as the megakernel does;
the oracle follows the reference's loops path by path. Same sampler streams, same
FP32 sequence: the films and the raw sums are bit-identical, and the tests hold
them to that (measured so on the GPU), on top of the 1e-4 relative per-pixel bar
the wavefront tests enforce."""
import numpy as np
import pytest

import oracle as O
import pbrtm
import scenes
from test_gpu_pathtracer import compare as compare_rel
from test_gpu_pathtracer import cornell_materials, log

pytestmark = pytest.mark.gpu


def compare(gpu, ref, label):
    assert compare_rel(gpu, ref, label) == 1.0, f"{label}: not bit-identical"


def render_both(ctx, sc, integrator, spp, depth, pool=0):
    log(f"{integrator}: {sc['width']}x{sc['height']} spp {spp} depth {depth} pool {pool}")
    s = ctx.scene(sc)
    try:
        rgb, aux, st = s.render(spp=spp, max_depth=depth, pool_size=pool, integrator=integrator)
    finally:
        s.free()
    orgb, oaux, ost = O.render_megakernel(dict(sc, spp=spp, max_depth=depth), integrator, threads=16)
    log(f"  gpu {st['ms_total']:.2f} ms rounds {st['rounds']} rays {st['rays']} shadow {st['shadow_rays']}; "
        f"oracle rays {ost['rays']} shadow {ost['shadow_rays']}")
    if integrator != "megakernelpath":  # one closest-hit ray and at most one any-hit ray per sample
        assert st["rays"] == ost["rays"] and st["shadow_rays"] == ost["shadow_rays"]
    return rgb, aux, orgb, oaux


@pytest.mark.parametrize("integrator", ["megakernelpath", "ambientocclusion", "surfacenormal"])
def test_cornell(ctx, integrator):
    sc = scenes.cornell34(48, 48)
    rgb, aux, orgb, oaux = render_both(ctx, sc, integrator, 4, 5)
    compare(aux, oaux, f"aux {integrator}")
    compare(rgb, orgb, f"film {integrator}")
    assert np.isfinite(rgb).all() and np.any(rgb)


@pytest.mark.parametrize("integrator", ["megakernelpath", "ambientocclusion", "surfacenormal"])
def test_pool_refills_non_square(ctx, integrator):
    """A pool of 1000 slots for 40 x 24 x 9 paths: slots refill every round, and the
    single-vertex paths' retired rays must not end the render early."""
    sc = scenes.cornell34(40, 24)
    rgb, aux, orgb, oaux = render_both(ctx, sc, integrator, 9, 4, pool=1000)
    compare(aux, oaux, f"aux {integrator} pool 1000")
    compare(rgb, orgb, f"film {integrator} pool 1000")
    assert np.all(aux[..., 3] != 0)


def test_megakernel_path_russian_roulette(ctx):
    sc = scenes.cornell34(32, 32)
    rgb, aux, orgb, oaux = render_both(ctx, sc, "megakernelpath", 4, 12)
    compare(aux, oaux, "aux megakernelpath depth 12")
    compare(rgb, orgb, "film megakernelpath depth 12")


@pytest.mark.parametrize("variant", ["rough", "glass", "smooth_glass"])
@pytest.mark.parametrize("depth", [2, 6])
def test_megakernel_path_materials(ctx, variant, depth):
    """Rough conductors and dielectrics; BK7 glass ends the secondary wavelengths in
    get_bsdf, also at the last vertex, where the megakernel stops after it."""
    sc = cornell_materials(variant)
    rgb, aux, orgb, oaux = render_both(ctx, sc, "megakernelpath", 4, depth)
    compare(aux, oaux, f"aux {variant} depth {depth}")
    compare(rgb, orgb, f"film {variant} depth {depth}")


@pytest.mark.parametrize("integrator", ["megakernelpath", "ambientocclusion"])
def test_soup_room(ctx, integrator):
    sc = scenes.soup_room(100_000, width=64, height=64)
    rgb, aux, orgb, oaux = render_both(ctx, sc, integrator, 4, 5)
    compare(aux, oaux, f"aux soup {integrator}")
    compare(rgb, orgb, f"film soup {integrator}")


def test_wavefront_only_entry_points(ctx):
    """Gradients and batches render with the wavefront path integrator only."""
    sc = scenes.cornell34(8, 8)
    s = ctx.scene(sc)
    try:
        with pytest.raises(pbrtm.PbrtmError, match="wavefront"):
            s.render_grad(np.zeros((8, 8, 3), np.float32), spp=4, max_depth=2, integrator="ambientocclusion")
    finally:
        s.free()


def test_cli_integrator_option(tmp_path):
    """pbrtm-render --integrator overrides the file's Integrator (scene_builder.cu:366-371)."""
    import os
    import subprocess
    from PIL import Image
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "pbr-rust_amd", "pbrtm-render")
    fixture = os.path.join(root, "tests", "fixtures", "cornell34.pbrt")
    text = open(fixture).read().replace("256", "64")
    p = tmp_path / "c.pbrt"
    p.write_text(text)
    out = tmp_path / "c.png"
    r = subprocess.run([exe, str(p), "--integrator", "ambientocclusion", "--outfile", str(out)],
                       capture_output=True, text=True, timeout=120)
    log(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    img = np.asarray(Image.open(out).convert("RGBA"))
    orgb, _, _ = O.render_megakernel(scenes.cornell34(64, 64), "ambientocclusion", threads=16)
    assert np.array_equal(img, O.srgb8(orgb))
