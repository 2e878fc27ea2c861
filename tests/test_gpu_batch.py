"""Batched rendering (pbrtm_render_batch) against one-by-one renders.

Every job of a batch shares one pool of path slots with the others; its film
must be bit-identical to pbrtm_render of the same scene and parameters (whose
films are in turn bit-identical to the oracle's, test_gpu_pathtracer.py), and
the round count the batch reports for it must equal the one-by-one render's.
"""
import numpy as np
import pytest

import oracle as O
import pbrtm
import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def mixed_jobs():
    """(scene dict, spp, max_depth, one-by-one pool size) covering refills, depths and materials."""
    rough = scenes.cornell34(48, 40)
    rough["materials"] = list(rough["materials"])
    rough["materials"][3] = {"type": "conductor", "metal": "Cu", "roughness": 0.2, "remap": True}
    rough["materials"][0] = {"type": "dielectric", "eta": "glass-BK7", "roughness": 0.0}
    return [
        (scenes.cornell34(40, 32), 4, 5, 0),
        (rough, 1, 6, 1000),
        (scenes.soup_room(20000, 64, 64), 16, 3, 5000),
        (scenes.soup_room(5000, 33, 17), 9, 8, 0),
        (scenes.cornell34(24, 24), 1, 12, 257),
        (scenes.soup_room(20000, 64, 64), 1, 1, 0),
    ]


@pytest.fixture(scope="module")
def rendered(ctx):
    jobs = mixed_jobs()
    handles = {}
    out = []
    for sc, spp, depth, vpool in jobs:
        key = id(sc)
        handles[key] = ctx.scene(sc)
        rgb, aux, st = handles[key].render(spp=spp, max_depth=depth, pool_size=vpool)
        out.append((rgb, aux, st))
        log(f"one-by-one {sc['width']}x{sc['height']} spp {spp} depth {depth} pool {vpool}: rounds {st['rounds']}")
    yield jobs, handles, out
    for h in handles.values():
        h.free()


@pytest.mark.parametrize("pool,graph,rpl,solo", [(0, True, 0, 0), (777, True, 3, 0xFFFFFFFF), (4096, False, 0, 20000),
                                                   (1 << 16, True, 1, 1)])
def test_batch_equals_one_by_one(ctx, rendered, pool, graph, rpl, solo):
    """solo: jobs with at least that many paths render alone before the shared pool
    (0 = the default, never; 20000 = some; 1 = all)."""
    jobs, handles, ref = rendered
    spec = [(handles[id(sc)], spp, depth, vpool) for sc, spp, depth, vpool in jobs]
    log(f"batch of {len(spec)} jobs, pool {pool}, graph {graph}, rounds per launch {rpl}, solo {solo}")
    rgbs, auxs, rounds, st = ctx.render_batch(spec, pool_size=pool, rounds_per_launch=rpl, graph=graph, aux=True,
                                              solo_paths=solo)
    log(f"  {st['ms_total']:.2f} ms, physical rounds {st['rounds']}, launches {st['launches']}, pool {st['pool_size']}, "
        f"solo jobs {st['solo_jobs']}")
    if solo in (0, 0xFFFFFFFF):
        assert st["solo_jobs"] == 0
    elif solo == 1:
        assert st["solo_jobs"] == len(spec)
    elif solo == 20000:
        assert 0 < st["solo_jobs"] < len(spec)
    for k, ((rgb0, aux0, st0), rgb, aux, r) in enumerate(zip(ref, rgbs, auxs, rounds)):
        assert r == st0["rounds"], (k, r, st0["rounds"])
        assert np.array_equal(aux.view(np.uint32), aux0.view(np.uint32)), k
        assert np.array_equal(rgb.view(np.uint32), rgb0.view(np.uint32)), k
    assert st["paths"] == sum(sc["width"] * sc["height"] * spp for sc, spp, _, _ in jobs)


def test_batch_job_against_oracle(ctx, rendered):
    """One job of the batch straight against the oracle (same bar as the single renders)."""
    jobs, handles, _ = rendered
    sc, spp, depth, vpool = jobs[2]
    rgbs, _, rounds, _ = ctx.render_batch([(handles[id(sc)], spp, depth, vpool)], pool_size=3000)
    orgb, _, ost = O.render(dict(sc, spp=spp, max_depth=depth), pool_size=vpool or (1 << 21), threads=16)
    assert rounds[0] == ost["rounds"]
    assert np.array_equal(rgbs[0].view(np.uint32), orgb.view(np.uint32))


def test_batch_errors(ctx, rendered):
    jobs, handles, _ = rendered
    h = handles[id(jobs[0][0])]
    with pytest.raises(pbrtm.PbrtmError):
        ctx.render_batch([(h, 3, 5, 0)])
    with pytest.raises(pbrtm.PbrtmError):
        ctx.render_batch([(h, 4, 4001, 0)])
    with pytest.raises(pbrtm.PbrtmError):
        ctx.render_batch([])


def test_batch_autotune_settles_on_the_fastest(ctx, rendered):
    """PBRTM_BATCH_AUTOTUNE: five calls each try one (pool size, traversal grid)
    candidate (three pool sizes at the full grid, then the fastest pool at 3 and 2
    quarters of it), the sixth and later use the fastest measured; every call's
    films equal the one-by-one films. Another batch (a different job list) starts
    its own tuning."""
    jobs, handles, ref = rendered
    spec = [(handles[id(sc)], spp, depth, vpool) for sc, spp, depth, vpool in jobs]
    seen = []
    for call in range(7):
        rgbs, _, rounds, st = ctx.render_batch(spec, autotune=True)
        log(f"call {call}: pool {st['pool_size']} quarters {st['grid_quarters']} state {st['tune_state']} "
            f"{st['ms_total']:.2f} ms")
        for k, ((rgb0, _, st0), rgb, r) in enumerate(zip(ref, rgbs, rounds)):
            assert r == st0["rounds"] and np.array_equal(rgb.view(np.uint32), rgb0.view(np.uint32)), (call, k)
        seen.append((st["pool_size"], st["grid_quarters"], st["tune_state"], st["ms_total"]))
    assert [s[2] for s in seen] == [1, 1, 1, 1, 1, 2, 2]
    assert [s[1] for s in seen[:3]] == [4, 4, 4] and [s[1] for s in seen[3:5]] == [3, 2]
    fastest = min(seen[:5], key=lambda s: s[3])
    assert seen[5][:2] == fastest[:2] and seen[6][:2] == fastest[:2]
    _, _, _, st = ctx.render_batch(spec[:2], autotune=True)
    assert st["tune_state"] == 1
