"""HLBVH build + traversal benchmark on one MI355X.

Workload (one step): build the HLBVH over 10M seeded FP32 triangles already
resident in HBM (centroids uniform in the unit cube, edge lengths log-normal
mu=-6, sigma=0.5), trace 1920x1080 primary rays (closest hit), spawn one
cosine-hemisphere diffuse bounce per hit, trace the bounce rays (closest hit).

Prints one JSON line: the headline metric is the HLBVH build time at 10M
triangles (lower is better); traversal throughput, per-phase times, the
roofline of the dominant kernel and a CPU baseline (the oracle, a host-only
restatement of the reference's build, timed on a bounded sample) ride along.

The "pathtracer" object measures the wavefront path tracer on the
displaced-sphere grid (2,097,218 triangles; diffuse, rough copper and glass;
64 emissive triangles) at 1920x1080, 64 spp stratified, max depth 8:
Msamples/s = camera samples / render time, the scene and its HLBVH resident
in HBM before the timed renders. Its roofline counts the render's algorithmic
bytes: path state read and written once per live slot and round, queue
entries, rays and hits, and the traversal's node and triangle bytes from an
untimed counting render of the same image. Its CPU baseline is the oracle's
render of the same scene, spp and depth on 1/16 of the pixels. The
"integrators" object measures the megakernel integrators (megakernelpath,
ambientocclusion) on the same scene at 16 spp the same way.

CPU baselines run on the host's cores: min(hardware_concurrency, 16) threads
(16 = the host share of one GPU on the machines this runs on); the CPU model
and the core counts ride along in "host".
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pbr-rust_amd"))

HBM_PEAK_GBS = 8000.0
WNODE_BYTES = 132    # 8-wide quantized node (128 B) + its octant order word (4 B), per visit
TRI_BYTES = 48       # leaf-ordered triangle, 9 floats in a 48-byte slot
RAY_BYTES = 32
BUILD_BYTES_PER_TRI = 216  # algorithmic bytes of an HLBVH build per triangle
# Path state of one pool slot (pathtrace.hip State): path id, flags, sample
# dimension (4 B each), RNG state (8 B), wavelengths, their pdfs, L, beta, the
# NEE radiance, (filter weight, bsdf pdf, eta scale) and the previous vertex
# (16 B each).
PATH_STATE_BYTES = 3 * 4 + 8 + 7 * 16
PATH_HIT_BYTES = 16  # prim, t and two barycentrics read back by the shading stage
QUEUE_BYTES = 4
FILM_RECORD_BYTES = 16 + 4  # per path: sensor RGB + filter weight, finishing round

# HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE x 2.0 + WRITE_SIZE
# x 1.0 in KiB per dispatch, the factors read off a 256 MiB elementwise copy run
# with --calibrate-copy), for the default workload. Taken from the profile named
# beside it, not measured by this run.
TRAFFIC = {
    "trace_wide_kernel": (2_014_720_512, "profiles/trace_wide_hlbvh10m_1080p_pmc_traffic.md"),
    "build": (13_166_864_448, "profiles/trace_wide_hlbvh10m_1080p_pmc_traffic.md"),
}


def note(msg):
    """Progress on stderr (the JSON line is the only stdout output)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """CPU model (/proc/cpuinfo), hardware_concurrency, the cores this process may
    use, and the thread count the CPU baselines use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    hc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = hc
    return {"model": model, "hardware_concurrency": hc, "affinity": avail, "threads": max(1, min(hc, avail, 16))}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tris", type=int, default=10_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--cpu-sample-tris", type=int, default=0,
                    help="build the CPU baseline on this many triangles (0: the full workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pt", action="store_true", help="skip the path-tracer measurement")
    ap.add_argument("--pt-steps", type=int, default=2)
    ap.add_argument("--pt-spp", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-render measurement")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiable-render measurement")
    ap.add_argument("--grad-steps", type=int, default=2)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (for scaling profiler byte counters)")
    return ap.parse_args()


def cpu_baseline(args, tri, host):
    """The oracle's build of the benchmark's own triangles (the full workload
    unless --cpu-sample-tris caps it; then scaled linearly, which slightly
    favours the CPU), threaded as the reference's host build: parallel Morton
    pass and sort, top-level SAH subtrees on threads of their own, parallel
    bottom splits. Best of two runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    n = args.tris if args.cpu_sample_tris <= 0 else min(args.cpu_sample_tris, args.tris)
    sample = tri if n == args.tris else tri[:n]
    runs = []
    for _ in range(2):
        t0 = time.perf_counter()
        b = oracle.OracleBvh(sample, threads=host["threads"])
        runs.append((time.perf_counter() - t0) * 1e3)
        del b
    ms = min(runs)
    scaled = n != args.tris
    return {
        "value": round(ms * args.tris / n, 1),
        "unit": "ms",
        "cores": host["threads"],
        "kind": "port",
        "sample": (f"oracle (host C++ restatement of hlbvh.cu) build of the benchmark's {n} triangles with "
                   f"{host['threads']} threads on {host['model']}, best of 2: {ms:.0f} ms"
                   + (f", scaled x{args.tris / n:g} to {args.tris}" if scaled else ", unscaled")),
    }


def render_bytes(st, paths, pixels):
    """Algorithmic HBM bytes of one wavefront render, by stage, from its ray counts
    and the counting traversal's totals: every closest-hit ray is one live slot in
    one round, whose path state is read and written once by the shading stage,
    which also writes the ray (32 B) and its queue entry and reads the hit (16 B);
    the trace reads the ray and queue entry, writes the hit (8 B) and reads 132 B
    per node visit and 48 B per triangle test. Shadow rays: written with their
    queue entry by shading, read by the trace, which writes one occlusion word and
    reads nodes and triangles as above. Film: one record and round word per path
    written and read, four floats per pixel written."""
    rays, srays = st["rays"], st["shadow_rays"]
    (cv, _, ct), (sv, _, stt) = st["closest_counts"], st["shadow_counts"]
    trace = rays * (RAY_BYTES + QUEUE_BYTES + 8) + cv * WNODE_BYTES + ct * TRI_BYTES
    shadow = srays * (RAY_BYTES + QUEUE_BYTES + 4) + sv * WNODE_BYTES + stt * TRI_BYTES
    shade = (rays * (2 * PATH_STATE_BYTES + PATH_HIT_BYTES + RAY_BYTES + QUEUE_BYTES)
             + srays * (RAY_BYTES + QUEUE_BYTES + 4))
    film = paths * 2 * FILM_RECORD_BYTES + pixels * 16
    return {"trace": trace, "shadow": shadow, "shade": shade, "film": film}


def render_roofline(b, runs):
    """Roofline of a render: its algorithmic bytes over the render time, and per
    stage over that stage's measured time."""
    import numpy as np

    ms = float(np.mean([r["ms_total"] for r in runs]))
    total = sum(b.values())
    stage_ms = {"trace": "ms_trace", "shadow": "ms_shadow", "shade": "ms_shade", "film": "ms_film"}
    stages = {}
    for k, key in stage_ms.items():
        t = float(np.mean([r[key] for r in runs]))
        gbs = b[k] / (t * 1e-3) / 1e9 if t > 0 else 0.0
        stages[k] = {"bytes": int(b[k]), "ms": round(t, 2), "gbs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    gbs = total / (ms * 1e-3) / 1e9
    return {"kernel": "whole render (trace_wide_kernel, shade_kernel, compaction, film)", "bound": "hbm",
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": None, "bytes_per_render": int(total), "stages": stages,
            "algorithmic_bytes": f"per closest ray: {2 * PATH_STATE_BYTES} B path state + {PATH_HIT_BYTES} B hit "
                                 f"+ 2 x ({RAY_BYTES} B ray + {QUEUE_BYTES} B queue) + 8 B hit out; per shadow ray: "
                                 f"2 x ({RAY_BYTES} + {QUEUE_BYTES}) + 4 + 4 B; {WNODE_BYTES} B per node visit, "
                                 f"{TRI_BYTES} B per triangle test (counted); {2 * FILM_RECORD_BYTES} B per path "
                                 "+ 16 B per pixel of film"}


def pathtracer(args, ctx, pbrtm, host):
    """Msamples/s of the wavefront path tracer at 1920x1080, 64 spp, depth 8 (sphere grid)."""
    import numpy as np
    from pbrtm import scenes

    W, H, spp, depth = 1920, 1080, args.pt_spp, 8
    sc = scenes.sphere_grid(width=W, height=H, spp=spp, max_depth=depth)
    scene = ctx.scene(sc)
    try:
        scene.render(spp=spp, max_depth=depth)  # warm-up: allocates the path pool and records
        runs = []
        t0 = time.perf_counter()
        for _ in range(args.pt_steps):
            _, _, st = scene.render(spp=spp, max_depth=depth)
            runs.append(st)
        wall = (time.perf_counter() - t0) / args.pt_steps
        # the same render with a 2^22-slot pool: each pixel then sums the same samples
        # in another round order, so it is reported beside the reference-pool figure,
        # not as it
        big = 4 << 20
        scene.render(spp=spp, max_depth=depth, pool_size=big)
        ms_big = float(np.mean([scene.render(spp=spp, max_depth=depth, pool_size=big)[2]["ms_total"]
                                for _ in range(args.pt_steps)]))
        # untimed: the same render with the counting traversal (same films and rays)
        counted = scene.render(spp=spp, max_depth=depth, count=True)[2]
    finally:
        scene.free()
    ms = float(np.mean([r["ms_total"] for r in runs]))
    last = runs[-1]
    out = {
        "metric": "Msamples/s at 1080p 64spp",
        "value": round(W * H * spp / (ms * 1e3), 2),
        "unit": "Msamples/s",
        "higher_is_better": True,
        "ms_per_image": round(ms, 2),
        "wall_ms_per_image": round(wall * 1e3, 2),
        "stages_ms": {k: round(float(np.mean([r[k] for r in runs])), 2)
                      for k in ("ms_trace", "ms_shadow", "ms_shade", "ms_film")},
        "rounds": last["rounds"], "rays": last["rays"], "shadow_rays": last["shadow_rays"],
        "ray_order": {"closest_sorted": bool(last["ray_order"] & 1), "shadow_sorted": bool(last["ray_order"] & 2)},
        "grid_quarters": last["grid_quarters"],
        "mrays_per_s": round((last["rays"] + last["shadow_rays"]) / (ms * 1e3), 1),
        "pool_4m": {"value": round(W * H * spp / (ms_big * 1e3), 2), "ms_per_image": round(ms_big, 2),
                    "note": "2^22 path slots instead of the reference's 2^21: each pixel sums the same samples in another order"},
        "config": {"workload": "sphere grid 32x32 UV spheres, 2,097,218 tris, 64 emitters", "width": W,
                   "height": H, "spp": spp, "max_depth": depth, "seed": 0xC0FFEE, "dtype": "f32"},
    }
    assert (counted["rays"], counted["shadow_rays"]) == (last["rays"], last["shadow_rays"])
    rb = render_bytes(counted, W * H * spp, W * H)
    out["roofline"] = render_roofline(rb, runs)
    rays_all = counted["rays"] + counted["shadow_rays"]
    out["traversal"] = {
        "node_visits_per_ray": round((counted["closest_counts"][0] + counted["shadow_counts"][0]) / rays_all, 2),
        "tri_tests_per_ray": round((counted["closest_counts"][2] + counted["shadow_counts"][2]) / rays_all, 2),
        "closest_counts": counted["closest_counts"], "shadow_counts": counted["shadow_counts"]}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = render_cpu_baseline(scenes, W, H, spp, depth, host, None)
    return out


def render_cpu_baseline(scenes, W, H, spp, depth, host, integrator):
    """The oracle's render of the same scene, spp and depth at a quarter of the
    resolution per axis (1/16 of the samples), on host["threads"] threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    cw, ch = W // 4, H // 4
    small = scenes.sphere_grid(width=cw, height=ch, spp=spp, max_depth=depth)
    if integrator is None:
        _, _, ost = oracle.render(small, threads=host["threads"])
        what = "wavefront"
    else:
        _, _, ost = oracle.render_megakernel(small, integrator, threads=host["threads"])
        what = integrator
    return {
        "value": round(cw * ch * spp / (ost["render_ms"] * 1e3), 4),
        "unit": "Msamples/s", "cores": host["threads"], "kind": "port",
        "sample": f"oracle {what} render of the same scene, {spp} spp, depth {depth}, at {cw}x{ch} "
                  f"(1/16 of the configured {W}x{H} samples: {cw * ch * spp}), {ost['render_ms']:.0f} ms on "
                  f"{host['threads']} threads of {host['model']}; set-up and BVH build excluded",
    }


def integrators(args, ctx, pbrtm, host):
    """Msamples/s of the megakernel integrators (megakernelpath at depth 8,
    ambientocclusion) on the same sphere grid at 1920x1080, 16 spp, each next to
    the oracle's sample-by-sample render of the same scene at 128x72."""
    import numpy as np
    from pbrtm import scenes

    W, H, spp, depth = 1920, 1080, 16, 8
    sc = scenes.sphere_grid(width=W, height=H, spp=spp, max_depth=depth)
    out = {}
    scene = ctx.scene(sc)
    try:
        for name in ("megakernelpath", "ambientocclusion"):
            scene.render(spp=spp, max_depth=depth, integrator=name)  # warm-up
            runs = [scene.render(spp=spp, max_depth=depth, integrator=name)[2] for _ in range(args.pt_steps)]
            ms = float(np.mean([r["ms_total"] for r in runs]))
            last = runs[-1]
            out[name] = {
                "value": round(W * H * spp / (ms * 1e3), 2), "unit": "Msamples/s", "ms_per_image": round(ms, 2),
                "rounds": last["rounds"], "rays": last["rays"], "shadow_rays": last["shadow_rays"],
                "mrays_per_s": round((last["rays"] + last["shadow_rays"]) / (ms * 1e3), 1),
            }
    finally:
        scene.free()
    if not args.no_cpu_baseline:
        for name in out:
            out[name]["cpu_baseline"] = render_cpu_baseline(scenes, W, H, spp, depth, host, name)
    out["config"] = {"workload": "sphere grid 32x32 UV spheres, 2,097,218 tris, 64 emitters", "width": W,
                     "height": H, "spp": spp, "max_depth": depth, "dtype": "f32"}
    return out


def batch(args, ctx, pbrtm):
    """64 mixed jobs: one by one (pbrtm_render each) against one batched launch
    (pbrtm_render_batch), after checking that every batched film equals its
    one-by-one film bit for bit. The shared pool size is tuned over --batch-pools."""
    import numpy as np
    from pbrtm import scenes

    t0 = time.perf_counter()
    spec = scenes.batch_jobs(args.batch_jobs)
    handles = [ctx.scene(sc) for sc, _, _ in spec]
    setup_s = time.perf_counter() - t0
    jobs = [(h, spp, depth, 0) for h, (_, spp, depth) in zip(handles, spec)]
    paths = sum(sc["width"] * sc["height"] * spp for sc, spp, _ in spec)
    try:
        for h, spp, depth, _ in jobs:  # warm-up: per-scene pools and records
            h.render(spp=spp, max_depth=depth)
        ref, one_ms, per_job = [], 0.0, []
        t0 = time.perf_counter()
        for h, spp, depth, _ in jobs:
            rgb, _, st = h.render(spp=spp, max_depth=depth)
            ref.append((rgb, st["rounds"]))
            one_ms += st["ms_total"]
            per_job.append((st["ms_total"], st["rays"] + st["shadow_rays"]))
        one_wall = (time.perf_counter() - t0) * 1e3
        # the library's autotuner: the first calls each try one (pool size, traversal
        # grid) candidate, later calls use the fastest; every call's films are checked
        tune, same_all = [], True
        for _ in range(6):
            rgbs, _, rounds, st = ctx.render_batch(jobs, autotune=True)
            same_all &= all(np.array_equal(a.view(np.uint32), r[0].view(np.uint32)) and n == r[1]
                            for a, n, r in zip(rgbs, rounds, ref))
            tune.append(st)
            if st["tune_state"] == 2:
                break
        assert tune[-1]["tune_state"] == 2
        st = tune[-1]
        best = st["pool_size"]
        # the floor under any batching: the jobs that fill the shared pool by
        # themselves, each rendered alone with that pool (their rays must still be
        # traced; other jobs could at best cost nothing)
        big = [(h, spp, depth) for (sc, _, _), (h, spp, depth, _) in zip(spec, jobs)
               if sc["width"] * sc["height"] * spp >= best]
        floor_ms = 0.0
        for h, spp, depth in big:
            h.render(spp=spp, max_depth=depth, pool_size=best)
            floor_ms += h.render(spp=spp, max_depth=depth, pool_size=best)[2]["ms_total"]
    finally:
        for h in handles:
            h.free()
    # the same comparison on 64 small jobs (64..256 square), where one job alone
    # leaves most of the GPU idle
    small = scenes.batch_jobs(args.batch_jobs, seed=0x5A11, res=(64, 256), tris=(10_000, 100_000))
    sh = [ctx.scene(sc) for sc, _, _ in small]
    try:
        sjobs = [(h, spp, depth, 0) for h, (_, spp, depth) in zip(sh, small)]
        for h, spp, depth, _ in sjobs:
            h.render(spp=spp, max_depth=depth)
        small_one = sum(h.render(spp=spp, max_depth=depth)[2]["ms_total"] for h, spp, depth, _ in sjobs)
        ctx.render_batch(sjobs)
        small_batch = ctx.render_batch(sjobs)[3]["ms_total"]
    finally:
        for h in sh:
            h.free()
    return {
        "metric": "batched render throughput (64 mixed jobs)",
        "value": round(paths / (st["ms_total"] * 1e3), 2),
        "unit": "Msamples/s",
        "higher_is_better": True,
        "ms_batch": round(st["ms_total"], 2),
        "ms_one_by_one": round(one_ms, 2),
        "wall_ms_one_by_one": round(one_wall, 2),
        "speedup_vs_one_by_one": round(one_ms / st["ms_total"], 3),
        "speedup_bound": {"value": round(one_ms / floor_ms, 3), "floor_ms": round(floor_ms, 2), "big_jobs": len(big),
                          "note": "one-by-one time over the time of the jobs with at least pool_size paths, each "
                                  "rendered alone with that pool: even if every other job cost nothing, no "
                                  "batching of these jobs beats one-by-one by more than this"},
        "films_identical": bool(same_all),
        "pool_size": best,
        "grid_quarters": st["grid_quarters"],
        "autotune": [{"pool_size": t["pool_size"], "grid_quarters": t["grid_quarters"], "ms": round(t["ms_total"], 2),
                      "state": "tried" if t["tune_state"] == 1 else "chosen"} for t in tune],
        "stages_ms": {"rounds": round(st["ms_rounds"], 2), "schedule": round(st["ms_schedule"], 2),
                      "film": round(st["ms_film"], 2)},
        "physical_rounds": st["rounds"], "graph_launches": st["launches"], "paths": paths,
        "scene_setup_s": round(setup_s, 2),
        "solo_jobs": st["solo_jobs"],
        "small_jobs": {"ms_one_by_one": round(small_one, 2), "ms_batch": round(small_batch, 2),
                       "speedup_vs_one_by_one": round(small_one / small_batch, 3),
                       "workload": "64 jobs: width and height U[64, 256], spp in {1, 4, 16, 64}, depth 1..8, "
                                   "the same scenes with 1e4..1e5 triangles, seed 0x5A11"},
        "config": {"workload": "64 jobs: width and height each U[64, 1024], spp in {1, 4, 16, 64}, max depth "
                               "U[1, 8]; scene = HLBVH benchmark soup of 1e4..1e6 triangles (log-uniform) under "
                               "one emissive quad, all diffuse albedo 0.5", "seed": 0xBA7C, "dtype": "f32"},
    }


def gradient(args, ctx, pbrtm):
    """Loss and gradient of a 512x512, 16 spp, depth 4 render of a 1M-triangle room
    with one diffuse albedo per triangle (3M parameters) and 8 emitter scales."""
    import numpy as np
    from pbrtm import scenes

    W = H = 512
    spp, depth = 16, 4
    t0 = time.perf_counter()
    sc = scenes.grad_room(1_000_000 - 28, W, H, spp=spp, max_depth=depth)
    tsc = ctx.scene(scenes.perturbed(sc))
    target, _, _ = tsc.render(spp=spp, max_depth=depth)
    tsc.free()
    scene = ctx.scene(sc)
    setup_s = time.perf_counter() - t0
    try:
        scene.render_grad(target, spp=spp, max_depth=depth)  # warm-up: allocations
        runs = [scene.render_grad(target, spp=spp, max_depth=depth) for _ in range(args.grad_steps)]
    finally:
        scene.free()
    ms = float(np.mean([r[4]["ms_total"] for r in runs]))
    fwd = float(np.mean([r[4]["ms_forward"] for r in runs]))
    stages = {k: round(float(np.mean([r[4][k] for r in runs])), 2) for k in ("ms_forward", "ms_adjoint", "ms_replay")}
    # everything after the forward render ends is backward work: the film
    # adjoint, the path replay, the coefficient download and the host-side
    # transpose through the RGB-to-spectrum table
    stages["ms_download_and_table_jacobian"] = round(ms - sum(float(np.mean([r[4][k] for r in runs]))
                                                               for k in ("ms_forward", "ms_adjoint", "ms_replay")), 2)
    loss, d_alb, d_sc, _, st = runs[-1]
    return {
        "metric": "gradient throughput (forward + path replay)",
        "value": round(W * H * spp / (ms * 1e3), 2),
        "unit": "Msamples/s",
        "higher_is_better": True,
        "ms_per_gradient": round(ms, 2),
        "backward_over_forward": round((ms - fwd) / fwd, 3),
        "stages_ms": stages,
        "loss": loss,
        "params": int(d_alb.size + d_sc.size),
        "grad_norm": {"albedo": float(np.linalg.norm(d_alb)), "emitters": float(np.linalg.norm(d_sc))},
        "scene_setup_s": round(setup_s, 2),
        "config": {"workload": "grad_room: 1,000,000 triangles (999,972 soup + 12 walls + 16 emitter triangles), "
                               "per-triangle albedo U[0.1, 0.9]^3, 8 emitters; target from +-10 % parameters",
                   "width": W, "height": H, "spp": spp, "max_depth": depth, "dtype": "f32"},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import pbrtm

    dev = torch.device("cuda:0")
    host = host_cpu()
    tri = pbrtm.gen_soup(args.tris, seed=args.seed, threads=16)
    rays = pbrtm.gen_camera_rays(args.width, args.height)
    nr = rays.shape[0]
    ctx = pbrtm.Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
    d_tri = torch.from_numpy(tri).to(dev)
    d_rays = torch.from_numpy(rays).to(dev)
    d_prim = torch.empty(nr, dtype=torch.int32, device=dev)
    d_t = torch.empty(nr, dtype=torch.float32, device=dev)
    d_bary = torch.empty((nr, 3), dtype=torch.float32, device=dev)
    d_bounce = torch.empty((nr, 8), dtype=torch.float32, device=dev)
    d_prim2 = torch.empty(nr, dtype=torch.int32, device=dev)
    d_t2 = torch.empty(nr, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def step(counts=None):
        bvh = ctx.build(d_tri, device_ptrs=True)
        s1 = s2 = None
        if counts is not None:
            s1, s2 = np.zeros(3, np.uint64), np.zeros(3, np.uint64)
        ev[0].record(stream)
        bvh.intersect(d_rays, device_ptrs=True, prim=d_prim, t=d_t, bary=d_bary, stats=s1)
        ev[1].record(stream)
        bvh.ao_bounce(d_rays, d_prim, d_bary, device_ptrs=True, out=d_bounce)
        ev[2].record(stream)
        bvh.intersect(d_bounce, device_ptrs=True, prim=d_prim2, t=d_t2, bary=None, stats=s2)
        ev[3].record(stream)
        stats = bvh.stats.as_dict()
        bvh.free()
        if counts is not None:
            counts.append((s1, s2))
        return stats

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    builds, prim_ms, bounce_ms, sec_ms = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
        torch.cuda.synchronize()
        builds.append(st)
        prim_ms.append(ev[0].elapsed_time(ev[1]))
        bounce_ms.append(ev[1].elapsed_time(ev[2]))
        sec_ms.append(ev[2].elapsed_time(ev[3]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # Untimed counting pass: the same traversals, with per-launch totals of box
    # and triangle tests for the algorithmic byte count.
    counts = []
    step(counts)
    torch.cuda.synchronize()
    (v1, b1, t1), (v2, b2, t2) = (tuple(int(x) for x in s) for s in counts[0])
    active2 = int((d_bounce[:, 3] > 0).sum().item())
    hits1 = int((d_prim != -1).sum().item())
    hits2 = int((d_prim2 != -1).sum().item())

    build_ms = float(np.mean([b["ms_total"] for b in builds]))
    phase = {k: round(float(np.mean([b[k] for b in builds])), 3)
             for k in ("ms_sort", "ms_treelets", "ms_top", "ms_bottom", "ms_layout")}
    trace_ms = float(np.mean(prim_ms) + np.mean(sec_ms))
    rays_traced = nr + active2
    mrays = rays_traced / (trace_ms * 1e-3) / 1e6
    # primary rays write prim + t + barycentrics (20 B), bounce rays prim + t (8 B)
    bytes1 = nr * (RAY_BYTES + 20) + v1 * WNODE_BYTES + t1 * TRI_BYTES
    bytes2 = nr * (RAY_BYTES + 8) + v2 * WNODE_BYTES + t2 * TRI_BYTES
    trace_gbs = (bytes1 + bytes2) / (trace_ms * 1e-3) / 1e9
    build_gbs = BUILD_BYTES_PER_TRI * args.tris / (build_ms * 1e-3) / 1e9
    default_workload = (args.tris, args.width, args.height, args.seed) == (10_000_000, 1920, 1080, 0x5EED)

    # The dominant single kernel of the step (rocprof: trace_wide_kernel, ~2 ms per
    # launch; the build is ~40 smaller kernels). Achieved = algorithmic bytes of the
    # two launches / their HIP-event durations on the library stream. Its limit is
    # latency, not bandwidth: a ray's node fetches are dependent, and the profile's
    # counters show partial lane use and occupancy (the bound below is the roofline
    # the bytes are priced against).
    traffic, traffic_src = TRAFFIC["trace_wide_kernel"]
    roof = {"kernel": "trace_wide_kernel", "bound": "hbm", "achieved": round(trace_gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(trace_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic if default_workload else None,
            "traffic_source": traffic_src if default_workload else None,
            "limiter": "latency of dependent node fetches (VALU lane use and waves per SIMD in the profile)",
            "bytes_per_launch": int((bytes1 + bytes2) / 2),
            "algorithmic_bytes": "32 B ray + 20/8 B hit + 132 B per 8-wide node visit + 48 B per triangle test"}
    # The headline: the whole build priced at 216 algorithmic bytes per triangle
    # (vertices read twice, sort keys and values, treelet scan, two 32 B nodes).
    build_traffic, build_traffic_src = TRAFFIC["build"]
    build_roof = {"kernel": "whole HLBVH build (about 100 launches)", "bound": "hbm",
                  "achieved": round(build_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(build_gbs / HBM_PEAK_GBS, 4),
                  "traffic": build_traffic if args.tris == 10_000_000 else None,
                  "traffic_source": build_traffic_src,
                  "traffic_gbs": round(build_traffic / (build_ms * 1e6), 1) if args.tris == 10_000_000 else None,
                  "bytes_per_build": BUILD_BYTES_PER_TRI * args.tris,
                  "algorithmic_bytes": f"{BUILD_BYTES_PER_TRI} B per triangle",
                  "limiter": "the scattered triangle gather into Morton order, the wide collapse's dependent "
                             "node fetches and the subtree finishers' per-split latency"}

    line = {
        "metric": "HLBVH build ms @10M tris",
        "value": round(build_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hlbvh build + 1080p primary + 1 diffuse bounce", "tris": args.tris,
                   "width": args.width, "height": args.height, "seed": args.seed, "mu": -6.0, "sigma": 0.5},
        "build_phases_ms": phase,
        "build_gbs_algorithmic": round(build_gbs, 1),
        "build_mtris_per_s": round(args.tris / build_ms / 1e3, 1),
        "trace_ms": {"primary": round(float(np.mean(prim_ms)), 3), "bounce_spawn": round(float(np.mean(bounce_ms)), 3),
                     "secondary": round(float(np.mean(sec_ms)), 3)},
        "mrays_per_s": round(mrays, 1),
        "rays": {"primary": nr, "primary_hits": hits1, "bounce": active2, "bounce_hits": hits2,
                 "node_visits_per_ray": round((v1 + v2) / rays_traced, 2),
                 "box_tests_per_ray": round((b1 + b2) / rays_traced, 2),
                 "tri_tests_per_ray": round((t1 + t2) / rays_traced, 2)},
        "roofline": roof,
        "build_roofline": build_roof,
        "host": host,
    }
    if args.calibrate_copy:  # a plain copy kernel of known size (256 MiB in, 256 MiB out), for counter scaling
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel; copy_ would go to the DMA engine
        torch.cuda.synchronize()
        del src, dst
    note(f"build {build_ms:.3f} ms, trace {trace_ms:.3f} ms")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args, tri, host)
        note(f"cpu build baseline {line['cpu_baseline']['value']} ms")
    del tri
    if not args.no_pt:
        del d_tri, d_rays, d_prim, d_t, d_bary, d_bounce, d_prim2, d_t2
        torch.cuda.empty_cache()
        line["pathtracer"] = pathtracer(args, ctx, pbrtm, host)
        note(f"pathtracer {line['pathtracer']['value']} Msamples/s")
        line["integrators"] = integrators(args, ctx, pbrtm, host)
        note("integrators done")
    if not args.no_batch:
        line["batch"] = batch(args, ctx, pbrtm)
        note(f"batch speed-up {line['batch']['speedup_vs_one_by_one']}")
    if not args.no_grad:
        line["gradient"] = gradient(args, ctx, pbrtm)
        note(f"gradient {line['gradient']['ms_per_gradient']} ms")
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
