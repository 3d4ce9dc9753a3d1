"""Benchmark of the Nova hot path on one MI355X.

Workload: the 1,048,576-triangle scene (512x512-cell fBm terrain + 64 octahedral
spheres), 1920x1080, 64 spp, depth 8, Russian roulette from the 4th scatter, FP32 sky.
A step is the whole job: 64 frames of one path per pixel (the reference renders 1 spp
per frame and accumulates), 132,710,400 paths. The timed region renders K jobs after W
warm-up jobs; value = width*height*64*K / t / 1e6 Msamples/s, and the per-job device
times give the median. Scene upload and BVH build are outside the timed region and
reported separately.

roofline: the dominant kernel is the per-ray closest-hit walk (extend_kernel, bounces 1-8;
camera rays take the per-wave walk, extend_coherent_kernel). Its algorithmic bytes per ray
are 48 (ray in, hit out) + 64 per 4-wide node visited (the whole 64-byte node) + 48 per
triangle tested, with visits and tests counted by an instrumented render of the same job;
its duration comes from HIP events the library records around every extend launch on the
render stream during the timed region. The kernel is bound by VALU issue, not by HBM
(rocprof counters: the per-launch VALU instructions and busy fraction are reported beside
the byte rate, from the counter summary in profiles/).

cpu_baseline: the oracle (the CPU restatement of the reference renderer) on the host's
CPUs, timed on whole frames 1, 2, ... of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s (paths) at 1080p, depth 8; adjoint Msamples/s vs primal, 1 GPU"
HBM_PEAK_GBS = 8000.0
HIT_CACHE_BYTES = 2 << 30  # the library's adjoint hit-cache budget (include/nova_hip.h)
NODE_BYTES = 64.0  # bytes of a 4-wide node one visit reads (child boxes and references)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """The host CPUs this process may use: affinity set, cgroup CPU quota and model."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = os.cpu_count()
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        info["cgroup_cpus"] = None if quota == "max" else round(int(quota) / int(period), 2)
    except (OSError, ValueError):
        info["cgroup_cpus"] = None
    try:
        info["model"] = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        info["model"] = None
    return info


def host_threads(info):
    """One worker thread per CPU this process may run on: the affinity set, capped by the
    cgroup quota when there is one."""
    n = info["affinity"] or 1
    if info.get("cgroup_cpus"):
        n = min(n, max(1, int(info["cgroup_cpus"])))
    return max(1, n)


def cpu_baseline(scene, width, height, depth, rr_start, target_s):
    """The oracle's tile-parallel CPU renderer on whole frames of the same workload (frames
    1, 2, ... until about target_s seconds of rendering), one thread per usable host CPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    info = host_info()
    cores = host_threads(info)
    t0 = time.perf_counter()
    osc = pyoracle.OracleScene(scene)
    build_s = time.perf_counter() - t0
    log(f"[cpu] oracle scene + reference BVH build: {build_s:.1f} s; host {info}")
    frames, dt = 0, 0.0
    while dt < target_s and frames < 64:
        t0 = time.perf_counter()
        pyoracle.render_rows(osc, width, height, 1 + frames, 0, height, depth, rr_start, cores)
        dt += time.perf_counter() - t0
        frames += 1
        log(f"[cpu] frame {frames}: {dt:.1f} s")
    paths = frames * width * height
    return {"value": paths / dt / 1e6, "unit": "Msamples/s", "cores": cores, "kind": "port", "host": info,
            "sample": f"frames 1..{frames} in full ({paths} paths, 1 spp each, depth {depth}, RR from scatter {rr_start}) of the same "
                      f"workload, {dt:.1f} s on {cores} threads (every CPU the process may use); oracle BVH build {build_s:.1f} s excluded"}


def adjoint_leg(width=1280, height=720, frames=32, depth=4):
    """Primal and path-replay adjoint throughput on the differentiation workload: ~100k
    triangles, FP32 512x512 albedo and roughness textures, L2 loss against a render of
    perturbed textures (sigma 0.05). Adjoint time covers the whole backward call (its
    per-batch primal re-render and the replay), median of three calls. The default replay
    reuses the primal pass's hits of its leading bounces from a cache of at most 2 GiB; the
    recompute-only replay (ADJOINT_RETRACE) traces every bounce again and stores nothing."""
    import numpy as np
    import torch
    from axomae_amd import scenes
    from axomae_amd.nova_hip import ADJOINT_RETRACE, HipEngine
    sc, alb, rough = scenes.differentiable_scene(width=width, height=height)
    eng = HipEngine(0)
    eng.upload(sc)
    eng.build_bvh()
    rng = np.random.default_rng(0xA70AE004)
    ta = np.clip(sc.textures[alb].data + rng.normal(0, 0.05, sc.textures[alb].data.shape), 0, 1).astype(np.float32)
    tr = np.clip(sc.textures[rough].data + rng.normal(0, 0.05, sc.textures[rough].data.shape), 0.01, 1).astype(np.float32)
    eng.update_texture(alb, torch.from_numpy(ta))
    eng.update_texture(rough, torch.from_numpy(tr))
    target = eng.render(width, height, frames, depth)[0]
    eng.update_texture(alb, torch.from_numpy(sc.textures[alb].data))
    eng.update_texture(rough, torch.from_numpy(sc.textures[rough].data))
    eng.render(width, height, 1, depth)  # warm-up

    def median_time(fn):
        runs = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            runs.append(time.perf_counter() - t0)
        return sorted(runs)[1], out

    t_primal, img = median_time(lambda: eng.render(width, height, frames, depth)[0])
    dimg = 2 * (img - target)
    dimg[..., 3] = 0
    loss = float(((img[..., :3].double() - target[..., :3].double()) ** 2).sum())
    eng.adjoint(width, height, 1, depth, dimg, 1.0)  # warm-up
    t_adj, grad = median_time(lambda: eng.adjoint(width, height, frames, depth, dimg, 1.0 / frames))
    t_re, _ = median_time(lambda: eng.adjoint(width, height, frames, depth, dimg, 1.0 / frames, flags=ADJOINT_RETRACE))
    ga, gr = eng.texture_grad(grad, alb), eng.texture_grad(grad, rough)
    paths = width * height * frames
    primal, adjoint, recompute = paths / t_primal / 1e6, paths / t_adj / 1e6, paths / t_re / 1e6
    cache_bounces = min(depth + 1, HIT_CACHE_BYTES // (paths * 16))
    log(f"[bench] adjoint leg: primal {t_primal * 1e3:.1f} ms, adjoint {t_adj * 1e3:.1f} ms (hits of {cache_bounces} bounces cached), "
        f"recompute-only {t_re * 1e3:.1f} ms, loss {loss:.4e}")
    eng.close()
    return {"workload": "differentiable_100k_720p", "triangles": sc.triangle_count, "width": width, "height": height, "spp": frames,
            "max_depth": depth, "textures": "albedo 512x512x3 + roughness 512x512 FP32", "primal_msamples_s": round(primal, 3),
            "adjoint_msamples_s": round(adjoint, 3), "adjoint_vs_primal": round(adjoint / primal, 4),
            "hit_cache": {"bytes_max": HIT_CACHE_BYTES, "bounces_cached": cache_bounces, "bytes": cache_bounces * paths * 16},
            "recompute_only_msamples_s": round(recompute, 3), "recompute_only_vs_primal": round(recompute / primal, 4),
            "grad_finite": bool(torch.isfinite(ga).all().item() and torch.isfinite(gr).all().item()),
            "grad_nonzero_texels": int((ga.abs().sum(-1) > 0).sum().item() + (gr.abs().sum(-1) > 0).sum().item())}


def bake_leg(samples=1024, cpu=True):
    """The IBL bake batch: 32 FP32 equirect maps of uneven sizes (512x256 to 4096x2048,
    seeded sun), per map a 64x32 irradiance map (Sobol sampler) and 6 GGX-prefiltered mips
    from 256x128 at roughness i/5, `samples` samples per texel, all in one launch."""
    import numpy as np
    import torch
    from axomae_amd import nova_hip, scenes
    specs = scenes.bake_envmap_specs()
    maps = [(scenes.sky_envmap_device(w, h, d, si, ss), (w, h, 4)) for (w, h, d, si, ss) in specs]
    jobs = []
    for m in maps:
        jobs.append((m, 64, 32, nova_hip.BAKE_IRRADIANCE_SOBOL, 0.0))
        for i in range(6):
            jobs.append((m, 256 >> i, 128 >> i, nova_hip.BAKE_PREFILTER, i / 5.0))
    nova_hip.bake_batch(jobs[:2], 16, keep_on_device=True)  # warm-up (tables, code objects)
    runs = []
    for _ in range(3):  # median of three launches
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = nova_hip.bake_batch(jobs, samples, keep_on_device=True)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
    t = sorted(runs)[1]
    texels = sum(j[1] * j[2] for j in jobs)
    prefilter_r0 = sum(j[1] * j[2] for j in jobs if j[3] == nova_hip.BAKE_PREFILTER and j[4] == 0.0)
    n_samples = (texels - prefilter_r0) * samples + prefilter_r0
    finite = all(bool(torch.isfinite(o).all().item()) for o in outs)
    log(f"[bench] bake batch: {len(jobs)} jobs, {texels} texels, {t * 1e3:.1f} ms (runs {', '.join(f'{r * 1e3:.1f}' for r in runs)})")
    res = {"workload": "ibl_bake_batch_32_maps", "maps": len(maps), "jobs": len(jobs), "irradiance": "64x32 sobol",
           "prefilter": "6 mips from 256x128, roughness i/5", "samples_per_texel": samples, "launch_ms": round(t * 1e3, 2),
           "texels": texels, "gsamples_s": round(n_samples / t / 1e9, 3), "maps_per_s": round(len(maps) / t, 2), "finite": finite,
           "algorithmic_GBs": round(16.0 * n_samples / t / 1e9, 1)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        threads = host_threads(host_info())
        w, h, d, si, ss = specs[0]
        env = scenes.sky_envmap(w, h, d, si, ss)
        t0 = time.perf_counter()
        po.bake_irradiance(env, 64, 32, samples, threads=threads, sobol=True)
        tc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(64 * 32 * samples / tc / 1e9, 4), "unit": "Gsamples/s", "cores": threads, "kind": "port",
                               "sample": f"one {w}x{h} map, 64x32 irradiance, {samples} Sobol samples per texel"}
    return res


def batch_leg(depth=6):
    """64 seeded scenes (triangles log-uniform in [1e3, 5e5], resolution in {128..1024}^2,
    spp in {4, 16, 64}, depth 6) rendered by nova_hip_render_batch with the autotuned launch
    shape, shortest scene first; beside it the same scenes rendered one by one with
    nova_hip_render_frames."""
    import numpy as np
    import torch
    from axomae_amd import scenes
    from axomae_amd.nova_hip import BatchRenderer, HipEngine
    specs = scenes.batch_scene_specs()
    t0 = time.perf_counter()
    jobs = []
    for seed, tris, res, spp in specs:
        e = HipEngine(0)
        e.upload(scenes.batch_scene(seed, tris, res))
        e.build_bvh()
        jobs.append((e, res, res, spp, depth, {}))
    torch.cuda.synchronize()
    log(f"[bench] batch: {len(jobs)} scenes generated, uploaded and built in {time.perf_counter() - t0:.1f} s")
    jobs.sort(key=lambda j: j[1] * j[2] * j[3])
    paths = sum(j[1] * j[2] * j[3] for j in jobs)
    tris = sum(j[0].scene.triangle_count for j in jobs)
    br = BatchRenderer(0)
    cfg, ns = br.autotune(jobs)
    br.render(jobs[:2], cfg)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    finals, st, comp = br.render(jobs, cfg)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(f).all().item()) for f in finals)
    del finals
    br.close()
    # the same scenes one render call each (each engine's queues freed after its call)
    torch.cuda.synchronize()
    t_seq = 0.0
    for e, w, h, spp, d, _ in jobs:
        t1 = time.perf_counter()
        e.render(w, h, spp, d)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t1
        e.close()
    log(f"[bench] batch: {paths / 1e6:.0f} M paths, batched {t * 1e3:.0f} ms ({st.waves} waves), one by one {t_seq * 1e3:.0f} ms")
    return {"workload": "batch_64_scenes_depth6", "scenes": len(jobs), "triangles": tris, "paths": paths, "max_depth": depth,
            "value": round(paths / t / 1e6, 3), "unit": "Msamples/s", "device_ms": round(st.total_ms, 2), "waves": int(st.waves),
            "completion_ms_p50": round(float(np.percentile(comp, 50)), 2), "completion_ms_p90": round(float(np.percentile(comp, 90)), 2),
            "completion_ms_p99": round(float(np.percentile(comp, 99)), 2), "completion_ms_max": round(float(np.max(comp)), 2),
            "completion_note": "device time from the batch start to the end of each scene's last wave; with 64 scenes p99 lies "
                               "between the two latest completions (linear interpolation)",
            "order": "shortest first", "autotuned": cfg.as_dict(), "autotune_ns_per_path": round(ns, 3),
            "one_by_one_msamples_s": round(paths / t_seq / 1e6, 3), "finite": finite}


def frame_by_frame_leg(eng, sc, W, H, frames, depth, rr_start):
    """The drop-in's per-frame call: the reference engine draws one 1-spp frame per
    traverse call (engine.cpp:461-494), so the library is called frames times with
    frame_count = 1 through nova_hip_render_to_renderbuffer (render, accumulate, swap on the
    device). Median of three runs of the whole frame loop."""
    import numpy as np
    import torch
    from axomae_amd.nova_hip import RenderBuffer
    rb = RenderBuffer(eng, W, H)
    eng.render_to(rb, 1, depth, frame_start=0, rr_start=rr_start)  # warm-up
    runs = []
    for _ in range(3):
        rb.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):
            eng.render_to(rb, 1, depth, frame_start=f, rr_start=rr_start)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
    from axomae_amd.nova_hip import RB_ACCUMULATOR, RENDER_ASYNC
    acc_plain = rb.read(RB_ACCUMULATOR)
    t = sorted(runs)[1]
    log(f"[bench] frame by frame: {frames} calls of 1 frame, {t * 1e3:.1f} ms (runs {', '.join(f'{r * 1e3:.1f}' for r in runs)})")
    out = {"calls": frames, "frame_count_per_call": 1, "entry_point": "nova_hip_render_to_renderbuffer",
           "ms": round(t * 1e3, 2), "msamples_s": round(W * H * frames / t / 1e6, 3)}
    # the same loop with calls that return once their frame is queued (NOVA_HIP_RENDER_ASYNC):
    # the host queues the next frame while the device renders
    runs = []
    for _ in range(3):
        rb.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):
            eng.render_to(rb, 1, depth, frame_start=f, rr_start=rr_start, flags=RENDER_ASYNC)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
    same = bool(np.array_equal(rb.read(RB_ACCUMULATOR), acc_plain))
    rb.close()
    t = sorted(runs)[1]
    log(f"[bench] frame by frame, calls that do not wait: {t * 1e3:.1f} ms (runs {', '.join(f'{r * 1e3:.1f}' for r in runs)}); "
        f"accumulator equal to the waiting calls' {same}")
    out["queued"] = {"flags": "NOVA_HIP_RENDER_ASYNC", "ms": round(t * 1e3, 2), "msamples_s": round(W * H * frames / t / 1e6, 3),
                     "accumulator_equal": same}
    # the same engine loop with frames rendered ahead: a call that misses renders `ahead`
    # frames at once, the following ones copy theirs (nova_hip_render_to_renderbuffer_lookahead)
    ahead = 8
    rb = RenderBuffer(eng, W, H)
    eng.render_to_lookahead(rb, 0, depth, 1, rr_start=rr_start)  # warm-up
    runs = []
    for _ in range(3):
        rb.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):
            eng.render_to_lookahead(rb, f, depth, ahead, rr_start=rr_start)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
    same = bool(np.array_equal(rb.read(RB_ACCUMULATOR), acc_plain))
    rb.close()
    t = sorted(runs)[1]
    log(f"[bench] frame by frame, {ahead} frames rendered ahead: {t * 1e3:.1f} ms (runs {', '.join(f'{r * 1e3:.1f}' for r in runs)}); "
        f"accumulator equal to the one-frame calls' {same}")
    out["frames_ahead"] = {"lookahead": ahead, "entry_point": "nova_hip_render_to_renderbuffer_lookahead", "ms": round(t * 1e3, 2),
                           "msamples_s": round(W * H * frames / t / 1e6, 3), "accumulator_equal": same}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="whole jobs (64 frames of 1 spp each) in the timed region")
    ap.add_argument("--warmup", type=int, default=1, help="untimed whole jobs before the timed region")
    ap.add_argument("--frames", type=int, default=64, help="frames (1 spp each) of one job")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--rr-start", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-vs-primal leg (1280x720, 32 spp, depth 4, ~100k triangles)")
    ap.add_argument("--no-bake", action="store_true", help="skip the IBL bake batch leg (32 maps, irradiance + 6 prefilter mips)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched-render leg")
    ap.add_argument("--no-frame-loop", action="store_true", help="skip the frame-by-frame (frame_count = 1 per call) leg")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="before the render, copy a buffer of this size 3 times (a copy of known bytes for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): the multi-GPU split of frames and scenes "
              "is a design, not built here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from axomae_amd import scenes
    from axomae_amd.nova_hip import HipEngine, RENDER_COUNT_TRAVERSAL, RENDER_TIMING

    W, H, K, F = args.width, args.height, args.steps, args.frames
    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):  # an elementwise kernel of its own (NegFunctor): reads and writes copy_probe_mb MiB
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 x torch.neg over {args.copy_probe_mb} MiB (read + written)")
        del src, dst
    t0 = time.perf_counter()
    sc = scenes.million_triangles(width=W, height=H)
    log(f"[bench] scene {sc.triangle_count} triangles generated in {time.perf_counter() - t0:.1f} s")
    eng = HipEngine(0)
    t0 = time.perf_counter()
    eng.upload(sc)
    torch.cuda.synchronize()
    upload_s = time.perf_counter() - t0
    info = eng.build_bvh()
    log(f"[bench] upload {upload_s * 1e3:.0f} ms, BVH build {info.build_ms:.1f} ms, binary nodes {info.binary_nodes} depth {info.binary_depth}, "
        f"wide nodes {info.wide_nodes} depth {info.wide_depth}, leaves {info.leaves}")

    # traversal counts of one job (deterministic per ray: every timed job traces the same rays)
    eng.render(W, H, F, args.depth, rr_start=args.rr_start, flags=RENDER_COUNT_TRAVERSAL)
    torch.cuda.synchronize()
    cst = eng.render_stats()
    log(f"[bench] instrumented job: segments {cst.segments} (camera {cst.camera_rays}), node visits {cst.node_visits} "
        f"(camera {cst.node_visits_first}), triangle tests {cst.triangle_tests} (camera {cst.triangle_tests_first}), escaped {cst.escaped}; "
        f"per-ray walk: {cst.walk_wave_turns} wave turns, {cst.walk_lane_turns} lane steps, {cst.walk_tri_rounds} triangle rounds; "
        f"last bounces in one launch: {cst.tail_segments} segments, {cst.tail_node_visits} node visits")

    for _ in range(args.warmup):
        eng.render(W, H, F, args.depth, rr_start=args.rr_start)
    torch.cuda.synchronize()
    steps = []
    t0 = time.perf_counter()
    for _ in range(K):
        final, accum = eng.render(W, H, F, args.depth, rr_start=args.rr_start, flags=RENDER_TIMING)
        steps.append(eng.render_stats())
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    dev_ms = [s_.total_ms for s_ in steps]
    ext_ms = sum(s_.extend_ms for s_ in steps)
    ext0_ms = sum(s_.extend_first_ms for s_ in steps)
    shade_ms = sum(s_.shade_ms for s_ in steps)
    tail_ms = sum(s_.tail_ms for s_ in steps)
    launches = sum(s_.extend_launches for s_ in steps)
    cam_launches = K  # one camera-ray launch per job (the 64 frames are one batch)
    log(f"[bench] timed {K} jobs of {F} frames: {t * 1e3:.1f} ms (device per job {', '.join(f'{d:.1f}' for d in dev_ms)} ms; "
        f"extend {ext_ms / K:.1f} (camera walk {ext0_ms / K:.1f}), shade {shade_ms / K:.1f}, last bounces {tail_ms / K:.2f} ms per job)")
    finite = bool(torch.isfinite(final).all().item())

    paths = W * H * F
    value = paths * K / t / 1e6
    med_ms = float(np.median(dev_ms))
    # the per-ray walk (bounces 1..depth in extend launches): its rays, visits and tests are the
    # job's minus the camera rays' and minus the last bounces' that tail_kernel finished
    seg_r = cst.segments - cst.camera_rays - cst.tail_segments
    vis_r = cst.node_visits - cst.node_visits_first - cst.tail_node_visits
    tst_r = cst.triangle_tests - cst.triangle_tests_first - cst.tail_triangle_tests
    launches_r = (launches - cam_launches) // K
    bytes_r = (48.0 * seg_r + NODE_BYTES * vis_r + 48.0 * tst_r) / launches_r
    avg_r_ms = (ext_ms - ext0_ms) / (launches - cam_launches)
    achieved = bytes_r / (avg_r_ms / 1e3) / 1e9
    bytes_c = 48.0 * cst.camera_rays + NODE_BYTES * cst.node_visits_first + 48.0 * cst.triangle_tests_first
    avg_c_ms = ext0_ms / cam_launches
    # whole-path algorithmic bytes: 48 + 424 per segment + node/triangle bytes + 20 per escape + 16 per path
    path_bytes = 472.0 * cst.segments + NODE_BYTES * cst.node_visits + 48.0 * cst.triangle_tests + 20.0 * cst.escaped + 16.0 * paths
    counters = None
    cfile = os.path.join(ROOT, "profiles", "extend_counters.json")
    if os.path.exists(cfile):
        counters = json.load(open(cfile)).get("extend_kernel")
    traffic, valu = None, None
    if counters:
        traffic = counters.get("hbm_bytes_per_launch")
        valu = {"valu_busy": counters.get("valu_busy"), "texture_data_busy": counters.get("td_busy"),
                "valu_insts_per_launch": counters.get("sq_insts_valu_per_launch"),
                "valu_lane_insts_per_node_visit": round(counters["sq_insts_valu_per_launch"] * 64.0 / (vis_r / launches_r), 1)
                if counters.get("sq_insts_valu_per_launch") else None,
                "source": "profiles/extend_counters.json (rocprofv3 --pmc on this bench command's workload)"}

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": round(t * 1e3 / K, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "million_triangles_1080p", "triangles": sc.triangle_count, "width": W, "height": H, "spp": F,
                   "step": f"one whole job: {F} frames of 1 spp", "max_depth": args.depth, "russian_roulette_from_scatter": args.rr_start,
                   "sky": "2048x1024 RGBA32F"},
        "median_job_ms": round(med_ms, 3), "median_job_msamples_s": round(paths / med_ms / 1e3, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "extend_kernel<false> (per-ray closest-hit walk, bounces 1-8)",
                     "algorithmic_bytes_per_launch": round(bytes_r), "avg_launch_ms": round(avg_r_ms, 4),
                     "measured_limiter": "VALU issue and texture data return (not HBM): see valu", "valu": valu,
                     "camera_walk": {"kernel": "extend_coherent_kernel<false> (bounce 0, one walk per wave)",
                                     "algorithmic_bytes_per_launch": round(bytes_c), "avg_launch_ms": round(avg_c_ms, 4),
                                     "achieved_GBs": round(bytes_c / (avg_c_ms / 1e3) / 1e9, 1) if avg_c_ms > 0 else None}},
        "path_algorithmic_GBs": round(path_bytes * K / t / 1e9, 1),
        "bvh": {"build_ms": round(info.build_ms, 2), "binary_nodes": info.binary_nodes, "binary_depth": info.binary_depth,
                "wide_nodes": info.wide_nodes, "wide_depth": info.wide_depth, "leaves": info.leaves, "upload_ms": round(upload_s * 1e3, 1)},
        "per_ray_walk_simd": {"lanes_stepping": round(cst.walk_lane_turns / max(64.0 * cst.walk_wave_turns, 1.0), 4),
                              "lanes_testing_per_triangle_round": round(tst_r / max(64.0 * cst.walk_tri_rounds, 1.0), 4),
                              "triangle_rounds_per_turn": round(cst.walk_tri_rounds / max(cst.walk_wave_turns, 1), 4),
                              "note": "instrumented job: fraction of a wave's lanes taking a node step per loop turn, and of lanes "
                                      "testing a triangle per round of the triangle loop"},
        "per_segment": {"node_visits": round(cst.node_visits / max(cst.segments, 1), 2),
                        "triangle_tests": round(cst.triangle_tests / max(cst.segments, 1), 2),
                        "segments_per_path": round(cst.segments / paths, 3)},
        "image_finite": finite,
    }
    if not args.no_frame_loop:
        out["frame_by_frame"] = frame_by_frame_leg(eng, sc, W, H, F, args.depth, args.rr_start)
    eng.close()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
    if not args.no_bake:
        out["bake"] = bake_leg(cpu=not args.no_cpu)
    if not args.no_batch:
        out["batch"] = batch_leg()
    if not args.no_cpu:
        cb = cpu_baseline(sc, W, H, args.depth, args.rr_start, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(value / cb["value"], 1)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
