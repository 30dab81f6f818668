#!/usr/bin/env python3
"""Benchmark of the gfx950 wavefront path tracer on one MI355X.

Headline workload (the single-large-frame configuration): 1920x1080, 64 spp,
max depth 8, displaced icosphere at subdivision 8 (1,310,720 tris) + ground
quad, three material queues (diffuse / GGX conductor / dielectric), 2048x1024
HDR environment with a sigma = 0.5 deg sun lobe of peak 1e4.  Synthetic,
seeded inputs.  A step is one full frame (all 64 samples of every pixel).

Prints ONE JSON line: value = incoherent Mrays/s (bounce >= 1 extension rays
plus all any-hit rays, over the whole-frame wall time), with spp/s at 1080p,
the extend-kernel roofline and the CPU oracle baseline beside it.  The
default run (--workload all) adds, in the same line, the batched-requests
workload (64 uneven requests, batched vs sequential), the differentiable
single view (forward + adjoint time ratio, gradient error against the FP64
CPU adjoint) and the batched differentiable multi-view workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "Mrays/s incoherent + spp/s at 1080p; adjoint/fwd time at grad rel-err≤1e-3"

# algorithmic bytes per unit of work: extend ray = 24 (ray) + 16 (hit) +
# 32 per box tested + 36 per triangle tested; shadow ray = 24 + 4 + same.
EXT_RAY_BYTES, SHD_RAY_BYTES, NODE_BYTES, TRI_BYTES = 40, 28, 32, 36


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _shape_of(variant: int, spp_batch: int) -> dict:
    """Launch shape of a fixed traversal variant (no autotuning)."""
    from firefly_amd import abi
    for v, block, lds, refill, ww, width in abi.traversal_variants():
        if v == variant:
            return {"variant": v, "block": block, "lds_entries": lds, "refill": refill, "while_while": ww,
                    "bvh_width": width, "spp_batch": spp_batch}
    return {"variant": variant, "spp_batch": spp_batch}


def load_counters(kernel: str, workload: str):
    """The committed rocprofv3 PMC summary of `kernel` on `workload`
    (profiles/<kernel>_<workload>_counters.json): HBM bytes per dispatch
    (FETCH_SIZE x 1024 + 64 x TCC_EA0_RDREQ_128B + WRITE_SIZE x 1024), L2 hit
    rate, VALU issue and lane utilisation, with the commit and commands that
    produced them.  None if absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def limiter_of(ctr, dram_frac):
    """The resource the committed counters put closest to its roof: VALU
    issue (SQ_INSTS_VALU per CU-cycle, ceiling 1) or HBM bandwidth (this
    run's traffic rate over the 8 TB/s peak); latency when neither reaches
    0.6 of its roof."""
    if not ctr or dram_frac is None:
        return None
    v = ctr["valu_issue_per_cu_cycle"]
    res = "latency" if max(v, dram_frac) < 0.6 else ("VALU issue" if v >= dram_frac else "HBM bandwidth")
    return {"resource": res, "valu_issue_frac": round(v, 3), "hbm_frac": round(dram_frac, 3),
            "valu_lane_utilization": round(ctr["valu_lane_utilization"], 3),
            "l2_hit_rate": round(ctr["l2_hit_rate"], 3),
            "basis": "SQ_INSTS_VALU / (256 CUs x cycles), ceiling 1; lanes SQ_THREAD_CYCLES_VALU / "
                     "(64 x SQ_ACTIVE_INST_VALU); L2 TCC_HIT / (TCC_HIT + TCC_MISS); from " + ctr["source"]}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _band(ps, spp, spp_batch, threads, budget_s, height):
    """Centre band of rows whose oracle render at the frame's full spp costs
    about budget_s on `threads` threads (threads render one row each, so the
    calibration band has one row per thread)."""
    mid = height // 2
    cal = max(2, threads)
    t0 = time.perf_counter()
    ps.render(spp=spp, spp_batch=spp_batch, n_threads=threads, rows=(mid - cal // 2, mid - cal // 2 + cal))
    per_row = (time.perf_counter() - t0) / cal
    n = int(max(2, min(height, round(budget_s / max(per_row, 1e-4)))))
    r0 = max(0, mid - n // 2)
    return r0, min(height, r0 + n)


def cpu_baseline(sd, nodes, ordered, budget_s: float, spp: int, spp_batch: int, gpu_film=None):
    """Oracle (CPU restatement) on bounded samples of the timed frame: a
    centre band of rows rendered at the frame's full spp with the frame's
    layer grouping (samples in flight), sized to cost about budget_s on all
    host threads, and a narrower band on one thread.  Scaled to incoherent
    rays/s.  The all-thread band is also an image check of the timed frame:
    the GPU film's same rows against it (relative RMSE, bit-identical
    pixels)."""
    import numpy as np
    from oracle import oracle as O
    ps = O.PreparedScene(sd, nodes, ordered)
    threads = len(os.sched_getaffinity(0))
    out = {"unit": "Mrays/s", "kind": "port", "cores": threads, "host_nproc": os.cpu_count(),
           "host_cpu": _cpu_model()}
    for label, nt, budget in (("all", threads, budget_s), ("one", 1, max(2.0, budget_s / 3))):
        r0, r1 = _band(ps, spp, spp_batch, nt, budget, sd.height)
        t0 = time.perf_counter()
        film, st = ps.render(spp=spp, spp_batch=spp_batch, n_threads=nt, rows=(r0, r1))
        t = time.perf_counter() - t0
        inc = st.rays_extension + st.rays_shadow
        rec = {"value": round(inc / t / 1e6, 4), "threads": nt,
               "sample": f"rows {r0}-{r1} of the {sd.width}x{sd.height} frame at {spp} spp "
                         f"({spp_batch} in flight), depth {sd.max_depth} ({int(st.paths)} paths, {t:.1f} s)",
               "total_mrays_s": round((inc + st.rays_primary) / t / 1e6, 4),
               "spp_per_s_1080p": round(st.paths / t / (1920 * 1080), 5),
               "frame_s_extrapolated": round(t * sd.height / (r1 - r0), 1)}
        if label == "all":
            out.update(value=rec["value"], sample=rec["sample"], total_mrays_s=rec["total_mrays_s"],
                       spp_per_s_1080p=rec["spp_per_s_1080p"], frame_s_extrapolated=rec["frame_s_extrapolated"])
            if gpu_film is not None:
                g = gpu_film[r0:r1, :, :3].astype(np.float64)
                c = film[r0:r1, :, :3].astype(np.float64)
                out["image_check"] = {
                    "rows": [r0, r1],
                    "rel_rmse": float(np.sqrt(np.mean((g - c) ** 2)) / max(np.sqrt(np.mean(c ** 2)), 1e-30)),
                    "identical_pixels": float(np.mean(np.all(gpu_film[r0:r1, :, :3] == film[r0:r1, :, :3], -1))),
                    "counts_exact": bool(np.all(gpu_film[r0:r1, :, 3] == film[r0:r1, :, 3]))}
        else:
            out["one_thread"] = rec
    return out


def run_batch(args) -> dict:
    """Batched-requests configuration: 64 seeded requests of uneven
    resolution, spp, depth and triangle count, rendered (a) as one batched
    wavefront replayed from a hipGraph and (b) as 64 sequential plain frames,
    both with K samples of each pixel in flight for K in --batch-k (films
    bit-identical at equal K), and (c) as 64 sequential frames with one
    sample per pixel in flight, the reference's per-frame scheme.  Same
    traversal launch shape as the frame.  Returns the throughputs, the
    ratios and the batched call's stage times and traversal counts."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from firefly_amd import abi, scenes

    t0 = time.perf_counter()
    specs = scenes.batch_request_specs(args.requests, seed=0)
    envs = [scenes.sun_env(2048, 1024, k) for k in range(4)]
    ctx = abi.Context(0)
    abi.set_traversal_variant(ctx, args.variant)
    if args.tail_variant >= 0:
        abi.set_traversal_tail(ctx, args.tail_variant, args.tail_from)

    def make(spec):
        sd = scenes.batch_request_scene(spec, envs[spec["index"] % 4])
        return sd, abi.Scene(ctx, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)

    with ThreadPoolExecutor(8) as ex:
        built = list(ex.map(make, specs))
    env_objs = [abi.EnvMap(ctx, e) for e in envs]
    log(f"{len(built)} requests: {sum(sd.n_tris for sd, _ in built) / 1e6:.1f} M tris, "
        f"{sum(sd.width * sd.height for sd, _ in built) / 1e6:.1f} M pixels, "
        f"{sum(sd.width * sd.height * sd.spp for sd, _ in built) / 1e6:.1f} M paths "
        f"(setup {time.perf_counter() - t0:.1f} s)")
    cams = [abi.make_camera(sd.inv_vp) for sd, _ in built]
    f_seq = [torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda") for sd, _ in built]
    f_bat = [torch.zeros_like(f) for f in f_seq]
    items = [(sc, env_objs[k % 4], cams[k], sd.width, sd.height, sd.spp, sd.max_depth, sd.seed, f_bat[k])
             for k, (sd, sc) in enumerate(built)]
    ks = [int(x) for x in str(args.batch_k).split(",") if x]

    def sequential(k_in_flight):
        for k, (sd, sc) in enumerate(built):
            abi.render(ctx, sc, env_objs[k % 4], cams[k], sd.width, sd.height, sd.spp, sd.max_depth, sd.seed,
                       f_seq[k], spp_batch=k_in_flight)

    def batched(k_in_flight, flags=abi.FF_RENDER_USE_GRAPH):
        abi.render_batch(ctx, items, flags=flags, spp_batch=k_in_flight)

    def best_time(name, fn, films):
        best = float("inf")
        for i in range(args.warmup + args.steps):
            for f in films:
                f.zero_()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t1
            if i >= args.warmup:
                best = min(best, dt)
            log(f"{name} run {i + 1}: {dt:.3f} s")
        return best

    if args.batch_only:  # profiling: the batched call alone, timed per launch as the stage times below are
        best_time(f"batched K={ks[0]}", lambda: batched(ks[0], abi.FF_RENDER_TIMING), f_bat)
        return {"workload": f"batch{len(built)}_uneven", "spp_batch": ks[0]}
    sweep = {}
    identical = True
    for k_in in ks:
        tb = best_time(f"batched K={k_in}", lambda: batched(k_in), f_bat)
        ts = best_time(f"sequential K={k_in}", lambda: sequential(k_in), f_seq)
        same = all(torch.equal(a, b) for a, b in zip(f_seq, f_bat))
        identical = identical and same
        sweep[k_in] = (tb, ts, same)
    tp = best_time("sequential K=1 (plain)", lambda: sequential(1), f_seq)
    rel_plain = max(float((a - b).abs().max() / b.abs().max().clamp_min(1e-30)) for a, b in zip(f_bat, f_seq))
    kb = min(sweep, key=lambda k: sweep[k][0])  # the batched call's best K
    ks_best = min(sweep, key=lambda k: sweep[k][1])
    kp = 16 if 16 in sweep else kb  # the samples in flight of the committed counter summary
    for f in f_bat:
        f.zero_()
    batched(kp, abi.FF_RENDER_TIMING)  # one more call, timed per stage
    st = abi.last_render_stats(ctx)
    stage_ms = {k: round(getattr(st, k), 2) for k in ("ms_extend", "ms_shadow", "ms_logic", "ms_shade", "ms_other")}
    inc = st.rays_extension + st.rays_shadow
    ext_rays = st.rays_primary + st.rays_extension
    tb = sweep[kb][0]
    ts_best = min(min(v[1] for v in sweep.values()), tp)
    wl = f"batch{len(built)}_uneven"
    ext_bytes = EXT_RAY_BYTES * ext_rays + NODE_BYTES * st.ext_nodes + TRI_BYTES * st.ext_tris
    alg = ext_bytes / (st.ms_extend / 1e3) / 1e9
    ctr = load_counters("k_extend", wl) if (args.variant == 23 and args.tail_variant == 0 and args.tail_from == 4
                                            and kp == 16) else None
    traffic = round(ctr["dram_bytes_per_dispatch"]) if ctr else None
    dram = traffic * st.launches_extend / (st.ms_extend / 1e3) / 1e9 if traffic else None
    out = {
        "metric": METRIC,
        "value": round(inc / tb / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tb * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl, "requests": len(built),
                   "pixels": int(sum(sd.width * sd.height for sd, _ in built)),
                   "paths": int(sum(sd.width * sd.height * sd.spp for sd, _ in built)),
                   "tris": int(sum(sd.n_tris for sd, _ in built)), "spp_batch": kb},
        "batched_incoherent_mrays_s": round(inc / tb / 1e6, 3),
        "sequential_plain_mrays_s": round(inc / tp / 1e6, 3),
        "batched_over_plain": round(tp / tb, 3),
        "sweep": {str(k): {"batched_ms": round(v[0] * 1e3, 2), "sequential_ms": round(v[1] * 1e3, 2),
                           "batched_over_sequential": round(v[1] / v[0], 3), "films_identical": v[2]}
                  for k, v in sweep.items()},
        "batched_over_sequential_same_spp_batch": round(sweep[kb][1] / tb, 3),
        "batched_best_over_sequential_best": round(ts_best / tb, 3),
        "sequential_best_spp_batch": ks_best,
        "films_identical": bool(identical),
        "film_max_rel_diff_vs_plain": rel_plain,
        "batched_stage_ms": stage_ms,
        "stage_and_roofline_spp_batch": kp,
        "rays": {"primary": st.rays_primary, "extension": st.rays_extension, "shadow": st.rays_shadow},
        "traversal": {"ext_nodes_per_ray": round(st.ext_nodes / max(1.0, ext_rays), 3),
                      "ext_tris_per_ray": round(st.ext_tris / max(1.0, ext_rays), 3),
                      "shadow_nodes_per_ray": round(st.shadow_nodes / max(1.0, st.rays_shadow), 3),
                      "shadow_tris_per_ray": round(st.shadow_tris / max(1.0, st.rays_shadow), 3)},
        "roofline": {"bound": "hbm", "kernel": "k_extend", "achieved": round(dram, 2) if dram else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dram / HBM_PEAK_GBS, 5) if dram else None,
                     "traffic": traffic,
                     "traffic_source": f"{ctr['source']} at commit {ctr['commit']}" if ctr else None,
                     "avg_launch_ms": round(st.ms_extend / max(1.0, st.launches_extend), 4),
                     "launches": int(st.launches_extend),
                     "algorithmic": {"bytes_per_launch": round(ext_bytes / max(1.0, st.launches_extend)),
                                     "gbs": round(alg, 2), "over_hbm_peak": round(alg / HBM_PEAK_GBS, 5)},
                     "limiter": limiter_of(ctr, dram / HBM_PEAK_GBS if dram else None)},
    }
    for sc in built:
        sc[1].close()
    for e in env_objs:
        e.close()
    ctx.close()
    return out


def _rel(a, b) -> float:
    import numpy as np
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _timed(fn, reps: int) -> float:
    """Best wall time of fn() over reps runs (synchronised)."""
    import torch
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def envmap_update_ms(env, rgb, reps: int = 20) -> float:
    """Mean wall time of ff_envmap_update from a device buffer (texel copy +
    the FP64 CDF rebuild), the per-iteration cost of an inverse-rendering
    loop that optimises the environment."""
    import torch
    dev = torch.from_numpy(rgb).cuda()
    env.update(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        env.update(dev)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def run_adjoint(args) -> dict:
    """Differentiable single view: 512x512, 16 spp, depth 4, the 8-sphere
    scene with the spheres' albedo from a 512x512 RGB texture, 1024x512 sky.
    Loss = mean squared error against the image of parameters perturbed by
    N(0, 0.05) (seed 1).  Times the forward (with path storage) and the
    path-replay adjoint (FP32 atomics, and the deterministic two-pass
    reduction); gradient error against the FP64 CPU adjoint of the same
    paths on all host threads."""
    import numpy as np
    import torch
    from firefly_amd import abi, scenes

    sd = scenes.textured_spheres_scene(512, 512, 16, 4, seed=0, tex_size=512, env_w=1024, env_h=512)
    ctx = abi.Context(0)
    scene = abi.Scene(ctx, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials, sd.textures)
    env = abi.EnvMap(ctx, sd.env)
    cam = abi.make_camera(sd.inv_vp)
    K = args.adj_spp_batch
    W, H, S, D = sd.width, sd.height, sd.spp, sd.max_depth
    film = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    paths = torch.zeros((S, H, W, 4), dtype=torch.float32, device="cuda")
    # target image: perturbed texture and environment
    t_tex, t_env = scenes.perturbed(sd, seed=1)
    scene.set_texture(0, t_tex[0])
    env.update(t_env)
    abi.render(ctx, scene, env, cam, W, H, S, D, sd.seed, film, spp_batch=K)
    torch.cuda.synchronize()
    target = (film[..., :3] / film[..., 3:]).clone()
    scene.set_texture(0, sd.textures[0])
    env.update(sd.env)

    def fwd():
        film.zero_()
        abi.render_paths(ctx, scene, env, cam, W, H, S, D, sd.seed, film, paths, spp_batch=K)

    fwd()
    t_fwd = _timed(fwd, args.steps + 1)
    st_f = abi.last_render_stats(ctx)
    img = film[..., :3] / float(S)
    d4 = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    d4[..., :3] = 2.0 * (img - target) / img.numel()
    g_env = torch.zeros(sd.env.shape, dtype=torch.float32, device="cuda")
    g_tex = torch.zeros(sd.textures[0].shape, dtype=torch.float32, device="cuda")

    def adj(flags=0):
        g_env.zero_()
        g_tex.zero_()
        abi.adjoint(ctx, scene, env, cam, W, H, S, D, sd.seed, paths, d4, env_grad=g_env, tex_grads=[g_tex],
                    spp_batch=K, flags=flags)

    adj()
    t_adj = _timed(adj, args.steps + 1)
    st_a = abi.last_render_stats(ctx)
    ge, gt = g_env.cpu().numpy(), g_tex.cpu().numpy()
    adj(abi.FF_ADJOINT_DETERMINISTIC)
    t_det = _timed(lambda: adj(abi.FF_ADJOINT_DETERMINISTIC), 2)
    de1, dt1 = g_env.cpu().numpy(), g_tex.cpu().numpy()
    adj(abi.FF_ADJOINT_DETERMINISTIC)
    reproducible = bool(np.array_equal(de1, g_env.cpu().numpy()) and np.array_equal(dt1, g_tex.cpu().numpy()))
    out = {
        "workload": "tspheres_512x512_16spp_d4_tex512_env1024x512",
        "forward_ms": round(t_fwd * 1e3, 3),
        "adjoint_ms": round(t_adj * 1e3, 3),
        "adjoint_deterministic_ms": round(t_det * 1e3, 3),
        "fwd_plus_adj_over_fwd": round((t_fwd + t_adj) / t_fwd, 3),
        "forward_mrays_s": round((st_f.rays_primary + st_f.rays_extension + st_f.rays_shadow) / t_fwd / 1e6, 2),
        "adjoint_mrays_s": round((st_a.rays_primary + st_a.rays_extension + st_a.rays_shadow) / t_adj / 1e6, 2),
        "deterministic_reproducible": reproducible,
        "spp_batch": K,
        "envmap_update_ms": round(envmap_update_ms(env, sd.env), 4),
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))
        ps = O.PreparedScene(sd, *scene.bvh())
        t0 = time.perf_counter()
        _, r_env, r_tex = ps.adjoint(d4[..., :3].cpu().numpy(), n_threads=threads)
        t_cpu = time.perf_counter() - t0
        out["grad_rel_err_env"] = _rel(ge, r_env)
        out["grad_rel_err_albedo"] = _rel(gt, r_tex[0])
        out["grad_rel_err_deterministic_albedo"] = _rel(dt1, r_tex[0])
        out["cpu_fp64_adjoint_s"] = round(t_cpu, 2)
        out["cpu_threads"] = threads
    scene.close()
    env.close()
    ctx.close()
    return out


def run_multiview(args) -> dict:
    """Batched differentiable multi-view: 32 views (radius 4, upper-hemisphere
    Fibonacci directions) of the displaced icosphere (subdivision 8), 256x256,
    8 spp, depth 6; gradients w.r.t. a 1024^2 roughness texture and a 1024^2
    RGB albedo texture.  One batched forward + one batched adjoint (graph
    replay) against 32 separate forward + adjoint runs; gradient error of two
    views against the FP64 CPU adjoint."""
    import numpy as np
    import torch
    from firefly_amd import abi, scenes

    sd, vps = scenes.multiview_scene(args.views, 256, 8, 6, seed=0, subdiv=args.subdiv, tex_size=1024)
    ctx = abi.Context(0)
    abi.set_traversal_variant(ctx, args.variant)
    if args.tail_variant >= 0:
        abi.set_traversal_tail(ctx, args.tail_variant, args.tail_from)
    scene = abi.Scene(ctx, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials, sd.textures)
    env = abi.EnvMap(ctx, sd.env)
    cams = [abi.make_camera(v) for v in vps]
    W, H, S, D, V = sd.width, sd.height, sd.spp, sd.max_depth, len(vps)
    films = [torch.zeros((H, W, 4), dtype=torch.float32, device="cuda") for _ in range(V)]
    paths = [torch.zeros((S, H, W, 4), dtype=torch.float32, device="cuda") for _ in range(V)]
    items = [(scene, env, c, W, H, S, D, sd.seed, f) for c, f in zip(cams, films)]
    kb = args.mv_spp_batch  # samples in flight per pixel, the same for batched and separate runs
    t_tex, _ = scenes.perturbed(sd, seed=1)
    for t in range(2):
        scene.set_texture(t, t_tex[t])
    abi.render_batch(ctx, items, flags=abi.FF_RENDER_USE_GRAPH, spp_batch=kb)
    torch.cuda.synchronize()
    targets = [(f[..., :3] / f[..., 3:]).clone() for f in films]
    for t in range(2):
        scene.set_texture(t, sd.textures[t])

    def fwd_batch():
        for f in films:
            f.zero_()
        abi.render_batch(ctx, items, path_radiance=paths, flags=abi.FF_RENDER_USE_GRAPH, spp_batch=kb)

    fwd_batch()
    n_total = 3 * W * H * V
    ds = []
    for f, tg in zip(films, targets):
        d = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        d[..., :3] = 2.0 * (f[..., :3] / float(S) - tg) / n_total
        ds.append(d)
    g_alb = torch.zeros(sd.textures[0].shape, dtype=torch.float32, device="cuda")
    g_rough = torch.zeros(sd.textures[1].shape, dtype=torch.float32, device="cuda")

    def adj_batch(flags=abi.FF_RENDER_USE_GRAPH):
        g_alb.zero_()
        g_rough.zero_()
        abi.adjoint_batch(ctx, items, paths, ds, tex_grads=[[g_alb, g_rough]] * V, flags=flags, spp_batch=kb)

    def separate():
        g_alb.zero_()
        g_rough.zero_()
        for k in range(V):
            films[k].zero_()
            abi.render_paths(ctx, scene, env, cams[k], W, H, S, D, sd.seed, films[k], paths[k], spp_batch=kb)
            abi.adjoint(ctx, scene, env, cams[k], W, H, S, D, sd.seed, paths[k], ds[k], tex_grads=[g_alb, g_rough],
                        spp_batch=kb)

    if args.mv_forward_only:  # profiling: the batched forward alone, timed per launch as the stage times below are
        for f in films:
            f.zero_()
        abi.render_batch(ctx, items, path_radiance=paths, flags=abi.FF_RENDER_USE_GRAPH | abi.FF_RENDER_TIMING,
                         spp_batch=kb)
        torch.cuda.synchronize()
        return {"workload": f"multiview{V}_icosphere{args.subdiv}_256x256_8spp_d6_tex1024", "spp_batch": kb}
    t_fb = _timed(fwd_batch, args.steps + 1)
    st_f = abi.last_render_stats(ctx)
    for f in films:
        f.zero_()
    abi.render_batch(ctx, items, path_radiance=paths, flags=abi.FF_RENDER_USE_GRAPH | abi.FF_RENDER_TIMING,
                     spp_batch=kb)
    st_t = abi.last_render_stats(ctx)
    adj_batch()
    t_ab = _timed(adj_batch, args.steps + 1)
    st_a = abi.last_render_stats(ctx)
    ga, gr = g_alb.cpu().numpy(), g_rough.cpu().numpy()
    t_det = _timed(lambda: adj_batch(abi.FF_RENDER_USE_GRAPH | abi.FF_ADJOINT_DETERMINISTIC), 2)
    separate()
    t_sep = _timed(separate, 2)
    rays_f = st_f.rays_primary + st_f.rays_extension + st_f.rays_shadow
    rays_a = st_a.rays_primary + st_a.rays_extension + st_a.rays_shadow
    wl = f"multiview{V}_icosphere{args.subdiv}_256x256_8spp_d6_tex1024"
    ext_rays = st_t.rays_primary + st_t.rays_extension
    ext_bytes = EXT_RAY_BYTES * ext_rays + NODE_BYTES * st_t.ext_nodes + TRI_BYTES * st_t.ext_tris
    alg = ext_bytes / (st_t.ms_extend / 1e3) / 1e9
    ctr = load_counters("k_extend", wl) if (args.variant == 23 and args.tail_variant == 0 and
                                            args.tail_from == 4 and kb == 8) else None
    traffic = round(ctr["dram_bytes_per_dispatch"]) if ctr else None
    dram = traffic * st_t.launches_extend / (st_t.ms_extend / 1e3) / 1e9 if traffic else None
    out = {
        "workload": wl,
        "batched_forward_stage_ms": {k: round(getattr(st_t, k), 2)
                                     for k in ("ms_extend", "ms_shadow", "ms_logic", "ms_shade", "ms_other")},
        "traversal": {"ext_nodes_per_ray": round(st_t.ext_nodes / max(1.0, ext_rays), 3),
                      "shadow_nodes_per_ray": round(st_t.shadow_nodes / max(1.0, st_t.rays_shadow), 3)},
        "roofline": {"bound": "hbm", "kernel": "k_extend (batched forward)",
                     "achieved": round(dram, 2) if dram else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dram / HBM_PEAK_GBS, 5) if dram else None, "traffic": traffic,
                     "traffic_source": f"{ctr['source']} at commit {ctr['commit']}" if ctr else None,
                     "avg_launch_ms": round(st_t.ms_extend / max(1.0, st_t.launches_extend), 4),
                     "launches": int(st_t.launches_extend),
                     "algorithmic": {"bytes_per_launch": round(ext_bytes / max(1.0, st_t.launches_extend)),
                                     "gbs": round(alg, 2), "over_hbm_peak": round(alg / HBM_PEAK_GBS, 5)},
                     "limiter": limiter_of(ctr, dram / HBM_PEAK_GBS if dram else None)},
        "batched_forward_ms": round(t_fb * 1e3, 2),
        "batched_adjoint_ms": round(t_ab * 1e3, 2),
        "batched_adjoint_deterministic_ms": round(t_det * 1e3, 2),
        "separate_fwd_adj_ms": round(t_sep * 1e3, 2),
        "batched_over_separate": round(t_sep / (t_fb + t_ab), 3),
        "fwd_plus_adj_over_fwd": round((t_fb + t_ab) / t_fb, 3),
        "forward_mrays_s": round(rays_f / t_fb / 1e6, 2),
        "adjoint_mrays_s": round(rays_a / t_ab / 1e6, 2),
        "tris": sd.n_tris,
        "spp_batch": kb,
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))
        nodes, ordered = scene.bvh()
        errs_a, errs_r = [], []
        t0 = time.perf_counter()
        for k in range(2):
            sdk = scenes.SceneData(**{**sd.__dict__, "inv_vp": vps[k]})
            ps = O.PreparedScene(sdk, nodes, ordered)
            _, _, r_tex = ps.adjoint(ds[k][..., :3].cpu().numpy(), n_threads=threads, env_grad=False)
            ga_k = torch.zeros_like(g_alb)
            gr_k = torch.zeros_like(g_rough)
            abi.adjoint(ctx, scene, env, cams[k], W, H, S, D, sd.seed, paths[k], ds[k], tex_grads=[ga_k, gr_k])
            torch.cuda.synchronize()
            errs_a.append(_rel(ga_k.cpu().numpy(), r_tex[0]))
            errs_r.append(_rel(gr_k.cpu().numpy(), r_tex[1]))
        out["grad_rel_err_albedo"] = max(errs_a)
        out["grad_rel_err_roughness"] = max(errs_r)
        out["grad_check"] = f"views 0-1 against the FP64 CPU adjoint ({threads} threads, " \
                            f"{time.perf_counter() - t0:.1f} s)"
    scene.close()
    env.close()
    ctx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--subdiv", type=int, default=8)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp-batch", type=int, default=0, help="samples of each pixel in flight (0: all of them)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", type=int, default=23,
                    help="traversal kernel variant (default: the autotuner's usual pick for the frame)")
    ap.add_argument("--workload", choices=("all", "frame", "batch", "adjoint", "multiview", "bvh"), default="all")
    ap.add_argument("--adj-spp-batch", type=int, default=16)
    ap.add_argument("--views", type=int, default=32)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--mv-spp-batch", type=int, default=8,
                    help="multiview workload: samples of each pixel in flight (batched and separate runs)")
    ap.add_argument("--batch-k", default="8,16,32",
                    help="batch workload: samples of each pixel in flight to measure, batched and sequential "
                         "(films bit-identical at equal values)")
    ap.add_argument("--mv-forward-only", action="store_true",
                    help="multiview workload: only the batched forward (for profiling)")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch workload: only the batched call at the first --batch-k value (for profiling)")
    ap.add_argument("--autotune", action="store_true",
                    help="frame: let the autotuner pick the launch shape instead of --variant/--tail-variant")
    ap.add_argument("--no-autotune", action="store_true", help="(the default; kept for older command lines)")
    ap.add_argument("--tail-variant", type=int, default=0,
                    help="traversal variant of the bounce iterations from --tail-from on (-1: none)")
    ap.add_argument("--tail-from", type=int, default=4)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (PMC calibration)")
    args = ap.parse_args()
    if args.spp_batch <= 0:
        args.spp_batch = args.spp
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build runs on a single MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    if args.workload in ("batch", "adjoint", "multiview", "bvh"):
        fn = {"batch": run_batch, "adjoint": run_adjoint, "multiview": run_multiview, "bvh": run_bvh}[args.workload]
        out = fn(args)
        out.update({"metric": METRIC, "n_gpus": 1})
        print(json.dumps(out), flush=True)
        return 0
    out = run_frame(args)
    if args.workload == "all":
        import torch
        torch.cuda.empty_cache()
        for name, fn in (("adjoint", run_adjoint), ("multiview", run_multiview), ("batch", run_batch),
                         ("bvh_build", run_bvh)):
            log(f"workload {name} ...")
            t0 = time.perf_counter()
            sub = dict(args.__dict__)
            a = argparse.Namespace(**sub)
            a.steps, a.warmup = 1, 1
            r = fn(a)
            for k in ("metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline", "dtype",
                      "data", "unit", "film_mismatch"):
                r.pop(k, None)
            out[name] = r
            log(f"workload {name}: {time.perf_counter() - t0:.1f} s: {json.dumps(r)[:400]}")
            torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)
    return 0


def run_bvh(args) -> dict:
    """BVH construction for the frame's mesh: the reference's host SAH tree
    (ff_bvh_build_host, all host threads) against the GPU linear BVH
    (ff_bvh_build_gpu: device time of its kernels, and wall time with the
    transfers), then scene creation with each builder and the frame traced
    through each tree at the frame's launch shape."""
    import torch
    from firefly_amd import abi, scenes

    sd = scenes.icosphere_scene(args.width, args.height, args.spp, args.depth, seed=0, subdiv=args.subdiv)
    ctx = abi.Context(0)
    variant = getattr(args, "tuned_variant", args.variant)
    abi.set_traversal_variant(ctx, variant)
    tail = getattr(args, "tuned_tail", (args.tail_variant, args.tail_from))
    if tail[0] >= 0:
        abi.set_traversal_tail(ctx, tail[0], tail[1])
    t0 = time.perf_counter()
    _, _, sah_depth = abi.bvh_build_host(sd.positions)
    sah_s = time.perf_counter() - t0
    builds = {}
    for name, b in (("lbvh", abi.FF_BVH_LBVH), ("ploc", abi.FF_BVH_PLOC)):
        abi.bvh_build_gpu(ctx, sd.positions, builder=b)  # warm-up
        best_wall, best_ms = float("inf"), float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            nodes, _, depth, ms = abi.bvh_build_gpu(ctx, sd.positions, builder=b)
            best_wall, best_ms = min(best_wall, time.perf_counter() - t0), min(best_ms, ms)
        builds[name] = (best_ms, best_wall, depth, len(nodes))
        log(f"bvh: GPU {name} {best_ms:.2f} ms kernels, {best_wall * 1e3:.1f} ms with transfers "
            f"({len(nodes)} nodes, depth {depth})")
    log(f"bvh: host SAH {sah_s:.3f} s (depth {sah_depth})")
    best_ms, best_wall, lb_depth, _ = builds["lbvh"]
    env = abi.EnvMap(ctx, sd.env)
    cam = abi.make_camera(sd.inv_vp, sd.lens_radius, sd.focal, sd.exposure)
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    out = {"tris": sd.n_tris, "host_sah_build_s": round(sah_s, 3), "sah_depth": sah_depth,
           "gpu_lbvh_build_ms": round(best_ms, 3), "gpu_lbvh_build_wall_ms": round(best_wall * 1e3, 2),
           "lbvh_depth": lb_depth, "build_speedup": round(sah_s / (best_ms / 1e3), 1),
           "gpu_ploc_build_ms": round(builds["ploc"][0], 3), "gpu_ploc_build_wall_ms": round(builds["ploc"][1] * 1e3, 2),
           "ploc_depth": builds["ploc"][2], "ploc_build_speedup": round(sah_s / (builds["ploc"][0] / 1e3), 1),
           "variant": variant, "spp_batch": args.spp_batch}
    for name, b in (("sah", abi.FF_BVH_SAH), ("lbvh", abi.FF_BVH_LBVH), ("ploc", abi.FF_BVH_PLOC)):
        t0 = time.perf_counter()
        sc = abi.Scene(ctx, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials, sd.textures, builder=b)
        create_s = time.perf_counter() - t0
        best = float("inf")
        for _ in range(2):
            film.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            abi.render(ctx, sc, env, cam, sd.width, sd.height, args.spp, args.depth, sd.seed, film,
                       spp_batch=args.spp_batch)
            st = abi.last_render_stats(ctx)
            best = min(best, time.perf_counter() - t0)
        rays = st.rays_primary + st.rays_extension
        out[f"{name}_scene_create_s"] = round(create_s, 3)
        out[f"{name}_frame_ms"] = round(best * 1e3, 2)
        out[f"{name}_incoherent_mrays_s"] = round((st.rays_extension + st.rays_shadow) / best / 1e6, 1)
        out[f"{name}_ext_nodes_per_ray"] = round(st.ext_nodes / max(1.0, rays), 2)
        out[f"{name}_shadow_nodes_per_ray"] = round(st.shadow_nodes / max(1.0, st.rays_shadow), 2)
        log(f"bvh {name}: scene create {create_s:.2f} s, frame {best * 1e3:.1f} ms")
        sc.close()
    env.close()
    ctx.close()
    return out


def run_frame(args) -> dict:
    import torch
    from firefly_amd import abi, scenes

    t0 = time.perf_counter()
    sd = scenes.icosphere_scene(args.width, args.height, args.spp, args.depth, seed=0, subdiv=args.subdiv)
    log(f"scene: {sd.n_tris} tris, env {sd.env.shape[1]}x{sd.env.shape[0]} ({time.perf_counter() - t0:.1f} s)")
    ctx = abi.Context(0)
    t0 = time.perf_counter()
    scene = abi.Scene(ctx, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials, sd.textures)
    env = abi.EnvMap(ctx, sd.env)
    cam = abi.make_camera(sd.inv_vp, sd.lens_radius, sd.focal, sd.exposure)
    log(f"upload + host BVH ({scene.n_nodes} nodes, depth {scene.max_depth}) + env CDF: "
        f"{time.perf_counter() - t0:.1f} s")
    env_ms = envmap_update_ms(env, sd.env)
    log(f"envmap update ({sd.env.shape[1]}x{sd.env.shape[0]}): {env_ms:.3f} ms")
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    if args.copy_mb:
        src = torch.ones(args.copy_mb * 2 ** 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_mb} MiB read + {args.copy_mb} MiB written")
        del src, dst

    tuned = None
    if not args.autotune:
        abi.set_traversal_variant(ctx, args.variant)
        if args.tail_variant >= 0:
            abi.set_traversal_tail(ctx, args.tail_variant, args.tail_from)
    else:
        from firefly_amd.autotune import autotune
        t0 = time.perf_counter()
        tuned = autotune(ctx, scene, env, cam, sd.width, sd.height, args.depth, sd.seed, probe_spp=4,
                         batches=(1, 2, 4), refine_batches=(8, 16, 32, 64), refine_spp=args.spp,
                         tail_candidates=(0, 1, 22, 25, 26), tail_from=(3, 4, 5))
        args.spp_batch = tuned["best"]["spp_batch"]
        args.tuned_variant = tuned["best"]["variant"]
        args.tuned_tail = (tuned["best"].get("tail_variant", -1), tuned["best"].get("tail_from", 4))
        for c in tuned["candidates"]:
            log(f"  variant {c['variant']} (block {c['block']}, lds {c['lds_entries']}, refill {c['refill']}, "
                f"ww {c['while_while']}, width {c['bvh_width']}) "
                f"spp_batch {c['spp_batch']} at {c.get('probe_spp', 4)} spp: {c['seconds'] * 1e3:.1f} ms")
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {tuned['best']}")

    def frame(flags=0):
        film.zero_()
        abi.render(ctx, scene, env, cam, sd.width, sd.height, args.spp, args.depth, sd.seed, film,
                   spp_batch=args.spp_batch, flags=flags)

    for i in range(args.warmup):
        frame()
        torch.cuda.synchronize()
        log(f"warmup frame {i + 1}/{args.warmup} done")

    keys = ("rays_primary", "rays_extension", "rays_shadow", "paths", "ext_nodes", "ext_tris",
            "shadow_nodes", "shadow_tris", "ms_extend", "ms_shadow", "ms_logic", "ms_shade", "ms_other",
            "launches_extend", "launches_shadow", "rays_shadow_gated")
    tot = dict.fromkeys(keys, 0.0)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        frame()
        st = abi.last_render_stats(ctx)  # synchronises the frame
        for k in keys:
            tot[k] += getattr(st, k)
        log(f"timed frame {i + 1}/{args.steps}: {time.perf_counter() - t_start:.2f} s elapsed")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_step = elapsed / args.steps * 1e3
    # stage times and the roofline: one more frame with every launch timed by
    # hipEvents (FF_RENDER_TIMING runs the stages one after another; the timed
    # frames above overlap each iteration's shadow and extension traversals)
    frame(abi.FF_RENDER_TIMING)
    tim = abi.last_render_stats(ctx).as_dict()

    incoherent = tot["rays_extension"] + tot["rays_shadow"]
    total_rays = incoherent + tot["rays_primary"]
    spp_s = args.steps * args.spp / elapsed * (sd.width * sd.height) / (1920 * 1080)
    ext_bytes = (EXT_RAY_BYTES * (tim["rays_primary"] + tim["rays_extension"]) +
                 NODE_BYTES * tim["ext_nodes"] + TRI_BYTES * tim["ext_tris"])
    shd_bytes = SHD_RAY_BYTES * tim["rays_shadow"] + NODE_BYTES * tim["shadow_nodes"] + TRI_BYTES * tim["shadow_tris"]
    kern = {
        "k_extend": (ext_bytes, tim["ms_extend"], tim["launches_extend"]),
        "k_shadow": (shd_bytes, tim["ms_shadow"], tim["launches_shadow"]),
    }
    dom = max(kern, key=lambda k: kern[k][1])
    b, ms, nl = kern[dom]
    workload = f"icosphere{args.subdiv}_{sd.width}x{sd.height}_{args.spp}spp_d{args.depth}"
    alg = b / (ms / 1e3) / 1e9  # GB/s of algorithmic bytes (every box and triangle tested)
    # the committed counters describe this launch shape only
    same_shape = (not args.autotune and args.variant == 23 and args.tail_variant == 0 and args.tail_from == 4 and
                  args.spp_batch == args.spp)
    ctr = load_counters(dom, workload) if same_shape else None
    traffic = round(ctr["dram_bytes_per_dispatch"]) if ctr else None
    dram = traffic * nl / (ms / 1e3) / 1e9 if traffic else None  # GB/s of measured HBM bytes
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(dram, 2) if dram else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dram / HBM_PEAK_GBS, 5) if dram else None,
        "traffic": traffic,
        "traffic_source": (f"{ctr['source']} at commit {ctr['commit']}: HBM bytes per dispatch "
                           f"({ctr['dram_bytes_formula']}), mean over the frame's {ctr['dispatches']} dispatches"
                           if ctr else None),
        "avg_launch_ms": round(ms / nl, 4),
        "launches": int(nl),
        "algorithmic": {
            "bytes_per_launch": round(b / nl),
            "gbs": round(alg, 2),
            "over_hbm_peak": round(alg / HBM_PEAK_GBS, 5),
            "basis": "40 B per ray + 32 B per box + 36 B per triangle tested (this run's counters); "
                     "cache-inclusive, so not bounded by the HBM peak",
        },
        "limiter": limiter_of(ctr, dram / HBM_PEAK_GBS if dram else None),
    }
    stages = {k: round(tim[k], 2) for k in ("ms_extend", "ms_shadow", "ms_logic", "ms_shade", "ms_other")}
    per_ray = {
        "ext_nodes_per_ray": round(tot["ext_nodes"] / max(1.0, tot["rays_primary"] + tot["rays_extension"]), 3),
        "ext_tris_per_ray": round(tot["ext_tris"] / max(1.0, tot["rays_primary"] + tot["rays_extension"]), 3),
        "shadow_nodes_per_ray": round(tot["shadow_nodes"] / max(1.0, tot["rays_shadow"]), 3),
        "shadow_tris_per_ray": round(tot["shadow_tris"] / max(1.0, tot["rays_shadow"]), 3),
    }
    cpu = None
    if not args.no_cpu:
        nodes, ordered = scene.bvh()
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(sd, nodes, ordered, args.cpu_seconds, args.spp, args.spp_batch, film.cpu().numpy())
        log(f"cpu baseline: {cpu}")
    out = {
        "metric": METRIC,
        "value": round(incoherent / elapsed / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "width": sd.width, "height": sd.height, "spp": args.spp,
                   "max_depth": args.depth, "tris": sd.n_tris, "env": f"{sd.env.shape[1]}x{sd.env.shape[0]}",
                   "spp_batch": args.spp_batch},
        "launch_shape": (tuned["best"] if tuned else dict(_shape_of(max(args.variant, 0), args.spp_batch),
                                                           tail_variant=args.tail_variant,
                                                           tail_from=args.tail_from if args.tail_variant >= 0 else 0)),
        "spp_per_s_1080p": round(spp_s, 3),
        "total_mrays_s": round(total_rays / elapsed / 1e6, 3),
        "rays_per_step": {k: int(tot[k] / args.steps) for k in ("rays_primary", "rays_extension", "rays_shadow")},
        # BRDF-sample visibility rays left untraced because their vertex's light
        # ray was occluded (no pixel depends on them); not counted in value
        "rays_shadow_gated_per_step": int(tot["rays_shadow_gated"] / args.steps),
        "stage_ms_per_step": stages,
        "stage_ms_note": ("one more frame with every launch timed by hipEvents, which runs the stages one after "
                          "another; the timed frames overlap each iteration's shadow and extension traversals"),
        "traversal": per_ray,
        "envmap_update_ms": round(env_ms, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    scene.close()
    env.close()
    ctx.close()
    return out


if __name__ == "__main__":
    sys.exit(main())
