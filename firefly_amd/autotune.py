"""Autotuner for the wavefront launch shapes.

Times renders of the actual request for a few paths-in-flight choices
(samples per pixel in flight, spp_batch) and for every traversal kernel
variant (workgroup size, LDS stack entries, refill threshold, loop shape, BVH
record width; see ff_traversal_variant_info), then keeps the fastest on the
context.  Results can be written as JSON and reused.
"""
from __future__ import annotations

import json
import time
from typing import Dict, Iterable, Optional

from . import abi


def _time_render(ctx, scene, env, cam, w, h, spp, depth, seed, film, spp_batch, reps):
    import torch
    best = float("inf")
    for _ in range(reps):
        film.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        abi.render(ctx, scene, env, cam, w, h, spp, depth, seed, film, spp_batch=spp_batch)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def autotune(ctx: abi.Context, scene: abi.Scene, env: abi.EnvMap, cam: abi.Camera, width: int, height: int,
             max_depth: int, seed: int = 0, probe_spp: int = 2, batches: Iterable[int] = (1, 2),
             reps: int = 2, out_json: Optional[str] = None, refine_batches: Iterable[int] = (),
             refine_spp: int = 0, tail_candidates: Iterable[int] = (), tail_from: Iterable[int] = (3, 4, 5)) -> Dict:
    """Pick the fastest traversal variant and spp_batch for this request.

    Without `refine_spp`: every variant renders `probe_spp` samples per pixel
    with each spp_batch of `batches` (timed, best of `reps`).

    With `refine_spp` (the request's spp): first the spp_batch, timed with
    variant 0 over `batches` and `refine_batches` at `refine_spp` samples --
    more paths in flight mean fewer traversal launches, each paying the
    latency of its slowest ray once; then every variant at that spp_batch
    (one batch of `min(refine_spp, spp_batch)` samples), since the ranking of
    variants depends on the launch size; then, if `tail_candidates` are given,
    a second variant for the bounce iterations from one of `tail_from` on
    (ff_set_traversal_tail), timed on the whole request.  The choice is set
    on `ctx`; returns a dict with the choice and every timing."""
    import torch
    film = torch.zeros((height, width, 4), dtype=torch.float32, device="cuda")
    results = []
    variants = abi.traversal_variants()
    # warm-up (first launches pay code-object load and allocation)
    abi.render(ctx, scene, env, cam, width, height, 1, max_depth, seed, film)
    torch.cuda.synchronize()

    def entry(var, k, t, spp):
        v, block, lds, refill, ww, width_bvh = var
        return {"variant": v, "block": block, "lds_entries": lds, "refill": refill, "while_while": ww,
                "bvh_width": width_bvh, "spp_batch": k, "seconds": t, "probe_spp": spp}

    if refine_spp > 0:
        ks = sorted(set(k for k in list(batches) + list(refine_batches) if k <= refine_spp))
        abi.set_traversal_variant(ctx, variants[0][0])
        stage1 = []
        for k in ks:
            t = _time_render(ctx, scene, env, cam, width, height, refine_spp, max_depth, seed, film, k, reps)
            stage1.append(entry(variants[0], k, t, refine_spp))
        k_best = min(stage1, key=lambda r: r["seconds"])["spp_batch"]
        spp2 = min(refine_spp, k_best)
        stage2 = []
        for var in variants:
            abi.set_traversal_variant(ctx, var[0])
            t = _time_render(ctx, scene, env, cam, width, height, spp2, max_depth, seed, film, k_best, reps)
            stage2.append(entry(var, k_best, t, spp2))
        best = min(stage2, key=lambda r: r["seconds"])
        # stage 3: a second variant for the late bounces, whose launches are
        # small (the shapes ranked on full launches are not the best there)
        abi.set_traversal_variant(ctx, best["variant"])
        stage3 = []
        if tail_candidates:
            base = _time_render(ctx, scene, env, cam, width, height, refine_spp, max_depth, seed, film, k_best, reps)
            best = dict(best, tail_variant=-1, tail_from=0, seconds=base)
            for tv in tail_candidates:
                for tf in tail_from:
                    if tv == best["variant"] or tf > max_depth + 1:
                        continue
                    abi.set_traversal_tail(ctx, tv, tf)
                    t = _time_render(ctx, scene, env, cam, width, height, refine_spp, max_depth, seed, film, k_best,
                                     reps)
                    e = dict(entry(variants[tv], k_best, t, refine_spp), tail_of=best["variant"], tail_from=tf)
                    stage3.append(e)
                    if t < best["seconds"]:
                        best = dict(best, tail_variant=tv, tail_from=tf, seconds=t)
            abi.set_traversal_tail(ctx, best["tail_variant"], max(1, best["tail_from"]))
        results = stage1 + stage2 + stage3
    else:
        for var in variants:
            abi.set_traversal_variant(ctx, var[0])
            for k in batches:
                if k > probe_spp:
                    continue
                t = _time_render(ctx, scene, env, cam, width, height, probe_spp, max_depth, seed, film, k, reps)
                results.append(entry(var, k, t, probe_spp))
        best = min(results, key=lambda r: r["seconds"])
    abi.set_traversal_variant(ctx, best["variant"])
    if best.get("tail_variant", -1) >= 0:
        abi.set_traversal_tail(ctx, best["tail_variant"], best["tail_from"])
    out = {"best": best, "candidates": results,
           "request": {"width": width, "height": height, "max_depth": max_depth, "probe_spp": probe_spp,
                       "refine_spp": refine_spp}}
    if out_json:
        with open(out_json, "w") as f:
            json.dump(out, f, indent=1)
    return out
