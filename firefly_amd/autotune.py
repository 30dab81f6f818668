"""Autotuner for the wavefront launch shapes.

Times short renders of the actual request for every traversal kernel variant
(workgroup size x LDS stack entries x refill threshold, see
ff_traversal_variant_info) and for a
few paths-in-flight choices (samples per pixel in flight, spp_batch), then
keeps the fastest on the context.  Results can be written as JSON and reused.
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
             refine_spp: int = 0, finalists: int = 4) -> Dict:
    """Pick the fastest traversal variant and spp_batch for this request.

    Each candidate renders `probe_spp` samples per pixel (timed, best of
    `reps`).  With `refine_batches`, a second stage keeps the chosen variant
    and times those larger spp_batch values (and the first stage's choice)
    at `refine_spp` samples per pixel: more paths in flight mean fewer
    traversal launches, each paying the latency of its slowest ray once.
    A third stage then re-times the `finalists` fastest variants of the first
    stage at the chosen spp_batch and `refine_spp` (the short probes rank
    close variants unreliably for long renders).  The chosen variant is set
    on `ctx`; returns a dict with the choice and every timing."""
    import torch
    film = torch.zeros((height, width, 4), dtype=torch.float32, device="cuda")
    results = []
    # warm-up (first launches pay code-object load and allocation)
    abi.render(ctx, scene, env, cam, width, height, 1, max_depth, seed, film)
    torch.cuda.synchronize()
    for v, block, lds, refill, ww, width_bvh in abi.traversal_variants():
        abi.set_traversal_variant(ctx, v)
        for k in batches:
            if k > probe_spp:
                continue
            t = _time_render(ctx, scene, env, cam, width, height, probe_spp, max_depth, seed, film, k, reps)
            results.append({"variant": v, "block": block, "lds_entries": lds, "refill": refill,
                            "while_while": ww, "bvh_width": width_bvh, "spp_batch": k, "seconds": t})
    best = min(results, key=lambda r: r["seconds"])
    refine = sorted(set(k for k in refine_batches if k <= refine_spp) | {best["spp_batch"]})
    if refine_spp > 0 and len(refine) > 1:
        abi.set_traversal_variant(ctx, best["variant"])
        stage2 = []
        for k in refine:
            t = _time_render(ctx, scene, env, cam, width, height, refine_spp, max_depth, seed, film, k, reps)
            stage2.append(dict(best, spp_batch=k, seconds=t, probe_spp=refine_spp))
        results += stage2
        best = min(stage2, key=lambda r: r["seconds"])
        ranked, seen = [], set()
        for r in sorted(results[:len(results) - len(stage2)], key=lambda r: r["seconds"]):
            if r["variant"] not in seen:
                seen.add(r["variant"])
                ranked.append(r)
        stage3 = [best]
        for r in ranked[:finalists]:
            if r["variant"] == best["variant"]:
                continue
            abi.set_traversal_variant(ctx, r["variant"])
            t = _time_render(ctx, scene, env, cam, width, height, refine_spp, max_depth, seed, film,
                             best["spp_batch"], reps)
            stage3.append(dict(r, spp_batch=best["spp_batch"], seconds=t, probe_spp=refine_spp))
        results += stage3[1:]
        best = min(stage3, key=lambda r: r["seconds"])
    abi.set_traversal_variant(ctx, best["variant"])
    out = {"best": best, "candidates": results,
           "request": {"width": width, "height": height, "max_depth": max_depth, "probe_spp": probe_spp,
                       "refine_spp": refine_spp}}
    if out_json:
        with open(out_json, "w") as f:
            json.dump(out, f, indent=1)
    return out
