"""The GPU wavefront's composed estimator against the same expectations that
pin the oracle's (estimator_expectation.py, computed without sampling): the
mean of 2^20 paths per pixel (32 seeds) within 4 standard errors (plus the
quadrature bound) for every material queue, a partly blocked sky (Q5
gating), and Russian roulette through a dielectric sheet (Q2).  Reference:
wf_logic + BRDF_L, CRT/kernel.cu:297-455, CRT/dMaterial.cu:363-469."""
import sys

import numpy as np
import pytest

import estimator_expectation as X
from firefly_amd import abi, scenes

pytestmark = pytest.mark.gpu

BATCHES, SPP = 32, 32768


def gpu_films(gpu, sd, batches=BATCHES, spp=SPP):
    import torch
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    cam = abi.make_camera(sd.inv_vp)
    out = []
    for b in range(batches):
        f = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
        abi.render(gpu, sc, env, cam, sd.width, sd.height, spp, sd.max_depth, 1000 + b, f, spp_batch=spp)
        out.append(f)
    torch.cuda.synchronize()
    films = [f.cpu().numpy() for f in out]
    sc.close()
    env.close()
    return films


@pytest.mark.parametrize("case", ["mix", "diffuse", "conductor", "mix_wall", "sheet"])
def test_gpu_estimator_matches_expectation(gpu, oracle, case):
    print(f"estimator expectation: {case}", file=sys.stderr, flush=True)
    if case == "sheet":
        sd, ex = X.sheet_expectations(oracle, scenes, 4.0, 12, 0.7)
    else:
        mtype = {"mix": 0, "diffuse": 1, "conductor": 2, "mix_wall": 0}[case]
        sd, ex = X.direct_expectations(oracle, scenes, mtype, case == "mix_wall")
    films = gpu_films(gpu, sd)
    assert all(np.all(f[..., 3] == SPP) for f in films)
    worst, rows = X.compare(films, SPP, ex)
    print(f"{case}: worst |z| {worst:.2f}", file=sys.stderr, flush=True)
    assert worst < 4.0, "\n".join(f"px {px} py {py} ch {c}: mean {m:.6f} expected {e:.6f} z {z:+.2f}"
                                  for px, py, c, m, e, s, z in rows)
