"""The C++ host interface (Scene -> dScene -> PathTracer::draw) renders the
same image as the C ABI driven directly: progressive one-sample draws add
the samples in the same order as one multi-sample request."""
import os
import subprocess

import numpy as np
import pytest

from firefly_amd import abi, scenes

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "firefly_amd", "render_demo")


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.fromfile(f, "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return data[::-1]


def test_cpp_pathtracer_matches_abi_render(gpu, tmp_path):
    import torch
    assert os.path.exists(DEMO), "render_demo not built"
    sd = scenes.spheres_scene(width=80, height=48, spp=6, max_depth=3)
    path = str(tmp_path / "s.ffs")
    scenes.write_scene_file(sd, path)
    out = str(tmp_path / "o.pfm")
    r = subprocess.run([DEMO, path, out, "6", "3", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = read_pfm(out)
    # the same request through the C ABI (camera matrix from the Python helper)
    sc = abi.Scene(gpu, sd.positions, sd.normals, sd.uvs, sd.material_ids, sd.materials)
    env = abi.EnvMap(gpu, sd.env)
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    abi.render(gpu, sc, env, abi.make_camera(sd.inv_vp), sd.width, sd.height, 6, 3, sd.seed, film)
    ref = film.cpu().numpy()
    ref = ref[..., :3] / ref[..., 3:4]
    err = float(np.sqrt(np.mean((img - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    print(f"C++ host interface vs C ABI: rel RMSE {err:.2e}")
    assert np.isfinite(img).all() and err < 0.01
