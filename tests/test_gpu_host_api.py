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


@pytest.mark.parametrize("textured", [False, True])
def test_cpp_pathtracer_renders_a_glb_model(gpu, tmp_path, textured):
    """Scene::load (binary glTF) feeds the same triangles, normals, texture
    coordinates, materials and embedded PNG textures to the renderer as the
    C ABI given the loaded arrays: the image matches the C ABI's render of
    them."""
    import torch
    from test_gltf_loader import Glb, read_dump, png_bytes
    sd = scenes.spheres_scene(width=80, height=48, spp=6, max_depth=3)
    g = Glb()
    prims, mats = [], []
    doc = {}
    if textured:
        rng = np.random.default_rng(3)
        base = png_bytes(rng.integers(0, 256, (24, 32, 3)).astype(np.uint8), 2, [0, 1, 2, 3, 4], "dynamic")
        mr = png_bytes(rng.integers(0, 256, (16, 16, 4)).astype(np.uint8), 6, [4], "fixed")
        bump = np.concatenate([rng.integers(70, 186, (20, 20, 2)), rng.integers(220, 256, (20, 20, 1))], axis=-1)
        nmap = png_bytes(bump.astype(np.uint8), 2, [2, 4], "dynamic")
        doc["images"] = [{"bufferView": g._view(base), "mimeType": "image/png"},
                         {"bufferView": g._view(mr), "mimeType": "image/png"},
                         {"bufferView": g._view(nmap), "mimeType": "image/png"}]
        doc["textures"] = [{"source": 0}, {"source": 1}, {"source": 2}]
    for m in np.unique(sd.material_ids):
        sel = sd.material_ids == m
        p = g.acc(sd.positions[sel].reshape(-1, 3).astype(np.float32), 5126, "VEC3")
        n = g.acc(sd.normals[sel].reshape(-1, 3).astype(np.float32), 5126, "VEC3")
        attrs = {"POSITION": p, "NORMAL": n}
        pbr = {"baseColorFactor": [float(x) for x in sd.materials[m, 1:4]] + [1.0]}
        if textured:
            attrs["TEXCOORD_0"] = g.acc(sd.uvs[sel].reshape(-1, 2).astype(np.float32), 5126, "VEC2")
            pbr["baseColorTexture"] = {"index": 0}
            if m % 2 == 0:
                pbr["metallicRoughnessTexture"] = {"index": 1}
        prims.append({"attributes": attrs, "material": len(mats)})
        mats.append({"name": f"m{m}", "pbrMetallicRoughness": pbr})
        if textured and m % 3 == 1:
            mats[-1]["normalTexture"] = {"index": 2}
    glb = str(tmp_path / "spheres.glb")
    g.write(glb, {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}], "meshes": [{"primitives": prims}],
                  "materials": mats, **doc})
    # the frame (camera, environment, seed) from a scene file without meshes
    empty = scenes.SceneData(sd.positions[:0], sd.normals[:0], sd.uvs[:0], sd.material_ids[:0],
                             sd.materials, sd.env, sd.inv_vp, sd.width, sd.height, sd.spp, sd.max_depth,
                             sd.seed, cam_eye=sd.cam_eye, cam_target=sd.cam_target, cam_fovy=sd.cam_fovy)
    ffs = str(tmp_path / "frame.ffs")
    scenes.write_scene_file(empty, ffs)
    out = str(tmp_path / "o.pfm")
    r = subprocess.run([DEMO, ffs, out, "6", "3", "1", glb], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = read_pfm(out)
    # the loaded arrays through the C ABI
    dump = str(tmp_path / "d.bin")
    assert subprocess.run([DEMO, "--dump-glb", glb, dump], capture_output=True, timeout=60).returncode == 0
    meshes, loaded = read_dump(dump)
    texs = read_dump.textures
    assert len(texs) == (4 if textured else 0)
    pos = np.concatenate([m["pos"][m["idx"]] for m in meshes]).reshape(-1, 3, 3)
    nrm = np.concatenate([m["nrm"][m["idx"]] for m in meshes]).reshape(-1, 3, 3)
    if textured:
        uv = np.concatenate([m["uv"][m["idx"]] for m in meshes]).reshape(-1, 3, 2)
        def frame(m):  # meshes without a normal texture carry no tangents
            if m["tan"] is None:
                return np.zeros((len(m["idx"]), 6), np.float32)
            return np.concatenate([m["tan"], m["bitan"]], axis=-1)[m["idx"]]
        tan = np.concatenate([frame(m) for m in meshes]).reshape(-1, 3, 6)
        assert any(m["tan"] is not None for m in meshes)
    else:
        uv = np.zeros((len(pos), 3, 2), np.float32)
        tan = None
    names = sorted(k for k in loaded if not k.endswith("#tex"))
    rec = np.zeros((len(names), abi.FF_MAT_FLOATS), np.float32)
    for k, nm in enumerate(names):
        b = loaded[nm]
        rec[k, :10] = [abi.FF_MAT_MIX, b[0], b[1], b[2], b[3], b[4], 0.04, 0.04, 0.04, 1.5]
        rec[k, 10:14] = loaded[nm + "#tex"]
    mid = np.concatenate([np.full(len(m["idx"]) // 3, names.index(m["material"]), np.int32) for m in meshes])
    sc = abi.Scene(gpu, pos, nrm, uv, mid, rec, texs, tangents=tan)
    env = abi.EnvMap(gpu, sd.env)
    film = torch.zeros((sd.height, sd.width, 4), dtype=torch.float32, device="cuda")
    abi.render(gpu, sc, env, abi.make_camera(sd.inv_vp), sd.width, sd.height, 6, 3, sd.seed, film)
    ref = film.cpu().numpy()
    ref = ref[..., :3] / ref[..., 3:4]
    err = float(np.sqrt(np.mean((img - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    print(f"glTF model through the C++ host interface vs C ABI: rel RMSE {err:.2e}")
    assert np.isfinite(img).all() and err < 0.01
    sc.close()
    env.close()


def test_png_output_is_the_tone_mapped_film(gpu, tmp_path):
    """render_demo writes PNG files through the resolve kernel (accum/n *
    exposure, c/(c+1), (int)(255c)): the PNG parses (chunk CRCs, zlib
    Adler-32) and its pixels are the tone map of the PFM written from the
    same frame."""
    import zlib
    sd = scenes.spheres_scene(width=72, height=40, spp=4, max_depth=3)
    ffs = str(tmp_path / "s.ffs")
    scenes.write_scene_file(sd, ffs)
    png, pfm = str(tmp_path / "o.png"), str(tmp_path / "o.pfm")
    for out in (png, pfm):
        r = subprocess.run([DEMO, ffs, out, "4", "3", "1"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
    data = open(png, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    off, chunks = 8, {}
    while off < len(data):
        n = int.from_bytes(data[off:off + 4], "big")
        typ, body = data[off + 4:off + 8], data[off + 8:off + 8 + n]
        assert int.from_bytes(data[off + 8 + n:off + 12 + n], "big") == zlib.crc32(typ + body)
        chunks[typ] = chunks.get(typ, b"") + body
        off += 12 + n
    w, h = int.from_bytes(chunks[b"IHDR"][:4], "big"), int.from_bytes(chunks[b"IHDR"][4:8], "big")
    assert (w, h) == (sd.width, sd.height) and chunks[b"IHDR"][8:10] == b"\x08\x06"
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(raw[:, 0] == 0)
    rgba = raw[:, 1:].reshape(h, w, 4)
    avg = read_pfm(pfm).astype(np.float32)
    c = avg * np.float32(1.0)
    c = c / (c + np.float32(1.0))
    want = (np.float32(255) * c).astype(np.int32).astype(np.uint8)
    assert np.all(rgba[..., 3] == 255)
    assert np.abs(rgba[..., :3].astype(int) - want.astype(int)).max() <= 1
    assert np.mean(rgba[..., :3] == want) > 0.99
