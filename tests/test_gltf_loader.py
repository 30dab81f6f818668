"""Scene::load, the binary glTF loader of the C++ host interface (in place of
the reference's assimp import, CRT/Scene.cpp:20-221), through
`render_demo --dump-glb` (no GPU): node transforms (TRS and matrix, nested),
inverse-transpose normals, generated smooth normals, flipped v, indexed and
unindexed primitives, interleaved buffer views, materials by name with the
reference loader's forced roughness 1 / metallic 0, and rejection of
malformed or unsupported files."""
import json
import math
import os
import struct
import subprocess

import numpy as np
import pytest

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "firefly_amd", "render_demo")


class Glb:
    """Minimal GLB writer: accessors over one binary buffer."""

    def __init__(self):
        self.bin = bytearray()
        self.views, self.accs = [], []

    def _view(self, data: bytes, stride=None):
        while len(self.bin) % 4:
            self.bin.append(0)
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.views.append(v)
        return len(self.views) - 1

    def acc(self, arr, ctype, typ, view=None, offset=0, count=None):
        arr = np.ascontiguousarray(arr)
        if view is None:
            view = self._view(arr.tobytes())
        a = {"bufferView": view, "componentType": ctype, "type": typ,
             "count": int(count if count is not None else len(arr))}
        if offset:
            a["byteOffset"] = offset
        self.accs.append(a)
        return len(self.accs) - 1

    def interleaved(self, pos, nrm):
        data = np.concatenate([pos, nrm], 1).astype(np.float32)
        v = self._view(data.tobytes(), stride=24)
        return (self.acc(None, 5126, "VEC3", view=v, count=len(pos)),
                self.acc(None, 5126, "VEC3", view=v, offset=12, count=len(pos)))

    def write(self, path, doc, patch=None):
        doc = dict(doc, asset={"version": "2.0"}, accessors=self.accs, bufferViews=self.views,
                   buffers=[{"byteLength": len(self.bin)}])
        js = json.dumps(doc).encode()
        if patch:
            js = patch(js)
        js += b" " * (-len(js) % 4)
        b = bytes(self.bin) + b"\0" * (-len(self.bin) % 4)
        total = 12 + 8 + len(js) + 8 + len(b)
        with open(path, "wb") as f:
            f.write(struct.pack("<III", 0x46546C67, 2, total))
            f.write(struct.pack("<II", len(js), 0x4E4F534A) + js)
            f.write(struct.pack("<II", len(b), 0x004E4942) + b)


def read_dump(path):
    with open(path, "rb") as f:
        data = f.read()
    off = [0]

    def take(fmt):
        r = struct.unpack_from(fmt, data, off[0])
        off[0] += struct.calcsize(fmt)
        return r

    def arr(dt, n):
        a = np.frombuffer(data, dt, n, off[0])
        off[0] += a.nbytes
        return a

    def s():
        (n,) = take("<i")
        r = data[off[0]: off[0] + n].decode()
        off[0] += n
        return r
    meshes = []
    (nm,) = take("<i")
    for _ in range(nm):
        name, mat = s(), s()
        nv, ni, has_uv, has_tan = take("<4i")
        m = {"name": name, "material": mat, "pos": arr("<f4", 3 * nv).reshape(-1, 3),
             "nrm": arr("<f4", 3 * nv).reshape(-1, 3)}
        m["uv"] = arr("<f4", 2 * nv).reshape(-1, 2) if has_uv else None
        m["tan"] = arr("<f4", 3 * nv).reshape(-1, 3) if has_tan else None
        m["bitan"] = arr("<f4", 3 * nv).reshape(-1, 3) if has_tan else None
        m["idx"] = arr("<u4", ni)
        meshes.append(m)
    mats = {}
    (nmat,) = take("<i")
    for _ in range(nmat):
        name = s()
        mats[name] = take("<5f")
        mats[name + "#tex"] = take("<4i")
    texs = []
    (ntex,) = take("<i")
    for _ in range(ntex):
        w, h, c = take("<3i")
        texs.append(arr("<f4", w * h * c).reshape(h, w, c))
    read_dump.textures = texs
    return meshes, mats


# ---- the loader's arithmetic restated (same operations, same order) ----
def mat_mul(a, b):
    r = [0.0] * 16
    for c in range(4):
        for row in range(4):
            s = 0.0
            for k in range(4):
                s += a[k * 4 + row] * b[c * 4 + k]
            r[c * 4 + row] = s
    return r


def local_of(node):
    if "matrix" in node:
        return [float(v) for v in node["matrix"]]
    t = node.get("translation", [0, 0, 0])
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    s = node.get("scale", [1, 1, 1])
    R = [1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
         2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
         2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]
    m = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]
    for c in range(3):
        for row in range(3):
            m[c * 4 + row] = R[c * 3 + row] * s[c]
    for k in range(3):
        m[12 + k] = float(t[k])
    return m


def place(m, p):
    out = np.zeros((len(p), 3), np.float32)
    for v, (x, y, z) in enumerate(p.astype(np.float64)):
        out[v] = (((m[0] * x + m[4] * y) + m[8] * z) + m[12], ((m[1] * x + m[5] * y) + m[9] * z) + m[13],
                  ((m[2] * x + m[6] * y) + m[10] * z) + m[14])
    return out


def normal_rows(m):
    a, b, c, dd, e, f, g, h, i = m[0], m[4], m[8], m[1], m[5], m[9], m[2], m[6], m[10]
    det = a * (e * i - f * h) - b * (dd * i - f * g) + c * (dd * h - e * g)
    return [(e * i - f * h) / det, -(dd * i - f * g) / det, (dd * h - e * g) / det,
            -(b * i - c * h) / det, (a * i - c * g) / det, -(a * h - b * g) / det,
            (b * f - c * e) / det, -(a * f - c * dd) / det, (a * e - b * dd) / det]


def unit(X, Y, Z):
    ln = math.sqrt((X * X + Y * Y) + Z * Z)
    return (X / ln, Y / ln, Z / ln) if ln > 0 else (0.0, 1.0, 0.0)


def transform_normals(m, n):
    N = normal_rows(m)
    out = np.zeros((len(n), 3), np.float32)
    for v, (x, y, z) in enumerate(n.astype(np.float64)):
        out[v] = unit((N[0] * x + N[1] * y) + N[2] * z, (N[3] * x + N[4] * y) + N[5] * z,
                      (N[6] * x + N[7] * y) + N[8] * z)
    return out


def smooth_normals(pos, idx):
    acc = np.zeros((len(pos), 3))
    for k in range(0, len(idx) - 2, 3):
        p0, p1, p2 = (pos[idx[k + q]].astype(np.float64) for q in range(3))
        ux, uy, uz = p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]
        vx, vy, vz = p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]
        f = (uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx)
        for q in range(3):
            for c in range(3):
                acc[idx[k + q], c] += f[c]
    return np.array([unit(*a) for a in acc], np.float32)


def _grid(n, seed):
    rng = np.random.default_rng(seed)
    xs, zs = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n))
    pos = np.stack([xs, rng.uniform(-0.1, 0.1, xs.shape), zs], -1).reshape(-1, 3).astype(np.float32)
    nrm = rng.normal(size=pos.shape).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    uv = rng.uniform(0, 1, (len(pos), 2)).astype(np.float32)
    idx = []
    for r in range(n - 1):
        for c in range(n - 1):
            a = r * n + c
            idx += [a, a + n, a + 1, a + 1, a + n, a + n + 1]
    return pos, nrm, uv, np.array(idx)


def _run(*args):
    return subprocess.run([DEMO, *args], capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    assert os.path.exists(DEMO), "render_demo not built"
    g = Glb()
    pa, na, ta, ia = _grid(5, 1)
    pos_a, nrm_a = g.interleaved(pa, na)
    uv_a = g.acc(ta, 5126, "VEC2")
    idx_a = g.acc(ia.astype(np.uint16), 5123, "SCALAR")
    pb, _, _, ib = _grid(4, 2)
    pos_b = g.acc(pb, 5126, "VEC3")
    idx_b = g.acc(ib.astype(np.uint32), 5125, "SCALAR")
    pc = np.random.default_rng(3).uniform(-1, 1, (6, 3)).astype(np.float32)
    pos_c = g.acc(pc, 5126, "VEC3")
    q = math.sin(math.pi / 4)
    mirror = [-2.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 0.5, 0, 0.25, -1.0, 3.0, 1.0]
    doc = {
        "scene": 0,
        "scenes": [{"nodes": [0, 2, 3]}],
        "nodes": [{"translation": [1, 2, 3], "scale": [2, 2, 2], "children": [1]},
                  {"rotation": [0, q, 0, q], "translation": [0.5, 0, 0], "mesh": 0},
                  {"matrix": mirror, "mesh": 1},
                  {"mesh": 2}],
        "meshes": [{"name": "grid", "primitives": [{"attributes": {"POSITION": pos_a, "NORMAL": nrm_a,
                                                                  "TEXCOORD_0": uv_a},
                                                    "indices": idx_a, "material": 0}]},
                   {"name": "bare", "primitives": [{"attributes": {"POSITION": pos_b}, "indices": idx_b}]},
                   {"primitives": [{"attributes": {"POSITION": pos_c}, "material": 1}]}],
        "materials": [{"name": "réd", "pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.1, 0.1, 1.0],
                                                                     "roughnessFactor": 0.3,
                                                                     "metallicFactor": 0.7}},
                      {"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.6, 1.0]}}],
    }
    d = tmp_path_factory.mktemp("glb")
    path = str(d / "m.glb")
    g.write(path, doc)
    out = str(d / "m.bin")
    r = _run("--dump-glb", path, out)
    assert r.returncode == 0, r.stderr
    return doc, read_dump(out), (pa, na, ta, ia, pb, ib, pc)


def test_meshes_placed_by_their_node_transforms(model):
    doc, (meshes, _), (pa, na, ta, ia, pb, ib, pc) = model
    assert [m["name"] for m in meshes] == ["grid/0", "bare/0", "mesh2/0"]
    n0, n1 = doc["nodes"][0], doc["nodes"][1]
    w_a = mat_mul(mat_mul([1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0], local_of(n0)), local_of(n1))
    a = meshes[0]
    assert np.array_equal(a["pos"], place(w_a, pa))
    assert np.array_equal(a["nrm"], transform_normals(w_a, na))
    assert np.array_equal(a["idx"], ia)
    assert np.array_equal(a["uv"], np.stack([ta[:, 0], np.float32(1) - ta[:, 1]], 1))
    # the child's rotation about y by 90 degrees, then the parent's scale 2 and translation
    assert np.allclose(a["pos"][0], 2 * (np.array([pa[0, 2], pa[0, 1], -pa[0, 0]]) + [0.5, 0, 0]) + [1, 2, 3],
                       atol=1e-5)
    w_b = local_of(doc["nodes"][2])
    b = meshes[1]
    assert np.array_equal(b["pos"], place(w_b, pb))
    assert np.array_equal(b["nrm"], smooth_normals(b["pos"], ib)) and b["uv"] is None
    c = meshes[2]
    assert np.array_equal(c["pos"], pc) and c["idx"].tolist() == list(range(6))
    assert np.array_equal(c["nrm"], smooth_normals(pc, np.arange(6)))


def test_mirrored_transform_keeps_normals_unit_and_outward(model):
    _, (meshes, _), _ = model
    b = meshes[1]
    assert np.allclose(np.linalg.norm(b["nrm"], axis=1), 1, atol=1e-6)
    p = b["pos"][b["idx"]].reshape(-1, 3, 3).astype(np.float64)
    face = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert np.all(np.sum(face * b["nrm"][b["idx"][0::3]], 1) > 0)


def test_materials_by_name_with_forced_roughness_and_metallic(model):
    _, (meshes, mats), _ = model
    assert [m["material"] for m in meshes] == ["réd", "default", "1"]
    assert mats["réd#tex"] == (-1, -1, -1, -1)
    assert mats["réd"] == pytest.approx((0.8, 0.1, 0.1, 1.0, 0.0))
    assert mats["1"] == pytest.approx((0.2, 0.4, 0.6, 1.0, 0.0))
    assert mats["default"] == pytest.approx((1.0, 1.0, 1.0, 1.0, 0.0))


def _bad(tmp_path, name, mutate, patch=None):
    g = Glb()
    p = g.acc(np.zeros((3, 3), np.float32), 5126, "VEC3")
    i = g.acc(np.array([0, 1, 2], np.uint16), 5123, "SCALAR")
    doc = {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": p}, "indices": i}]}]}
    mutate(g, doc)
    path = str(tmp_path / (name + ".glb"))
    g.write(path, doc, patch)
    return path


@pytest.mark.parametrize("name", ["sparse", "lines", "index_range", "view_range", "truncated", "magic", "deep"])
def test_malformed_files_are_rejected(tmp_path, name):
    if name == "sparse":
        path = _bad(tmp_path, name, lambda g, d: g.accs[0].update(sparse={"count": 1}))
    elif name == "lines":
        path = _bad(tmp_path, name, lambda g, d: d["meshes"][0]["primitives"][0].update(mode=1))
    elif name == "index_range":
        path = _bad(tmp_path, name, lambda g, d: g.acc(np.array([0, 1, 7], np.uint16), 5123, "SCALAR")
                    and d["meshes"][0]["primitives"][0].update(indices=2))
    elif name == "view_range":
        path = _bad(tmp_path, name, lambda g, d: g.accs[0].update(count=1000))
    elif name == "deep":  # JSON nested far deeper than any glTF file
        path = _bad(tmp_path, name, lambda g, d: d.update(extras="DEEP"),
                    lambda js: js.replace(b'"DEEP"', b"[" * 5000 + b"0" + b"]" * 5000))
    else:
        path = _bad(tmp_path, name, lambda g, d: None)
        data = open(path, "rb").read()
        data = data[:len(data) // 2] if name == "truncated" else b"glTx" + data[4:]
        open(path, "wb").write(data)
    r = _run("--dump-glb", path, str(tmp_path / "o.bin"))
    assert r.returncode == 1 and "render_demo:" in r.stderr, (r.returncode, r.stderr)


_HUGE_CASES = {
    # count * comps and stride * (count - 1) wrap a 64-bit size for these
    "count_2^62_vec4": lambda g, d: g.accs[0].update(count=2 ** 62, type="VEC4"),
    "count_2^64-1": lambda g, d: g.accs[0].update(count=2 ** 64 - 1),
    "accessor_offset_near_2^64": lambda g, d: g.accs[0].update(byteOffset=2 ** 64 - 8),
    "view_offset_near_2^64": lambda g, d: g.views[0].update(byteOffset=2 ** 64 - 16),
    "view_length_near_2^64": lambda g, d: g.views[0].update(byteLength=2 ** 64 - 4),
    "view_offset_plus_length_wraps": lambda g, d: g.views[0].update(byteOffset=2 ** 63, byteLength=2 ** 63),
    "negative_count": lambda g, d: g.accs[0].update(count=-3),
    "negative_offset": lambda g, d: g.accs[0].update(byteOffset=-36),
    "fractional_offset": lambda g, d: g.accs[0].update(byteOffset=0.5),
    "string_count": lambda g, d: g.accs[0].update(count="3"),
    "huge_stride": lambda g, d: g.views[0].update(byteStride=2 ** 40),
    "negative_accessor_index": lambda g, d: d["meshes"][0]["primitives"][0]["attributes"].update(POSITION=-1),
    "huge_node_index": lambda g, d: d["scenes"][0].update(nodes=[2 ** 63]),
    "zero_count_huge_without_view": lambda g, d: _drop_view(g.accs[0], 2 ** 40),
}


def _drop_view(acc, count):
    del acc["bufferView"]
    acc["count"] = count


@pytest.mark.parametrize("name", sorted(_HUGE_CASES))
def test_out_of_range_sizes_and_offsets_are_rejected(tmp_path, name):
    """Sizes, offsets and indices are integers in range before any cast, and
    the bounds checks are written so that no sum or product can wrap: each
    case here once wrapped (or was undefined) and passed the old checks."""
    path = _bad(tmp_path, name, _HUGE_CASES[name])
    r = _run("--dump-glb", path, str(tmp_path / "o.bin"))
    assert r.returncode == 1 and "render_demo:" in r.stderr, (r.returncode, r.stderr)


def test_image_view_outside_the_chunk_is_rejected(tmp_path):
    """An image's bufferView whose offset + length wraps past 2^64."""
    img = np.zeros((2, 2, 3), np.uint8)
    path = str(tmp_path / "img.glb")
    g = Glb()
    p = g.acc(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), 5126, "VEC3")
    uvs = g.acc(np.zeros((3, 2), np.float32), 5126, "VEC2")
    g._view(png_bytes(img, 2, [0, 0], "stored"))
    g.views[-1].update(byteOffset=2 ** 64 - 8, byteLength=64)
    doc = {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": p, "TEXCOORD_0": uvs}, "material": 0}]}],
           "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}],
           "textures": [{"source": 0}], "images": [{"bufferView": len(g.views) - 1, "mimeType": "image/png"}]}
    g.write(path, doc)
    r = _run("--dump-glb", path, str(tmp_path / "o.bin"))
    assert r.returncode == 1 and ("image outside" in r.stderr or "not an integer" in r.stderr), r.stderr


# ---- embedded PNG textures ----
def png_bytes(img, ctype, filters, zmode, depth=8, palette=None, trns=None, interlace=0, bad_adler=False):
    """PNG writer with a chosen filter per row and deflate mode ('stored',
    'fixed', 'dynamic'), so the decoder meets every block type and filter."""
    import zlib
    h, w = img.shape[:2]
    spp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = spp * depth // 8
    rows = img.reshape(h, -1)
    if depth == 16:
        rows = rows.astype(">u2").view(np.uint8).reshape(h, -1)
    raw = bytearray()
    prev = np.zeros(rows.shape[1], np.int32)
    for y in range(h):
        cur = rows[y].astype(np.int32)
        f = filters[y % len(filters)]
        left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        upleft = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if f == 0:
            out = cur
        elif f == 1:
            out = cur - left
        elif f == 2:
            out = cur - prev
        elif f == 3:
            out = cur - (left + prev) // 2
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            out = cur - pred
        raw.append(f)
        raw += (out & 0xFF).astype(np.uint8).tobytes()
        prev = cur
    if zmode == "stored":
        z = zlib.compress(bytes(raw), 0)
    else:
        co = zlib.compressobj(9, zlib.DEFLATED, 15, 9, zlib.Z_FIXED if zmode == "fixed" else zlib.Z_DEFAULT_STRATEGY)
        z = co.compress(bytes(raw)) + co.flush()
    if bad_adler:
        z = z[:-1] + bytes([z[-1] ^ 1])

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body))
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if palette is not None:
        out += chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    if trns is not None:
        out += chunk(b"tRNS", trns.astype(np.uint8).tobytes())
    return out + chunk(b"IDAT", z) + chunk(b"IEND", b"")


def _textured_glb(path, base_png, rough_png):
    g = Glb()
    p = g.acc(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), 5126, "VEC3")
    views = [g._view(base_png), g._view(rough_png)]
    g.write(path, {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
                   "meshes": [{"primitives": [{"attributes": {"POSITION": p}, "material": 0}]}],
                   "images": [{"bufferView": views[0], "mimeType": "image/png"},
                              {"bufferView": views[1], "mimeType": "image/png"}],
                   "textures": [{"source": 0}, {"source": 1}],
                   "materials": [{"name": "tex", "pbrMetallicRoughness": {
                       "baseColorTexture": {"index": 0}, "metallicRoughnessTexture": {"index": 1}}}]})


@pytest.mark.parametrize("ctype,zmode,depth", [(2, "dynamic", 8), (6, "fixed", 8), (0, "stored", 8),
                                               (3, "dynamic", 8), (4, "dynamic", 8), (2, "dynamic", 16)])
def test_png_textures_decoded_exactly(tmp_path, ctype, zmode, depth):
    """Base colour (RGB), roughness (green channel) and metalness (blue
    channel) textures from embedded PNGs of every colour type, all five row filters and all three
    deflate block types: texels = samples / 255, rows bottom first."""
    rng = np.random.default_rng(ctype * 10 + depth)
    h, w = 13, 17
    spp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    palette = trns = None
    if ctype == 3:
        palette = rng.integers(0, 256, (40, 3))
        trns = rng.integers(0, 256, 12)
        img = rng.integers(0, 40, (h, w, 1))
        rgb = palette[img[..., 0]]
    else:
        img = rng.integers(0, 256 if depth == 8 else 65536, (h, w, spp))
        top = img >> 8 if depth == 16 else img
        rgb = top[..., :3] if spp >= 3 else np.repeat(top[..., :1], 3, -1)
    dtype = np.uint16 if depth == 16 else np.uint8
    png = png_bytes(img.astype(dtype), ctype, [0, 1, 2, 3, 4], zmode, depth, palette, trns)
    path = str(tmp_path / "t.glb")
    _textured_glb(path, png, png)
    out = str(tmp_path / "t.bin")
    r = _run("--dump-glb", path, out)
    assert r.returncode == 0, r.stderr
    meshes, mats = read_dump(out)
    base, rough, metal = read_dump.textures
    assert mats["tex#tex"] == (0, 1, 2, -1)
    want = (rgb[::-1].astype(np.float32) / np.float32(255)).astype(np.float32)
    assert np.array_equal(base, want)
    assert np.array_equal(rough[..., 0], want[..., 1])
    assert np.array_equal(metal[..., 0], want[..., 2])


@pytest.mark.parametrize("flaw", ["adler", "interlaced", "truncated", "filter"])
def test_malformed_png_is_rejected(tmp_path, flaw):
    img = np.random.default_rng(1).integers(0, 256, (8, 8, 3)).astype(np.uint8)
    png = png_bytes(img, 2, [1, 2], "dynamic", bad_adler=flaw == "adler", interlace=int(flaw == "interlaced"))
    if flaw == "truncated":
        png = png[:len(png) // 2]
    if flaw == "filter":
        png = png_bytes(img, 2, [7], "stored")
    path = str(tmp_path / "t.glb")
    _textured_glb(path, png, png)
    r = _run("--dump-glb", path, str(tmp_path / "o.bin"))
    assert r.returncode == 1 and "PNG" in r.stderr, r.stderr


def _normal_mapped_glb(path, tangent=None, translation=(0.0, 0.0, 0.0)):
    """A quad in the xz plane (normal +y, texture u = x, v = z) whose
    material has a normal texture; TANGENT given or not."""
    g = Glb()
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [1, 0, 1]], np.float32)
    attrs = {"POSITION": g.acc(pos, 5126, "VEC3"),
             "NORMAL": g.acc(np.tile(np.array([0, 1, 0], np.float32), (4, 1)), 5126, "VEC3"),
             "TEXCOORD_0": g.acc(pos[:, [0, 2]].copy(), 5126, "VEC2")}
    if tangent is not None:
        attrs["TANGENT"] = g.acc(np.tile(np.array(tangent, np.float32), (4, 1)), 5126, "VEC4")
    idx = g.acc(np.array([0, 2, 1, 1, 2, 3], np.uint16), 5123, "SCALAR")
    png = png_bytes(np.full((2, 2, 3), 128, np.uint8), 2, [0], "stored")
    g.write(path, {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0, "translation": list(translation)}],
                   "meshes": [{"primitives": [{"attributes": attrs, "indices": idx, "material": 0}]}],
                   "images": [{"bufferView": g._view(png), "mimeType": "image/png"}],
                   "textures": [{"source": 0}], "materials": [{"name": "bumpy", "normalTexture": {"index": 0}}]})


def test_tangents_read_from_the_file(tmp_path):
    """TANGENT xyz and bitangent cross(NORMAL, xyz) * w, placed as the
    reference places them: normalize(M (v, 1)), the translation included."""
    path, out = str(tmp_path / "t.glb"), str(tmp_path / "t.bin")
    _normal_mapped_glb(path, tangent=(0, 0, 1, -1), translation=(2, 0, 0))
    assert _run("--dump-glb", path, out).returncode == 0
    meshes, mats = read_dump(out)
    assert mats["bumpy#tex"] == (-1, -1, -1, 0)
    (me,) = meshes
    # cross((0,1,0), (0,0,1)) * -1 = (-1, 0, 0); + (2, 0, 0) -> (1, 0, 0)
    np.testing.assert_allclose(me["tan"], np.tile([2, 0, 1] / np.sqrt(5.0), (4, 1)), rtol=0, atol=1e-7)
    np.testing.assert_array_equal(me["bitan"], np.tile([1, 0, 0], (4, 1)))


def test_tangents_computed_from_texture_coordinates(tmp_path):
    """Without TANGENT the frame follows the texture coordinates: on the
    quad u grows along +x, so the tangent is +x; the bitangent is -dp/dv of
    the file's v (z), -z, as assimp's step before the v flip gives it."""
    path, out = str(tmp_path / "t.glb"), str(tmp_path / "t.bin")
    _normal_mapped_glb(path)
    assert _run("--dump-glb", path, out).returncode == 0
    (me,), _ = read_dump(out)
    np.testing.assert_array_equal(me["tan"], np.tile([1, 0, 0], (4, 1)))
    np.testing.assert_array_equal(me["bitan"], np.tile([0, 0, -1], (4, 1)))
    assert np.all(np.abs(np.sum(me["tan"] * me["nrm"], axis=-1)) < 1e-7)


def test_no_tangents_without_a_normal_texture(tmp_path):
    g = Glb()
    p = g.acc(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), 5126, "VEC3")
    uv = g.acc(np.array([[0, 0], [1, 0], [0, 1]], np.float32), 5126, "VEC2")
    path, out = str(tmp_path / "t.glb"), str(tmp_path / "t.bin")
    g.write(path, {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
                   "meshes": [{"primitives": [{"attributes": {"POSITION": p, "TEXCOORD_0": uv}}]}]})
    assert _run("--dump-glb", path, out).returncode == 0
    (me,), _ = read_dump(out)
    assert me["tan"] is None
