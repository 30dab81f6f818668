"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("bmax", "<f4", 3), ("offset", "<i4"),
                       ("n_prims", "<u2"), ("axis", "u1"), ("pad", "u1")])


class OrScene(C.Structure):
    _fields_ = [("n_tris", C.c_int32), ("pos", C.c_void_p), ("nrm", C.c_void_p), ("uv", C.c_void_p),
                ("mat", C.c_void_p), ("n_mats", C.c_int32), ("mats", C.c_void_p),
                ("n_tex", C.c_int32), ("tex", C.c_void_p), ("tex_w", C.c_void_p),
                ("tex_h", C.c_void_p), ("tex_c", C.c_void_p), ("env_w", C.c_int32),
                ("env_h", C.c_int32), ("env_rgb", C.c_void_p), ("env_color", C.c_float * 3),
                ("cam_inv_vp", C.c_float * 16), ("lens_radius", C.c_float), ("focal", C.c_float),
                ("exposure", C.c_float), ("n_nodes", C.c_int32), ("nodes", C.c_void_p),
                ("ordered", C.c_void_p), ("tan", C.c_void_p)]


class OrParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("max_depth", C.c_int32), ("seed", C.c_uint32), ("spp_batch", C.c_int32),
                ("n_threads", C.c_int32), ("row_begin", C.c_int32), ("row_end", C.c_int32)]


class OrStats(C.Structure):
    _fields_ = [("rays_primary", C.c_double), ("rays_extension", C.c_double),
                ("rays_shadow", C.c_double), ("paths", C.c_double), ("rays_shadow_gated", C.c_double)]


_L = None
_P = C.c_void_p


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB_PATH)
        protos = {
            "or_rng_base": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "or_rng": (C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "or_bvh_build": (C.c_int32, [_P, C.c_int32, C.c_int32, _P, _P]),
            "or_bvh_max_depth": (C.c_int32, [_P, C.c_int32]),
            "or_lbvh_build": (C.c_int32, [_P, C.c_int32, C.c_int32, _P, _P]),
            "or_ploc_build": (C.c_int32, [_P, C.c_int32, C.c_int32, _P, _P]),
            "or_trace_closest": (None, [_P, _P, _P, C.c_int32, C.c_int32, _P, _P, _P, _P]),
            "or_trace_closest_brute": (None, [C.c_int32, _P, _P, C.c_int32, C.c_int32, _P, _P]),
            "or_trace_any": (None, [_P, _P, _P, C.c_int32, _P]),
            "or_trace_any_brute": (None, [C.c_int32, _P, _P, C.c_int32, _P]),
            "or_math_array": (None, [C.c_int32, _P, _P, C.c_int32, _P]),
            "or_count_traversal": (None, [_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
            "or_trace_walk": (None, [_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
            "or_env_build": (None, [_P, C.c_int32, C.c_int32, _P, _P, _P, _P, _P]),
            "or_upper_bound": (C.c_int32, [_P, C.c_int32, C.c_float]),
            "or_env_sample_dir": (None, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_float, _P, _P, _P]),
            "or_bilinear": (None, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, _P]),
            "or_env_pdf": (C.c_float, [_P, C.c_int32, C.c_int32, _P]),
            "or_env_L": (None, [_P, C.c_int32, C.c_int32, _P, _P]),
            "or_concentric": (None, [_P, C.c_int32, _P]),
            "or_set_mutation": (None, [C.c_int32]),
            "or_set_env_tables": (None, [_P, _P, _P, _P, C.c_int32, C.c_int32]),
            "or_gen_ray": (None, [C.POINTER(OrScene), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_uint32, _P]),
            "or_ggx_ndf": (C.c_double, [_P, _P, C.c_float]),
            "or_geo_atten": (C.c_double, [_P, _P, _P, C.c_float]),
            "or_spec_pdf": (C.c_float, [_P, _P, _P, C.c_float]),
            "or_spec_f": (None, [_P, _P, _P, C.c_float, C.c_float, _P, _P, _P]),
            "or_diff_f": (None, [_P, _P, C.c_float, _P, _P, _P]),
            "or_diff_sample": (None, [_P, C.c_float, C.c_float, _P]),
            "or_spec_sample": (None, [_P, _P, C.c_float, C.c_float, C.c_float, _P]),
            "or_fresnel_dielectric": (C.c_float, [C.c_float, C.c_float, C.c_float]),
            "or_render": (None, [C.POINTER(OrScene), C.POINTER(OrParams), _P, C.POINTER(OrStats)]),
            "or_diff_adjoint": (None, [C.POINTER(OrScene), C.POINTER(OrParams), _P, _P, _P, _P]),
            "or_tapes_record": (_P, [C.POINTER(OrScene), C.POINTER(OrParams)]),
            "or_tapes_eval": (None, [_P, C.POINTER(OrScene), _P, _P, C.c_int32, C.c_int32, _P]),
            "or_tapes_len": (C.c_int32, [_P, C.c_int32, C.c_int32]),
            "or_tapes_free": (None, [_P]),
        }
        for name, (res, args) in protos.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _L = L
    return _L


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def bvh_build(positions: np.ndarray, max_prims: int = 8):
    pos = f32(positions).reshape(-1, 9)
    n = pos.shape[0]
    nodes = np.zeros(max(2 * n - 1, 1), NODE_DTYPE)
    ordered = np.zeros(n, np.int32)
    nn = lib().or_bvh_build(pos.ctypes.data, n, max_prims, nodes.ctypes.data, ordered.ctypes.data)
    return nodes[:nn].copy(), ordered


def lbvh_build(positions: np.ndarray, max_prims: int = 8):
    """The GPU linear-BVH builder restated (same node format as bvh_build)."""
    pos = f32(positions).reshape(-1, 9)
    n = pos.shape[0]
    nodes = np.zeros(max(2 * n - 1, 1), NODE_DTYPE)
    ordered = np.zeros(n, np.int32)
    nn = lib().or_lbvh_build(pos.ctypes.data, n, max_prims, nodes.ctypes.data, ordered.ctypes.data)
    return nodes[:nn].copy(), ordered


def ploc_build(positions, max_prims: int = 8):
    """The GPU PLOC builder restated (same node format as bvh_build)."""
    pos = f32(positions).reshape(-1, 9)
    n = pos.shape[0]
    nodes = np.zeros(max(2 * n - 1, 1), NODE_DTYPE)
    ordered = np.zeros(n, np.int32)
    nn = lib().or_ploc_build(pos.ctypes.data, n, max_prims, nodes.ctypes.data, ordered.ctypes.data)
    return nodes[:nn].copy(), ordered


def bvh_max_depth(nodes) -> int:
    return lib().or_bvh_max_depth(nodes.ctypes.data, len(nodes))


def ordered_tris(positions, ordered):
    return f32(f32(positions).reshape(-1, 9)[ordered])


def _rays6(rays):
    r = f32(rays)
    if r.shape[1] == 8:
        r = f32(np.concatenate([r[:, 0:3], r[:, 4:7]], 1))
    return r


def trace_closest(nodes, tris, rays, two_sided=False):
    r = _rays6(rays)
    n = r.shape[0]
    t, u, v = (np.zeros(n, np.float32) for _ in range(3))
    i = np.zeros(n, np.int32)
    lib().or_trace_closest(nodes.ctypes.data, tris.ctypes.data, r.ctypes.data, n, int(two_sided),
                           t.ctypes.data, u.ctypes.data, v.ctypes.data, i.ctypes.data)
    return t, u, v, i


def trace_closest_brute(tris, rays, two_sided=False):
    r = _rays6(rays)
    n = r.shape[0]
    t = np.zeros(n, np.float32)
    i = np.zeros(n, np.int32)
    lib().or_trace_closest_brute(tris.shape[0], tris.ctypes.data, r.ctypes.data, n, int(two_sided),
                                 t.ctypes.data, i.ctypes.data)
    return t, i


def trace_any(nodes, tris, rays):
    r = _rays6(rays)
    o = np.zeros(r.shape[0], np.uint8)
    lib().or_trace_any(nodes.ctypes.data, tris.ctypes.data, r.ctypes.data, r.shape[0], o.ctypes.data)
    return o


def trace_any_brute(tris, rays):
    r = _rays6(rays)
    o = np.zeros(r.shape[0], np.uint8)
    lib().or_trace_any_brute(tris.shape[0], tris.ctypes.data, r.ctypes.data, r.shape[0], o.ctypes.data)
    return o


MATH_FNS = {"sin": 0, "cos": 1, "atan": 2, "atan2": 3, "asin": 4, "acos": 5}


def math_d(fn, x, y=None):
    """The oracle's (and the GPU's) double transcendental `fn` element-wise;
    atan2 takes (y, x)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    yy = np.ascontiguousarray(x if y is None else y, dtype=np.float64)
    out = np.zeros_like(x)
    lib().or_math_array(MATH_FNS[fn], x.ctypes.data, yy.ctypes.data, x.size, out.ctypes.data)
    return out


def count_traversal(nodes, tris, rays, any_hit=False, prune=True, width=2):
    """(boxes, triangles) tested over all rays by the GPU's visit scheme for
    BVH records of `width` (2: child pairs, 4: grandchild quads)."""
    r = _rays6(rays)
    cn = np.zeros(1, np.int64)
    ct = np.zeros(1, np.int64)
    lib().or_count_traversal(nodes.ctypes.data, tris.ctypes.data, r.ctypes.data, r.shape[0],
                             int(any_hit), int(prune), int(width), cn.ctypes.data, ct.ctypes.data)
    return int(cn[0]), int(ct[0])


def trace_walk(nodes, tris, rays, any_hit=False, prune=True, width=2):
    """Closest/any hit by the visit scheme the GPU uses for BVH records of
    `width` (with its pruning); returns (t, slot)."""
    r = _rays6(rays)
    t = np.zeros(r.shape[0], np.float32)
    i = np.zeros(r.shape[0], np.int32)
    lib().or_trace_walk(nodes.ctypes.data, tris.ctypes.data, r.ctypes.data, r.shape[0], int(any_hit),
                        int(prune), int(width), t.ctypes.data, i.ctypes.data)
    return t, i


def env_build(rgb):
    a = f32(rgb)
    h, w = a.shape[0], a.shape[1]
    my, mp = np.zeros(h), np.zeros(h)
    conds = np.zeros((h, w))
    pdf = np.zeros((h, w), np.float32)
    denom = np.zeros(1)
    lib().or_env_build(a.ctypes.data, w, h, my.ctypes.data, mp.ctypes.data, conds.ctypes.data,
                       pdf.ctypes.data, denom.ctypes.data)
    return my, mp, conds, pdf, float(denom[0])


class env_tables:
    """Context manager: renders inside it use the given environment tables
    (marginal_y [h], marginal_p [h], conditional CDFs [h][w], pdf [h][w]) in
    place of the oracle's own serial FP64 build."""

    def __init__(self, marginal_y, marginal_p, conds, pdf):
        self.a = [np.ascontiguousarray(marginal_y, np.float64), np.ascontiguousarray(marginal_p, np.float64),
                  np.ascontiguousarray(conds, np.float64), np.ascontiguousarray(pdf, np.float32)]
        h = self.a[0].size
        assert self.a[1].size == h and self.a[2].size == self.a[3].size and self.a[2].size % h == 0
        self.w, self.h = self.a[2].size // h, h

    def __enter__(self):
        lib().or_set_env_tables(*[x.ctypes.data for x in self.a], self.w, self.h)
        return self

    def __exit__(self, *exc):
        lib().or_set_env_tables(None, None, None, None, 0, 0)
        return False


def upper_bound(lst, val) -> int:
    a = np.ascontiguousarray(lst, dtype=np.float64)
    return lib().or_upper_bound(a.ctypes.data, len(a), float(val))


def concentric(u):
    """ConcentricSampleDisk of the points u [n][2] as the oracle (and the
    GPU) compute it: [n][2]."""
    a = f32(u).reshape(-1, 2)
    out = np.zeros_like(a)
    lib().or_concentric(a.ctypes.data, a.shape[0], out.ctypes.data)
    return out


def bilinear(tex, u, v):
    t = f32(tex)
    h, w = t.shape[0], t.shape[1]
    c = t.shape[2] if t.ndim == 3 else 1
    out = np.zeros(c, np.float32)
    lib().or_bilinear(t.ctypes.data, w, h, c, float(u), float(v), out.ctypes.data)
    return out


class PreparedScene:
    """Oracle view of a SceneData (keeps the arrays alive)."""

    def __init__(self, sd, nodes=None, ordered=None):
        self.sd = sd
        self.pos = f32(sd.positions).reshape(-1, 9)
        self.nrm = f32(sd.normals).reshape(-1, 9)
        self.uv = f32(sd.uvs).reshape(-1, 6)
        self.mat = np.ascontiguousarray(sd.material_ids, dtype=np.int32)
        self.mats = f32(sd.materials).reshape(-1, 16)
        self.env = f32(sd.env)
        if nodes is None:
            nodes, ordered = bvh_build(self.pos)
        self.nodes, self.ordered = nodes, np.ascontiguousarray(ordered, dtype=np.int32)
        self.texs = [f32(t) for t in sd.textures]
        s = OrScene()
        s.n_tris = self.pos.shape[0]
        s.pos, s.nrm, s.uv, s.mat = (self.pos.ctypes.data, self.nrm.ctypes.data, self.uv.ctypes.data,
                                     self.mat.ctypes.data)
        s.n_mats, s.mats = self.mats.shape[0], self.mats.ctypes.data
        s.n_tex = len(self.texs)
        if self.texs:
            self._tp = (C.c_void_p * len(self.texs))(*[t.ctypes.data for t in self.texs])
            self._tw = np.array([t.shape[1] for t in self.texs], np.int32)
            self._th = np.array([t.shape[0] for t in self.texs], np.int32)
            self._tc = np.array([t.shape[2] if t.ndim == 3 else 1 for t in self.texs], np.int32)
            s.tex = C.cast(self._tp, C.c_void_p)
            s.tex_w, s.tex_h, s.tex_c = self._tw.ctypes.data, self._th.ctypes.data, self._tc.ctypes.data
        s.env_h, s.env_w = self.env.shape[0], self.env.shape[1]
        s.env_rgb = self.env.ctypes.data
        for i, x in enumerate(f32(sd.inv_vp).reshape(16)):
            s.cam_inv_vp[i] = float(x)
        s.lens_radius, s.focal, s.exposure = sd.lens_radius, sd.focal, sd.exposure
        s.n_nodes, s.nodes, s.ordered = len(self.nodes), self.nodes.ctypes.data, self.ordered.ctypes.data
        self.tan = None if sd.tangents is None else f32(sd.tangents).reshape(-1, 18)
        s.tan = None if self.tan is None else self.tan.ctypes.data
        self.s = s

    def tris_ordered(self):
        return ordered_tris(self.pos, self.ordered)

    def gen_ray(self, px, py, sample, seed=None, width=None, height=None):
        w = width or self.sd.width
        h = height or self.sd.height
        sd = self.sd.seed if seed is None else seed
        base = lib().or_rng_base(sd, py * w + px, sample)
        out = np.zeros(6, np.float32)
        lib().or_gen_ray(C.byref(self.s), w, h, px, py, base, out.ctypes.data)
        return out

    def _params(self, spp, max_depth, seed, n_threads, width=None, height=None):
        sd = self.sd
        p = OrParams()
        p.width, p.height = width or sd.width, height or sd.height
        p.spp = spp or sd.spp
        p.max_depth = max_depth or sd.max_depth
        p.seed = sd.seed if seed is None else seed
        p.spp_batch = 1
        p.n_threads = n_threads or os.cpu_count() or 1
        p.row_begin, p.row_end = 0, p.height
        return p

    def adjoint(self, d_image, spp=None, max_depth=None, seed=None, n_threads=None, env_grad=True,
                tex_grads=None):
        """FP64 adjoint of the detached-sampling estimator.  d_image [H][W][3] is
        d loss / d (mean image).  Returns (image [H][W][3] float64 mean of the
        FP64 estimator, g_env [eh][ew][3] or None, {texture: gradient})."""
        p = self._params(spp, max_depth, seed, n_threads)
        d = f32(d_image).reshape(p.height, p.width, 3)
        img = np.zeros((p.height, p.width, 3), np.float64)
        g_env = np.zeros(self.env.shape[:2] + (3,), np.float64) if env_grad else None
        tex_grads = list(range(len(self.texs))) if tex_grads is None else list(tex_grads)
        gt = {t: np.zeros(self.texs[t].shape, np.float64) for t in tex_grads}
        ptrs = None
        if self.texs:
            ptrs = (C.c_void_p * len(self.texs))(*[gt[t].ctypes.data if t in gt else None
                                                    for t in range(len(self.texs))])
        lib().or_diff_adjoint(C.byref(self.s), C.byref(p), d.ctypes.data, img.ctypes.data,
                              g_env.ctypes.data if g_env is not None else None,
                              C.cast(ptrs, C.c_void_p) if ptrs is not None else None)
        return img, g_env, gt

    def record_tapes(self, spp=None, max_depth=None, seed=None, n_threads=None):
        """Records every path of the float forward; .eval(textures, env)
        evaluates the FP64 detached-sampling estimator for other parameters."""
        p = self._params(spp, max_depth, seed, n_threads)
        return Tapes(self, lib().or_tapes_record(C.byref(self.s), C.byref(p)), p.width, p.height, p.spp)

    def render(self, spp=None, max_depth=None, seed=None, spp_batch=1, n_threads=None,
               rows=None, width=None, height=None):
        sd = self.sd
        p = OrParams()
        p.width, p.height = width or sd.width, height or sd.height
        p.spp = spp or sd.spp
        p.max_depth = max_depth or sd.max_depth
        p.seed = sd.seed if seed is None else seed
        p.spp_batch = spp_batch
        p.n_threads = n_threads or os.cpu_count() or 1
        p.row_begin, p.row_end = rows if rows else (0, p.height)
        film = np.zeros((p.height, p.width, 4), np.float32)
        st = OrStats()
        lib().or_render(C.byref(self.s), C.byref(p), film.ctypes.data, C.byref(st))
        return film, st


class Tapes:
    """Recorded paths of the float forward (oracle path tapes)."""

    def __init__(self, ps, handle, width, height, spp):
        self.ps, self.h, self.width, self.height, self.spp = ps, handle, width, height, spp

    def eval(self, textures=None, env=None, rows=None):
        """FP64 mean image [H][W][3] of the recorded paths with the given
        textures (list, shapes as the scene's) and environment."""
        ps = self.ps
        texs = [f32(t) for t in (textures if textures is not None else ps.texs)]
        e = f32(env if env is not None else ps.env)
        ptrs = (C.c_void_p * max(1, len(texs)))(*[t.ctypes.data for t in texs]) if texs else None
        r0, r1 = rows if rows else (0, self.height)
        img = np.zeros((self.height, self.width, 3), np.float64)
        lib().or_tapes_eval(self.h, C.byref(ps.s), C.cast(ptrs, C.c_void_p) if ptrs is not None else None,
                            e.ctypes.data, r0, r1, img.ctypes.data)
        return img

    def length(self, pixel, sample):
        return lib().or_tapes_len(self.h, pixel, sample)

    def close(self):
        if self.h:
            lib().or_tapes_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
