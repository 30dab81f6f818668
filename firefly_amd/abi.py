"""ctypes declarations of the C ABI (include/ff/ff_abi.h) and thin wrappers.

Device buffers are torch tensors on the GPU (PyTorch is only used for device
memory and streams); every compute call goes through libff_hip.so.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

FF_OK = 0
FF_MAT_FLOATS = 16
FF_MAT_MIX, FF_MAT_DIFFUSE, FF_MAT_CONDUCTOR, FF_MAT_DIELECTRIC = 0, 1, 2, 3
FF_RENDER_USE_GRAPH = 1
FF_RENDER_TIMING = 2
FF_ADJOINT_DETERMINISTIC = 4
# host SAH (the reference's tree) / linear BVH built on the GPU / PLOC clustering built on the GPU
FF_BVH_SAH, FF_BVH_LBVH, FF_BVH_PLOC = 0, 1, 2

NODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("bmax", "<f4", 3), ("offset", "<i4"),
                       ("n_prims", "<u2"), ("axis", "u1"), ("pad", "u1")])
assert NODE_DTYPE.itemsize == 32


class FFError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"ff status {status}: {msg}")
        self.status = status


class SceneDesc(C.Structure):
    _fields_ = [("n_tris", C.c_int32), ("positions", C.c_void_p), ("normals", C.c_void_p),
                ("uvs", C.c_void_p), ("material_ids", C.c_void_p), ("n_materials", C.c_int32),
                ("materials", C.c_void_p), ("n_textures", C.c_int32), ("textures", C.c_void_p),
                ("tex_w", C.c_void_p), ("tex_h", C.c_void_p), ("tex_c", C.c_void_p),
                ("max_prims_in_node", C.c_int32), ("builder", C.c_int32), ("tangents", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [("inv_view_proj", C.c_float * 16), ("lens_radius", C.c_float),
                ("focal_distance", C.c_float), ("exposure", C.c_float), ("pad", C.c_float)]


class RenderReq(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("max_depth", C.c_int32), ("seed", C.c_uint32), ("spp_batch", C.c_int32),
                ("sample_offset", C.c_int32), ("flags", C.c_int32)]


class Grads(C.Structure):
    _fields_ = [("env", C.c_void_p), ("n_textures", C.c_int32), ("textures", C.c_void_p)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("env", C.c_void_p), ("camera", Camera), ("width", C.c_int32),
                ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32), ("seed", C.c_uint32),
                ("sample_offset", C.c_int32), ("film", C.c_void_p), ("path_radiance", C.c_void_p),
                ("d_image", C.c_void_p), ("grads", C.c_void_p), ("spp_batch", C.c_int32)]


class RayStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "rays_primary", "rays_extension", "rays_shadow", "paths", "ext_nodes", "ext_tris",
        "shadow_nodes", "shadow_tris", "ms_extend", "ms_shadow", "ms_logic", "ms_shade", "ms_other",
        "launches_extend", "launches_shadow", "rays_shadow_gated")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_P = C.c_void_p
_PROTOS = {
    "ff_last_error": (C.c_char_p, []),
    "ff_abi_version": (C.c_int32, []),
    "ff_context_create": (C.c_int, [C.c_int32, C.POINTER(_P)]),
    "ff_context_destroy": (C.c_int, [_P]),
    "ff_bvh_build_host": (C.c_int, [_P, C.c_int32, C.c_int32, _P, _P, C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32)]),
    "ff_bvh_build_gpu": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _P, _P, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "ff_bvh_build_ploc": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _P, _P, C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "ff_scene_create": (C.c_int, [_P, C.POINTER(SceneDesc), C.POINTER(_P)]),
    "ff_scene_destroy": (C.c_int, [_P]),
    "ff_scene_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32)]),
    "ff_scene_get_bvh": (C.c_int, [_P, _P, _P]),
    "ff_envmap_create": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "ff_envmap_read": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "ff_envmap_destroy": (C.c_int, [_P]),
    "ff_trace_closest": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int32, _P, _P]),
    "ff_trace_any": (C.c_int, [_P, _P, C.c_uint32, _P, _P, _P]),
    "ff_trace_closest_noalloc": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int32, _P, _P]),
    "ff_generate_rays": (C.c_int, [_P, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_uint32,
                                   C.c_int32, _P, _P]),
    "ff_render": (C.c_int, [_P, _P, _P, C.POINTER(Camera), C.POINTER(RenderReq), _P, _P]),
    "ff_last_render_stats": (C.c_int, [_P, C.POINTER(RayStats)]),
    "ff_render_batch": (C.c_int, [_P, C.POINTER(BatchItem), C.c_int32, C.c_int32, _P]),
    "ff_envmap_update": (C.c_int, [_P, _P, _P]),
    "ff_scene_set_texture": (C.c_int, [_P, C.c_int32, _P, _P]),
    "ff_render_paths": (C.c_int, [_P, _P, _P, C.POINTER(Camera), C.POINTER(RenderReq), _P, _P, _P]),
    "ff_adjoint": (C.c_int, [_P, _P, _P, C.POINTER(Camera), C.POINTER(RenderReq), _P, _P, C.POINTER(Grads), _P]),
    "ff_adjoint_batch": (C.c_int, [_P, C.POINTER(BatchItem), C.c_int32, C.c_int32, _P]),
    "ff_traversal_variant_count": (C.c_int32, []),
    "ff_traversal_variant_info": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ff_set_traversal_variant": (C.c_int, [_P, C.c_int32]),
    "ff_set_traversal_tail": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "ff_film_resolve": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_float, _P, _P, _P]),
}

EXPORTED = sorted(_PROTOS)
_L: Optional[C.CDLL] = None


def load(path: str) -> C.CDLL:
    global _L
    L = C.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _L = L
    return L


def _lib() -> C.CDLL:
    if _L is None:
        from . import lib
        lib()
    return _L


def check(status: int) -> None:
    if status != FF_OK:
        raise FFError(status, _lib().ff_last_error().decode())


def _ptr(a) -> int:
    """Address of a numpy array or torch tensor."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def bvh_build_gpu(ctx: "Context", positions: np.ndarray, max_prims: int = 8, builder: int = FF_BVH_LBVH):
    """BVH built on the GPU, linear (FF_BVH_LBVH, ff_bvh_build_gpu) or by
    clustering (FF_BVH_PLOC, ff_bvh_build_ploc): (nodes, ordered, max_depth,
    device ms of the build kernels)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 9)
    n = pos.shape[0]
    nodes = np.zeros(max(2 * n - 1, 1), dtype=NODE_DTYPE)
    ordered = np.zeros(n, dtype=np.int32)
    nn, md, ms = C.c_int32(), C.c_int32(), C.c_float()
    fn = _lib().ff_bvh_build_ploc if builder == FF_BVH_PLOC else _lib().ff_bvh_build_gpu
    check(fn(ctx.h, pos.ctypes.data, n, max_prims, nodes.ctypes.data, ordered.ctypes.data,
             C.byref(nn), C.byref(md), C.byref(ms)))
    return nodes[: nn.value].copy(), ordered, md.value, ms.value


def bvh_build_host(positions: np.ndarray, max_prims: int = 8):
    """Host SAH BVH (no GPU): returns (nodes structured array, ordered, max_depth)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 9)
    n = pos.shape[0]
    nodes = np.zeros(max(2 * n - 1, 1), dtype=NODE_DTYPE)
    ordered = np.zeros(n, dtype=np.int32)
    nn, md = C.c_int32(), C.c_int32()
    check(_lib().ff_bvh_build_host(pos.ctypes.data, n, max_prims, nodes.ctypes.data,
                                   ordered.ctypes.data, C.byref(nn), C.byref(md)))
    return nodes[: nn.value].copy(), ordered, md.value


def make_camera(inv_vp_colmajor: Sequence[float], lens_radius=0.0, focal=1.0, exposure=1.0) -> Camera:
    cam = Camera()
    for i, v in enumerate(np.asarray(inv_vp_colmajor, dtype=np.float32).reshape(16)):
        cam.inv_view_proj[i] = float(v)
    cam.lens_radius = float(lens_radius)
    cam.focal_distance = float(focal)
    cam.exposure = float(exposure)
    return cam


class Context:
    def __init__(self, device: int = 0):
        h = _P()
        check(_lib().ff_context_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            check(_lib().ff_context_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Device scene built from host arrays (positions/normals/uvs [n,3,...]); the BVH by the
    reference's host SAH (FF_BVH_SAH), or on the GPU as a linear BVH (FF_BVH_LBVH) or by clustering (FF_BVH_PLOC)."""

    def __init__(self, ctx: Context, positions, normals, uvs, material_ids, materials,
                 textures: Sequence[np.ndarray] = (), max_prims: int = 8, builder: int = FF_BVH_SAH,
                 tangents=None):
        self.ctx = ctx
        self._keep = []
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 9)
        nrm = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 9)
        uv = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 6)
        mid = np.ascontiguousarray(material_ids, dtype=np.int32).reshape(-1)
        mats = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, FF_MAT_FLOATS)
        texs = [np.ascontiguousarray(t, dtype=np.float32) for t in textures]
        d = SceneDesc()
        d.n_tris = pos.shape[0]
        d.positions, d.normals = pos.ctypes.data, nrm.ctypes.data
        d.uvs = None if uv is None else uv.ctypes.data
        d.material_ids = mid.ctypes.data
        d.n_materials = mats.shape[0]
        d.materials = mats.ctypes.data
        d.n_textures = len(texs)
        if texs:
            ptrs = (C.c_void_p * len(texs))(*[t.ctypes.data for t in texs])
            tw = np.array([t.shape[1] for t in texs], dtype=np.int32)
            th = np.array([t.shape[0] for t in texs], dtype=np.int32)
            tc = np.array([t.shape[2] if t.ndim == 3 else 1 for t in texs], dtype=np.int32)
            d.textures = C.cast(ptrs, C.c_void_p)
            d.tex_w, d.tex_h, d.tex_c = tw.ctypes.data, th.ctypes.data, tc.ctypes.data
            self._keep += [ptrs, tw, th, tc]
        d.max_prims_in_node = max_prims
        d.builder = builder
        tan = None if tangents is None else np.ascontiguousarray(tangents, dtype=np.float32).reshape(-1, 18)
        if tan is not None and tan.shape[0] != pos.shape[0]:
            raise ValueError("tangents: one [3][6] record per triangle")
        d.tangents = None if tan is None else tan.ctypes.data
        self._keep += [pos, nrm, uv, mid, mats, texs, tan]
        h = _P()
        check(_lib().ff_scene_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self._keep = []
        nt, nn, md = C.c_int32(), C.c_int32(), C.c_int32()
        check(_lib().ff_scene_info(h, C.byref(nt), C.byref(nn), C.byref(md)))
        self.n_tris, self.n_nodes, self.max_depth = nt.value, nn.value, md.value

    def bvh(self):
        nodes = np.zeros(self.n_nodes, dtype=NODE_DTYPE)
        ordered = np.zeros(self.n_tris, dtype=np.int32)
        check(_lib().ff_scene_get_bvh(self.h, nodes.ctypes.data, ordered.ctypes.data))
        return nodes, ordered

    def trace_closest(self, rays_dev, hits_dev, two_sided=False, counts_dev=None, stream=None):
        n = rays_dev.shape[0]
        check(_lib().ff_trace_closest(self.h, _ptr(rays_dev), n, _ptr(hits_dev), int(two_sided),
                                      _ptr(counts_dev), stream))

    def trace_closest_noalloc(self, rays_dev, hits_dev, two_sided=False, counts_dev=None, stream=None):
        """ff_trace_closest_noalloc: as trace_closest, with no device allocation."""
        n = rays_dev.shape[0]
        check(_lib().ff_trace_closest_noalloc(self.h, _ptr(rays_dev), n, _ptr(hits_dev), int(two_sided),
                                              _ptr(counts_dev), stream))

    def trace_any(self, rays_dev, occ_dev, counts_dev=None, stream=None):
        n = rays_dev.shape[0]
        check(_lib().ff_trace_any(self.h, _ptr(rays_dev), n, _ptr(occ_dev), _ptr(counts_dev), stream))

    def set_texture(self, t: int, data, stream=None):
        """New texels for texture t (host array or device tensor, same shape)."""
        if isinstance(data, np.ndarray):
            data = np.ascontiguousarray(data, dtype=np.float32)
        check(_lib().ff_scene_set_texture(self.h, t, _ptr(data), stream))

    def close(self):
        if self.h:
            check(_lib().ff_scene_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnvMap:
    def __init__(self, ctx: Context, rgb: np.ndarray):
        a = np.ascontiguousarray(rgb, dtype=np.float32)
        h_, w_ = a.shape[0], a.shape[1]
        h = _P()
        check(_lib().ff_envmap_create(ctx.h, a.ctypes.data, w_, h_, C.byref(h)))
        self.h, self.w, self.hgt = h, w_, h_

    def read(self):
        my = np.zeros(self.hgt, np.float64)
        mp = np.zeros(self.hgt, np.float64)
        conds = np.zeros((self.hgt, self.w), np.float64)
        pdf = np.zeros((self.hgt, self.w), np.float32)
        denom = np.zeros(1, np.float64)
        check(_lib().ff_envmap_read(self.h, my.ctypes.data, mp.ctypes.data, conds.ctypes.data,
                                    pdf.ctypes.data, denom.ctypes.data))
        return my, mp, conds, pdf, float(denom[0])

    def update(self, rgb, stream=None):
        """New texels (host array or device tensor [h][w][3]); CDFs rebuilt."""
        if isinstance(rgb, np.ndarray):
            rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        check(_lib().ff_envmap_update(self.h, _ptr(rgb), stream))

    def close(self):
        if self.h:
            check(_lib().ff_envmap_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render(ctx: Context, scene: Scene, env: EnvMap, cam: Camera, width: int, height: int,
           spp: int, max_depth: int, seed: int, film_dev, spp_batch: int = 1,
           sample_offset: int = 0, flags: int = 0, stream=None) -> None:
    r = RenderReq(width, height, spp, max_depth, seed, spp_batch, sample_offset, flags)
    check(_lib().ff_render(ctx.h, scene.h, env.h, C.byref(cam), C.byref(r), _ptr(film_dev), stream))


def _per_item(v, n):
    return list(v) if isinstance(v, (list, tuple)) else [v] * n


def render_batch(ctx: Context, items, flags: int = 0, stream=None, path_radiance=None, spp_batch=1) -> None:
    """items: sequence of (scene, env, camera, width, height, spp, max_depth,
    seed, film_dev[, sample_offset]); path_radiance: optional per-item
    [spp][h][w][4] device buffers for the adjoint; spp_batch: samples of each
    pixel in flight, one value or one per item."""
    arr = (BatchItem * len(items))()
    kb = _per_item(spp_batch, len(items))
    for k, it in enumerate(items):
        sc, env, cam, w, h, spp, depth, seed, film = it[:9]
        off = it[9] if len(it) > 9 else 0
        pr = _ptr(path_radiance[k]) if path_radiance is not None else None
        arr[k] = BatchItem(sc.h, env.h, cam, w, h, spp, depth, seed, off, _ptr(film), pr, None, None, kb[k])
    check(_lib().ff_render_batch(ctx.h, arr, len(items), flags, stream))


def make_grads(env_grad=None, tex_grads=None):
    """ff_grads for device buffers (torch tensors); tex_grads: list indexed by
    scene texture (None entries skipped).  Returns (Grads, keep-alive)."""
    tex_grads = list(tex_grads or [])
    arr = (C.c_void_p * max(1, len(tex_grads)))(*[_ptr(t) for t in tex_grads]) if tex_grads else None
    g = Grads(_ptr(env_grad), len(tex_grads), C.cast(arr, C.c_void_p) if arr is not None else None)
    return g, arr


def render_paths(ctx: Context, scene: Scene, env: EnvMap, cam: Camera, width: int, height: int, spp: int,
                 max_depth: int, seed: int, film_dev, path_radiance, spp_batch: int = 1, sample_offset: int = 0,
                 flags: int = 0, stream=None) -> None:
    """ff_render that also stores every path's radiance ([spp][h][w][4])."""
    r = RenderReq(width, height, spp, max_depth, seed, spp_batch, sample_offset, flags)
    check(_lib().ff_render_paths(ctx.h, scene.h, env.h, C.byref(cam), C.byref(r), _ptr(film_dev),
                                 _ptr(path_radiance), stream))


def adjoint(ctx: Context, scene: Scene, env: EnvMap, cam: Camera, width: int, height: int, spp: int,
            max_depth: int, seed: int, path_radiance, d_image, env_grad=None, tex_grads=None, spp_batch: int = 1,
            sample_offset: int = 0, flags: int = 0, stream=None) -> None:
    """Path-replay adjoint: ADDS d loss / d env texel and d loss / d texel of
    each scene texture to env_grad / tex_grads[t] (device buffers), given
    d_image = d loss / d mean image ([h][w][4], xyz used)."""
    r = RenderReq(width, height, spp, max_depth, seed, spp_batch, sample_offset, flags)
    g, keep = make_grads(env_grad, tex_grads)
    check(_lib().ff_adjoint(ctx.h, scene.h, env.h, C.byref(cam), C.byref(r), _ptr(path_radiance), _ptr(d_image),
                            C.byref(g), stream))
    del keep


def adjoint_batch(ctx: Context, items, path_radiance, d_images, env_grads=None, tex_grads=None, flags: int = 0,
                  stream=None, spp_batch=1) -> None:
    """Batched adjoint of render_batch items; env_grads / tex_grads are per
    item (items may share buffers)."""
    n = len(items)
    arr = (BatchItem * n)()
    kb = _per_item(spp_batch, n)
    keep = []
    for k, it in enumerate(items):
        sc, env, cam, w, h, spp, depth, seed = it[:8]
        off = it[9] if len(it) > 9 else 0
        g, a = make_grads(env_grads[k] if env_grads is not None else None,
                          tex_grads[k] if tex_grads is not None else None)
        keep.append((g, a))
        arr[k] = BatchItem(sc.h, env.h, cam, w, h, spp, depth, seed, off, None, _ptr(path_radiance[k]),
                           _ptr(d_images[k]), C.cast(C.pointer(g), C.c_void_p), kb[k])
    check(_lib().ff_adjoint_batch(ctx.h, arr, n, flags, stream))
    del keep


def traversal_variants():
    """[(variant, block, lds_entries, refill, while_while, bvh_width)] of the traversal kernels."""
    out = []
    for v in range(_lib().ff_traversal_variant_count()):
        b, l, r, w, k = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(_lib().ff_traversal_variant_info(v, C.byref(b), C.byref(l), C.byref(r), C.byref(w), C.byref(k)))
        out.append((v, b.value, l.value, r.value, w.value, k.value))
    return out


def set_traversal_variant(ctx: Context, variant: int) -> None:
    check(_lib().ff_set_traversal_variant(ctx.h, variant))


def set_traversal_tail(ctx: Context, variant: int, from_iteration: int = 4) -> None:
    """Traversal variant of the bounce iterations from `from_iteration` on
    (-1: the context's variant everywhere)."""
    check(_lib().ff_set_traversal_tail(ctx.h, variant, from_iteration))


def last_render_stats(ctx: Context) -> RayStats:
    s = RayStats()
    check(_lib().ff_last_render_stats(ctx.h, C.byref(s)))
    return s


def generate_rays(ctx: Context, cam: Camera, width: int, height: int, seed: int, sample: int,
                  rays_dev, stream=None) -> None:
    check(_lib().ff_generate_rays(ctx.h, C.byref(cam), width, height, seed, sample, _ptr(rays_dev), stream))


def film_resolve(film_dev, width: int, height: int, exposure: float, avg_dev=None, rgba_dev=None,
                 stream=None) -> None:
    check(_lib().ff_film_resolve(_ptr(film_dev), width, height, exposure, _ptr(avg_dev),
                                 _ptr(rgba_dev), stream))
