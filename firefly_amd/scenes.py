"""Procedural scene inputs for the benchmark configurations.

Host-side input generation only (numpy): triangle soups with smooth normals,
material tables, environment maps and camera matrices.  World is y-up, as the
reference's camera and spherical mapping are (CRT/Camera.h:31,
CRT/dMath.cuh:1483-1501).  Everything is a deterministic function of the seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .abi import FF_MAT_CONDUCTOR, FF_MAT_DIELECTRIC, FF_MAT_DIFFUSE, FF_MAT_FLOATS, FF_MAT_MIX


@dataclass
class SceneData:
    positions: np.ndarray      # [n,3,3] float32
    normals: np.ndarray        # [n,3,3] float32
    uvs: np.ndarray            # [n,3,2] float32
    material_ids: np.ndarray   # [n] int32
    materials: np.ndarray      # [m,16] float32
    env: np.ndarray            # [h,w,3] float32
    inv_vp: np.ndarray         # [16] float32 column-major inverse(proj*view)
    width: int
    height: int
    spp: int
    max_depth: int
    seed: int
    textures: List[np.ndarray] = field(default_factory=list)
    lens_radius: float = 0.0
    focal: float = 1.0
    exposure: float = 1.0
    name: str = ""
    cam_eye: tuple = (0.0, 0.0, 0.0)
    cam_target: tuple = (0.0, 0.0, -1.0)
    cam_fovy: float = 45.0
    mesh_sizes: tuple = ()     # triangles per mesh, in order
    mesh_materials: tuple = ()  # material index of each mesh
    tangents: Optional[np.ndarray] = None  # [n,3,6] tangent, bitangent per vertex (normal textures)

    @property
    def n_tris(self) -> int:
        return int(self.positions.shape[0])


# ------------------------------------------------------------------ camera --
def perspective(fovy_rad: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    """Right-handed perspective with clip z in [-1, 1] (math layout, float64)."""
    f = 1.0 / np.tan(fovy_rad / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = -(zfar + znear) / (zfar - znear)
    m[2, 3] = -(2.0 * zfar * znear) / (zfar - znear)
    m[3, 2] = -1.0
    return m


def look_at(eye, center, up) -> np.ndarray:
    eye, center, up = (np.asarray(v, dtype=np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def inv_view_proj(eye, center, fovy_deg: float, aspect: float, znear=0.1, zfar=100.0) -> np.ndarray:
    """Column-major float32 inverse(proj*view), the dCamera matrix (CRT/dScene.cpp:381-391)."""
    vp = perspective(np.radians(fovy_deg), aspect, znear, zfar) @ look_at(eye, center, (0, 1, 0))
    inv = np.linalg.inv(vp)
    return np.ascontiguousarray(inv.T.reshape(16), dtype=np.float32)


# ---------------------------------------------------------------- geometry --
def _orient(pos: np.ndarray, nrm: np.ndarray, uv: np.ndarray, outward: np.ndarray):
    """Swap v1/v2 where the face normal cross(v1-v0, v2-v0) opposes `outward`
    so that every triangle is front-facing from outside (culling test)."""
    e1 = pos[:, 1] - pos[:, 0]
    e2 = pos[:, 2] - pos[:, 0]
    flip = np.einsum("ij,ij->i", np.cross(e1, e2), outward) < 0
    for a in (pos, nrm, uv):
        a[flip, 1], a[flip, 2] = a[flip, 2].copy(), a[flip, 1].copy()
    return pos, nrm, uv


def uv_sphere(center, radius: float, segments: int = 32, rings: int = 16):
    """UV sphere with triangle-fan poles: segments*(2*(rings-2)+2) triangles."""
    c = np.asarray(center, dtype=np.float64)
    th = np.pi * np.arange(rings + 1) / rings
    ph = 2.0 * np.pi * np.arange(segments + 1) / segments

    def vert(i, j):
        d = np.stack([np.sin(th[i]) * np.cos(ph[j]), np.cos(th[i]), np.sin(th[i]) * np.sin(ph[j])], -1)
        return c + radius * d, d, np.stack([ph[j] / (2 * np.pi), th[i] / np.pi], -1)

    tris = []
    for i in range(rings):
        for j in range(segments):
            a, b, cc, d = (i, j), (i, j + 1), (i + 1, j), (i + 1, j + 1)
            if i == 0:
                tris.append((a, d, cc))
            elif i == rings - 1:
                tris.append((a, b, cc))
            else:
                tris.append((a, b, d))
                tris.append((a, d, cc))
    idx = np.array(tris)  # [n,3,2]
    p, n, t = vert(idx[..., 0], idx[..., 1])
    pos, nrm, uv = _orient(p, n, t, p.mean(axis=1) - c)
    return pos.astype(np.float32), nrm.astype(np.float32), uv.astype(np.float32)


def ground_quad(y: float, half: float):
    p = np.array([[[-half, y, -half], [-half, y, half], [half, y, half]],
                  [[-half, y, -half], [half, y, half], [half, y, -half]]], dtype=np.float64)
    n = np.tile(np.array([0.0, 1.0, 0.0]), (2, 3, 1))
    uv = (p[..., [0, 2]] + half) / (2 * half)
    pos, nrm, uv = _orient(p, n, uv, np.tile([0.0, 1.0, 0.0], (2, 1)))
    return pos.astype(np.float32), nrm.astype(np.float32), uv.astype(np.float32)


def icosphere(subdiv: int):
    """Unit icosphere: vertices [V,3] float64 and faces [20*4^s,3] (CCW outward),
    plus the base icosahedron face of every triangle."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    region = np.arange(20, dtype=np.int64)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(e, axis=1)
        uniq, inv = np.unique(key[:, 0] * len(v) + key[:, 1], return_inverse=True)
        a, b = uniq // len(v), uniq % len(v)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        m = inv.reshape(3, -1).T + base  # midpoints of edges 01, 12, 20
        f = np.concatenate([np.stack([f[:, 0], m[:, 0], m[:, 2]], 1), np.stack([f[:, 1], m[:, 1], m[:, 0]], 1),
                            np.stack([f[:, 2], m[:, 2], m[:, 1]], 1), m])
        region = np.concatenate([region] * 4)
    return v, f, region


def value_noise3(p: np.ndarray, freq: float, seed: int) -> np.ndarray:
    """Seeded lattice value noise in [-1,1] with smoothstep trilinear blending."""
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1.0, 1.0, 4096)
    perm = rng.permutation(4096)
    q = p * freq
    i = np.floor(q).astype(np.int64)
    fr = q - i
    s = fr * fr * (3 - 2 * fr)

    def h(ix, iy, iz):
        return table[perm[(perm[(perm[ix & 4095] + iy) & 4095] + iz) & 4095]]

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) *
                     (s[:, 2] if dz else 1 - s[:, 2]))
                out = out + w * h(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
    return out


def displaced_icosphere(subdiv: int, center, radius: float, amp: float, seed: int, freq: float = 4.0):
    v, f, region = icosphere(subdiv)
    disp = 1.0 + amp * value_noise3(v, freq, seed)
    pv = v * (radius * disp)[:, None] + np.asarray(center, float)
    # smooth normals: area-weighted face normals accumulated per vertex
    fn = np.cross(pv[f[:, 1]] - pv[f[:, 0]], pv[f[:, 2]] - pv[f[:, 0]])
    vn = np.zeros_like(pv)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    uvv = np.stack([np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi) + 0.5, np.arccos(np.clip(v[:, 1], -1, 1)) / np.pi], 1)
    pos = pv[f].astype(np.float32)
    nrm = vn[f].astype(np.float32)
    uv = uvv[f].astype(np.float32)
    return pos, nrm, uv, region


def tangent_frames(normals: np.ndarray) -> np.ndarray:
    """Per-vertex tangent frames [n,3,6] (tangent xyz, bitangent xyz) for
    normals [n,3,3]: the tangent is the unit vector orthogonal to the normal
    in the plane of the normal and the world axis least aligned with it;
    bitangent = normal x tangent."""
    n = normals.astype(np.float64)
    ax = np.eye(3)[np.argmin(np.abs(n), axis=-1)]
    t = ax - n * np.sum(ax * n, axis=-1, keepdims=True)
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    b = np.cross(n, t)
    return np.concatenate([t, b], axis=-1).astype(np.float32)


# ---------------------------------------------------------------- materials --
def material(mtype: int, albedo=(0.5, 0.5, 0.5), roughness=1.0, metallic=0.0, fresnel=0.04,
             ior=1.5, albedo_tex=-1, rough_tex=-1, metal_tex=-1, normal_tex=-1) -> np.ndarray:
    m = np.zeros(FF_MAT_FLOATS, np.float32)
    m[0] = mtype
    m[1:4] = albedo
    m[4] = roughness
    m[5] = metallic
    m[6:9] = fresnel
    m[9] = ior
    m[10] = albedo_tex
    m[11] = rough_tex
    m[12] = metal_tex
    m[13] = normal_tex
    return m


# ----------------------------------------------------------------- env maps --
def texel_dirs(w: int, h: int) -> np.ndarray:
    """Direction of each texel centre under sample_spherical_direction
    (CRT/dMath.cuh:1491-1501): [h,w,3]."""
    u = (np.arange(w) + 0.5) / w
    v = (np.arange(h) + 0.5) / h
    th = 2 * np.pi * u[None, :]
    ph = np.pi * v[:, None]
    return -np.stack([np.cos(th) * np.sin(ph), np.cos(ph) * np.ones_like(th), np.sin(th) * np.sin(ph)], -1)


def sky_env(w: int = 512, h: int = 256) -> np.ndarray:
    v = (np.arange(h) + 0.5) / h
    el = np.pi * (v - 0.5)
    val = 0.5 + 0.3 * np.cos(el)
    return np.repeat(np.repeat(val[:, None, None], w, 1), 3, 2).astype(np.float32)


def sun_env(w: int, h: int, seed: int, sigma_deg=0.5, peak=1e4, base=0.5) -> np.ndarray:
    rng = np.random.default_rng(seed + 7)
    az = rng.uniform(0, 2 * np.pi)
    elv = rng.uniform(np.radians(25), np.radians(60))
    sd = np.array([np.cos(elv) * np.cos(az), np.sin(elv), np.cos(elv) * np.sin(az)])
    d = texel_dirs(w, h)
    ang = np.arccos(np.clip(d @ sd, -1, 1))
    sig = np.radians(sigma_deg)
    val = base + peak * np.exp(-0.5 * (ang / sig) ** 2)
    return np.repeat(val[..., None], 3, 2).astype(np.float32)


# ----------------------------------------------------------------- configs --
def _concat(parts):
    pos = np.concatenate([p[0] for p in parts])
    nrm = np.concatenate([p[1] for p in parts])
    uv = np.concatenate([p[2] for p in parts])
    mid = np.concatenate([np.full(len(p[0]), p[3], np.int32) if np.isscalar(p[3]) else p[3] for p in parts])
    return pos, nrm, uv, mid.astype(np.int32)


def spheres_scene(width=256, height=256, spp=16, max_depth=4, seed=0, n_spheres=8,
                  segments=32, rings=16, mtype=FF_MAT_MIX) -> SceneData:
    """Ground quad + 8 UV spheres under a procedural sky (256x256, 16 spp, depth 4).
    Materials follow what the reference loader produces for a plain albedo
    (roughness 1, metallic 0, CRT/Scene.cpp:203-204) on the mixed lobe pair."""
    rng = np.random.default_rng(seed)
    parts = [(*ground_quad(0.0, 10.0), 0)]
    mats = [material(mtype, (0.5, 0.5, 0.5))]
    for k in range(n_spheres):
        cx, cz = rng.uniform(-1, 1, 2)
        r = rng.uniform(0.2, 0.4)
        alb = rng.uniform(0.2, 0.8, 3)
        parts.append((*uv_sphere((cx, 0.3, cz), r, segments, rings), k + 1))
        mats.append(material(mtype, alb))
    pos, nrm, uv, mid = _concat(parts)
    eye, target = (0.0, 1.2, 3.5), (0.0, 0.0, 0.0)
    return SceneData(pos, nrm, uv, mid, np.stack(mats), sky_env(512, 256),
                     inv_view_proj(eye, target, 45.0, width / height),
                     width, height, spp, max_depth, seed, name="spheres256", cam_eye=eye, cam_target=target,
                     cam_fovy=45.0, mesh_sizes=tuple(len(p[0]) for p in parts),
                     mesh_materials=tuple(p[3] for p in parts))


def icosphere_scene(width=1920, height=1080, spp=64, max_depth=8, seed=0, subdiv=8,
                    env_w=2048, env_h=1024) -> SceneData:
    """Displaced icosphere (20*4^subdiv tris) + ground quad; three material
    queues (diffuse, GGX conductor, dielectric) assigned per base-icosahedron
    face; 2048x1024 HDR sky with a sigma=0.5 deg sun lobe of peak 1e4."""
    rng = np.random.default_rng(seed)
    pos, nrm, uv, region = displaced_icosphere(subdiv, (0.0, 1.1, 0.0), 1.0, 0.05, seed)
    mats = [material(FF_MAT_DIFFUSE, (0.5, 0.5, 0.5))]  # ground
    region_mat = np.zeros(20, np.int32)
    for r in range(20):
        kind = r % 3
        if kind == 0:
            mats.append(material(FF_MAT_DIFFUSE, rng.uniform(0.1, 0.9, 3)))
        elif kind == 1:
            f0 = rng.uniform(0.5, 1.0, 3)
            mats.append(material(FF_MAT_CONDUCTOR, f0, roughness=rng.uniform(0.05, 0.5), metallic=1.0))
        else:
            mats.append(material(FF_MAT_DIELECTRIC, (1, 1, 1), ior=1.5))
        region_mat[r] = len(mats) - 1
    parts = [(*ground_quad(0.0, 20.0), 0), (pos, nrm, uv, region_mat[region])]
    P, N, U, M = _concat(parts)
    eye, target = (0.0, 1.4, 4.2), (0.0, 1.0, 0.0)
    return SceneData(P, N, U, M, np.stack(mats), sun_env(env_w, env_h, seed),
                     inv_view_proj(eye, target, 40.0, width / height),
                     width, height, spp, max_depth, seed, name=f"icosphere{subdiv}", cam_eye=eye,
                     cam_target=target, cam_fovy=40.0)


def textured_spheres_scene(width=512, height=512, spp=16, max_depth=4, seed=0, tex_size=512, env_w=1024,
                           env_h=512, segments=32, rings=16) -> SceneData:
    """Differentiable single view: the 8-sphere scene with the spheres'
    albedo read from one tex_size^2 RGB texture of U[0.2, 0.8] texels
    (texture 0) under an env_w x env_h sky."""
    sd = spheres_scene(width, height, spp, max_depth, seed, segments=segments, rings=rings)
    rng = np.random.default_rng(seed + 101)
    sd.textures = [rng.uniform(0.2, 0.8, (tex_size, tex_size, 3)).astype(np.float32)]
    mats = sd.materials.copy()
    mats[1:, 10] = 0
    sd.materials = mats
    sd.env = sky_env(env_w, env_h)
    sd.name = f"tspheres{width}"
    return sd


def perturbed(sd: SceneData, seed: int = 1, sigma: float = 0.05):
    """Parameters of the target image: textures and environment plus
    N(0, sigma) noise (textures clipped to [0.01, 1])."""
    rng = np.random.default_rng(seed)
    tex = [np.clip(t + rng.normal(0.0, sigma, t.shape), 0.01, 1.0).astype(np.float32) for t in sd.textures]
    env = np.maximum(sd.env + rng.normal(0.0, sigma, sd.env.shape), 0.0).astype(np.float32)
    return tex, env


def view_directions(n: int):
    """n Fibonacci-spiral directions over the upper hemisphere (y >= 0.1)."""
    k = np.arange(n) + 0.5
    y = 0.1 + 0.85 * k / n
    phi = k * np.pi * (3.0 - np.sqrt(5.0))
    r = np.sqrt(1.0 - y * y)
    return np.stack([r * np.cos(phi), y, r * np.sin(phi)], 1)


def multiview_scene(views=32, side=256, spp=8, max_depth=6, seed=0, subdiv=8, tex_size=1024,
                    env_w=2048, env_h=1024, radius=4.0):
    """Differentiable multi-view: the displaced icosphere (20*4^subdiv tris)
    and ground quad, the sphere on one mixed GGX/diffuse material whose
    roughness (U[0.05, 0.6], texture 1) and RGB albedo (texture 0) are
    tex_size^2 textures, seen from `views` cameras on a sphere of `radius`
    around it.  Returns (SceneData of view 0, [inverse view-projection of
    every view])."""
    rng = np.random.default_rng(seed + 202)
    pos, nrm, uv, _ = displaced_icosphere(subdiv, (0.0, 1.1, 0.0), 1.0, 0.05, seed)
    mats = [material(FF_MAT_DIFFUSE, (0.5, 0.5, 0.5)),
            material(FF_MAT_MIX, (0.5, 0.5, 0.5), roughness=0.3, metallic=0.25, albedo_tex=0, rough_tex=1)]
    P, N, U, M = _concat([(*ground_quad(0.0, 20.0), 0), (pos, nrm, uv, 1)])
    textures = [rng.uniform(0.2, 0.8, (tex_size, tex_size, 3)).astype(np.float32),
                rng.uniform(0.05, 0.6, (tex_size, tex_size, 1)).astype(np.float32)]
    target = np.array([0.0, 1.1, 0.0])
    vps = [inv_view_proj(tuple(target + radius * d), tuple(target), 40.0, 1.0) for d in view_directions(views)]
    eye = tuple(target + radius * view_directions(views)[0])
    sd = SceneData(P, N, U, M, np.stack(mats), sun_env(env_w, env_h, seed), vps[0], side, side, spp, max_depth,
                   seed, textures=textures, name=f"multiview{views}", cam_eye=eye, cam_target=tuple(target),
                   cam_fovy=40.0)
    return sd, vps


def soup_in_cube(n: int, seed: int, edge: float = 0.01) -> tuple:
    """n triangles with edges ~`edge`, uniform in the unit cube, random facing."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (n, 1, 3))
    pos = c + rng.normal(scale=edge / np.sqrt(2), size=(n, 3, 3))
    fn = np.cross(pos[:, 1] - pos[:, 0], pos[:, 2] - pos[:, 0])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-12)
    nrm = np.repeat(fn[:, None, :], 3, 1)
    uv = rng.uniform(0, 1, (n, 3, 2))
    return pos.astype(np.float32), nrm.astype(np.float32), uv.astype(np.float32)


def batch_request_specs(n: int = 64, seed: int = 0):
    """Seeded uneven requests: side 2^U[7,10] (square), spp U{1..32}, depth
    U{2..12}, triangle count log-uniform in [1e4, 2e6]."""
    rng = np.random.default_rng(seed + 1000)
    out = []
    for i in range(n):
        side = int(round(2.0 ** rng.uniform(7.0, 10.0)))
        out.append({"index": i, "width": side, "height": side, "spp": int(rng.integers(1, 33)),
                    "max_depth": int(rng.integers(2, 13)),
                    "tris": int(round(10.0 ** rng.uniform(4.0, np.log10(2e6)))), "seed": seed * 1000 + i})
    return out


def batch_request_scene(spec: dict, env: np.ndarray) -> SceneData:
    """One request's scene: a displaced icosphere (largest subdivision under
    70% of the triangle budget) in the unit cube plus a triangle soup with
    edge ~0.01 making up the rest; materials as in the single-frame scene."""
    n = spec["tris"]
    sub = 2
    while 20 * 4 ** (sub + 1) <= 0.7 * n:
        sub += 1
    rng = np.random.default_rng(spec["seed"])
    pos, nrm, uv, region = displaced_icosphere(sub, (0.5, 0.5, 0.5), 0.3, 0.05, spec["seed"])
    mats = [material(FF_MAT_DIFFUSE, rng.uniform(0.1, 0.9, 3))]  # soup
    region_mat = np.zeros(20, np.int32)
    for r in range(20):
        kind = r % 3
        if kind == 0:
            mats.append(material(FF_MAT_DIFFUSE, rng.uniform(0.1, 0.9, 3)))
        elif kind == 1:
            mats.append(material(FF_MAT_CONDUCTOR, rng.uniform(0.5, 1.0, 3), roughness=rng.uniform(0.05, 0.5),
                                 metallic=1.0))
        else:
            mats.append(material(FF_MAT_DIELECTRIC, (1, 1, 1), ior=1.5))
        region_mat[r] = len(mats) - 1
    parts = [(pos, nrm, uv, region_mat[region])]
    n_soup = n - len(pos)
    if n_soup > 0:
        parts.append((*soup_in_cube(n_soup, spec["seed"] + 1), 0))
    P, N, U, M = _concat(parts)
    az, el = rng.uniform(0, 2 * np.pi), rng.uniform(0.1, 0.8)
    eye = (0.5 + 1.6 * np.cos(el) * np.cos(az), 0.5 + 1.6 * np.sin(el), 0.5 + 1.6 * np.cos(el) * np.sin(az))
    w, h = spec["width"], spec["height"]
    return SceneData(P, N, U, M, np.stack(mats), env, inv_view_proj(eye, (0.5, 0.5, 0.5), 45.0, w / h), w, h,
                     spec["spp"], spec["max_depth"], spec["seed"], name=f"request{spec['index']}", cam_eye=eye,
                     cam_target=(0.5, 0.5, 0.5), cam_fovy=45.0)


def write_scene_file(sd: SceneData, path: str) -> None:
    """Write the FFS1 scene file read by the C++ render_demo (one mesh per
    material run, in triangle order)."""
    mid = sd.material_ids
    cuts = np.flatnonzero(np.diff(mid)) + 1
    starts = np.concatenate([[0], cuts]) if len(mid) else np.zeros(0, np.int64)
    ends = np.concatenate([cuts, [len(mid)]]) if len(mid) else np.zeros(0, np.int64)
    with open(path, "wb") as f:
        f.write(b"FFS1")
        np.array([len(starts), len(sd.materials), sd.env.shape[1], sd.env.shape[0], sd.width, sd.height],
                 np.int32).tofile(f)
        np.array([sd.seed], np.uint32).tofile(f)
        np.array(list(sd.cam_eye) + list(sd.cam_target) + [sd.cam_fovy], np.float32).tofile(f)
        np.ascontiguousarray(sd.materials, np.float32).tofile(f)
        np.ascontiguousarray(sd.env, np.float32).tofile(f)
        for a, b in zip(starts, ends):
            np.array([mid[a], b - a], np.int32).tofile(f)
            for arr in (sd.positions, sd.normals, sd.uvs):
                np.ascontiguousarray(arr[a:b], np.float32).tofile(f)


def random_rays(n: int, seed: int, lo=-1.5, hi=1.5, y_lo=0.0, y_hi=1.5) -> np.ndarray:
    """Incoherent rays: uniform origins in a box, uniform directions; [n,8]."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(lo, hi, n), rng.uniform(y_lo, y_hi, n), rng.uniform(lo, hi, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    return r
