"""Expected pixel values of the composed path estimator on small scenes,
computed without sampling and without the oracle's path tracer, to pin the
composition of wf_logic + BRDF_L (CRT/kernel.cu:297-455,
CRT/dMaterial.cu:363-469):

* depth-1 direct lighting of a ground quad under a banded environment (and
  optionally a wall that blocks part of the sky): the light-sample term as a
  sum over every texel corner the sampler can return
  (dEnvironmentLight::get_direction, CRT/dEnvironmentLight.cu:41-58) weighted
  by its exact probability under the reference's upper_bound
  (CRT/Helpers.cu:16-31), and the two BRDF-sampled terms as midpoint
  quadratures over the sampler's two uniforms; the power-heuristic pairs,
  cosines and pdf divisions as BRDF_L writes them; the light ray's
  visibility gating the whole NEE sum (Q5) and the mixed material's diffuse
  visibility ray leaving along the spec sample (Q6).  With the light sample
  independent of the BRDF samples, the expectation is
  E[visL A] + P(visL) (E[visS Bs] + E[visD Cd]).
* Russian roulette through a dielectric sheet (IOR eta, beta = 1 at every
  vertex, constant environment): a closed-form sum over the number of
  internal reflections, with survival 0.95 per check at path lengths > 3 and
  no 1/(1-q) compensation (Q2).

The environment's CDF arrays come from the oracle's env_build (pinned by its
own closed-form tests) and the inverse-CDF indices from its upper_bound
(pinned by a hand-derived table); the lobe formulas are restated here in
float64 from CRT/dMaterial.cu:90-288."""
import numpy as np

PI = float(np.float32(np.pi))  # the reference's float M_PI
DIFF_PDF = 1.0 / (2.0 * PI)


def normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def dot(a, b):
    return np.sum(a * b, axis=-1)


# ---------------------------------------------------------------- lobes --
def ggx_ndf(n, h, r):
    a2 = r ** 4
    nh = np.maximum(0.0, dot(n, h))
    return a2 / (PI * (nh * nh * (a2 - 1.0) + 1.0) ** 2)


def geo_atten(wi, wo, n, r):
    k = (r + 1.0) ** 2 / 8.0
    nl = np.maximum(dot(n, wi), 0.0)
    nv = np.maximum(dot(n, wo), 0.0)
    return nl / (nl * (1 - k) + k) * (nv / (nv * (1 - k) + k))


def fresnel(f0, h, w):
    return f0 + (1.0 - f0) * ((1.0 - np.maximum(0.0, dot(h, w))) ** 5)[..., None]


def spec_pdf(n, wi, wo, r):
    wh = normalize(wo + wi)
    return ggx_ndf(n, wh, r) * np.maximum(dot(wh, n), 0.0) / np.maximum(4.0 * np.maximum(dot(wo, wh), 0.0), 0.001)


def spec_f(n, wi, wo, r, m, alb, fres):
    f0 = fres + (alb - fres) * m
    wh = normalize(wo + wi)
    ndwi = np.maximum(dot(n, wi), 0.0)
    ndwo = np.maximum(dot(n, wo), 0.0)
    dg = ggx_ndf(n, wh, r) * geo_atten(wi, wo, n, r)
    return fresnel(f0, wh, wi) * (dg / np.maximum(4.0 * ndwo * ndwi, 0.001))[..., None]


def diff_f(wi, wo, m, alb, fres):
    f0 = fres + (alb - fres) * m
    wh = normalize(wo + wi)
    return (1.0 - fresnel(f0, wh, wi)) * (1.0 - m) * alb / PI


def power(f, g):
    return f * f / (f * f + g * g)


# ---------------------------------------------------------- environment --
class Env:
    """Equirectangular map [h][w][3] with the reference's lookups."""

    def __init__(self, rgb, oracle):
        self.rgb = np.asarray(rgb, np.float32)
        self.h, self.w = self.rgb.shape[:2]
        self.O = oracle
        self.my, self.mp, self.conds, self.pdf_tex, _ = oracle.env_build(self.rgb)

    def sph_map(self, d):
        u = np.arctan2(d[..., 2], d[..., 0]) / (2 * PI) + 0.5
        v = np.arcsin(np.clip(d[..., 1], -1.0, 1.0)) / PI + 0.5
        return u, v

    def L(self, d):
        """dEnvironmentLight::L: bilinear, wrap, texel centres at (i + 0.5)/w."""
        u, v = self.sph_map(d)
        x = (u - np.floor(u)) * self.w - 0.5
        y = (v - np.floor(v)) * self.h - 0.5
        fx, fy = np.floor(x), np.floor(y)
        a, b = (x - fx)[..., None], (y - fy)[..., None]
        i0 = np.mod(fx.astype(np.int64), self.w)
        j0 = np.mod(fy.astype(np.int64), self.h)
        i1, j1 = (i0 + 1) % self.w, (j0 + 1) % self.h
        t = self.rgb.astype(np.float64)
        return ((1 - a) * (1 - b) * t[j0, i0] + a * (1 - b) * t[j0, i1] + (1 - a) * b * t[j1, i0] +
                a * b * t[j1, i1])

    def pdf(self, d):
        """get_pdf(isect, wi): pdf texel at (int)(u(w-1)), (int)(v(h-1)),
        divided by sin(acos(wi.z)) (CRT/dEnvironmentLight.cu:83-103)."""
        u, v = self.sph_map(d)
        px = np.clip((u * (self.w - 1)).astype(np.int64), 0, self.w - 1)
        py = np.clip((v * (self.h - 1)).astype(np.int64), 0, self.h - 1)
        with np.errstate(divide="ignore"):
            return self.pdf_tex[py, px].astype(np.float64) / np.sin(np.arccos(np.clip(d[..., 2], -1.0, 1.0)))

    def _pick_probs(self, lst):
        """{index: probability} of upper_bound(lst, e) - 1 for e uniform in
        [0, 1): the index is constant between consecutive list values."""
        brk = np.unique(np.clip(np.concatenate([[0.0, 1.0], lst[np.isfinite(lst)]]), 0.0, 1.0))
        out = {}
        for a, b in zip(brk[:-1], brk[1:]):
            if b <= a:
                continue
            k = self.O.upper_bound(lst, np.float32((a + b) / 2)) - 1
            out[max(k, 0)] = out.get(max(k, 0), 0.0) + (b - a)
        return out

    def corners(self):
        """[(probability, direction)] of every texel corner the light sampler
        returns: row y = upper_bound(marginal_y, ey) - 1, column x =
        upper_bound(conds[y], ex) - 1, direction sample_spherical_direction(
        x/w, y/h) (CRT/dMath.cuh:1491-1501)."""
        out = []
        for y, py in self._pick_probs(self.my).items():
            for x, px in self._pick_probs(self.conds[y]).items():
                u, v = np.float32(x) / np.float32(self.w), np.float32(y) / np.float32(self.h)
                th, ph = 2 * PI * float(u), PI * float(v)
                d = -np.array([np.cos(th) * np.sin(ph), np.cos(ph), np.sin(th) * np.sin(ph)])
                out.append((py * px, d))
        return out


# ------------------------------------------------------------- geometry --
class Blockers:
    """Occluders of the any-hit rays: the ground quad y = 0 (|x|, |z| <= g)
    and optionally a wall x = wx (y in [-1, wy], |z| <= wz)."""

    def __init__(self, ground_half, wall=None):
        self.g = ground_half
        self.wall = wall  # (wx, wy, wz) or None

    def visible(self, o, d):
        o = np.broadcast_to(o, d.shape)
        vis = np.ones(d.shape[:-1], bool)
        with np.errstate(divide="ignore", invalid="ignore"):
            t = -o[..., 1] / d[..., 1]
            p = o + t[..., None] * d
            hit = (t >= 0) & (np.abs(p[..., 0]) <= self.g) & (np.abs(p[..., 2]) <= self.g)
            vis &= ~hit
            if self.wall is not None:
                wx, wy, wz = self.wall
                t = (wx - o[..., 0]) / d[..., 0]
                p = o + t[..., None] * d
                hit = (t >= 0) & (p[..., 1] >= -1) & (p[..., 1] <= wy) & (np.abs(p[..., 2]) <= wz)
                vis &= ~hit
        return vis


def _frame(n):
    """get_orthogonal_vec frame (CRT/dMath.cuh:1450-1463) around n."""
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.57735026919 else (
        np.array([0, 1.0, 0]) if abs(n[1]) < 0.57735026919 else np.array([0, 0, 1.0]))
    t = normalize(np.cross(n, a))
    return t, normalize(np.cross(n, t))


def direct_expectation(env, blockers, P, wo, mtype, rough, metal, alb, fres=0.04, n_quad=1024):
    """Expected radiance (RGB) of a depth-1 path at ground point P seen from
    direction -wo, material type mtype (0 mix, 1 diffuse lobe, 2 GGX lobe),
    and a breakdown of the terms."""
    n = np.array([0.0, 1.0, 0.0])
    alb = np.asarray(alb, np.float64)
    fres = np.full(3, fres)
    do_spec, do_diff = mtype in (0, 2), mtype in (0, 1)
    # light-sample term over the sampler's texel corners
    E_L = np.zeros(3)
    P_L = 0.0
    lo = P + n * 0.001
    for p, d in env.corners():
        wi = normalize(d * 10000.0 - P)
        vis = bool(blockers.visible(lo, wi[None])[0])
        if not vis:
            continue
        P_L += p
        ndwi = float(n @ wi)
        if ndwi <= 0:
            continue
        lp = float(env.pdf(wi[None])[0])
        Li = env.L(wi[None])[0]
        if not (lp > 0 and np.isfinite(lp)) or not np.any(Li != 0):
            continue
        term = np.zeros(3)
        if do_spec:
            sp = float(spec_pdf(n[None], wi[None], wo[None], rough)[0])
            term += spec_f(n[None], wi[None], wo[None], rough, metal, alb, fres)[0] * ndwi * power(lp, sp)
        if do_diff:
            term += diff_f(wi[None], wo[None], metal, alb, fres)[0] * ndwi * power(lp, DIFF_PDF)
        E_L += p * term * Li / lp
    # BRDF-sampled terms: midpoint rule over the two uniforms of each sampler
    g = (np.arange(n_quad) + 0.5) / n_quad
    e0, e1 = np.meshgrid(g, g, indexing="ij")
    e0, e1 = e0.ravel(), e1.ravel()
    T, B = _frame(n)
    E_S = np.zeros(3)
    P_S = 1.0
    if do_spec:
        th = np.arctan(rough * rough * np.sqrt(e0 / (1.0 - e0)))
        ph = 2 * PI * e1
        h = normalize(T * (np.sin(th) * np.cos(ph))[:, None] + n * np.cos(th)[:, None] +
                      B * (np.sin(th) * np.sin(ph))[:, None])
        wi_s = 2.0 * dot(wo, h)[:, None] * h - wo
        nd = wi_s @ n
        sp = spec_pdf(n, wi_s, wo, rough)
        lp = env.pdf(wi_s)
        L2 = env.L(wi_s)
        vis = blockers.visible(P + 0.001 * wi_s, wi_s)
        ok = (nd > 0) & (sp > 0) & vis & np.any(L2 != 0, -1)
        with np.errstate(divide="ignore", invalid="ignore"):
            val = spec_f(n, wi_s, wo, rough, metal, alb, fres) * nd[:, None] * L2 * (power(sp, lp) / sp)[:, None]
        E_S = np.where(ok[:, None], val, 0.0).mean(0)
        if mtype == 0:  # the mixed material's diffuse ray starts at P + 0.001 wi_spec (Q6)
            P_S = float(np.mean(wi_s @ n > 0))
    E_D = np.zeros(3)
    if do_diff:
        st = np.sqrt(1.0 - e0 * e0)
        ph = 2 * PI * e1
        wi_d = normalize(T * (st * np.cos(ph))[:, None] + n * e0[:, None] + B * (st * np.sin(ph))[:, None])
        nd = wi_d @ n
        lp = env.pdf(wi_d)
        L3 = env.L(wi_d)
        vis = blockers.visible(P + 0.001 * wi_d, wi_d)  # the wall; the ground offset is in P_S
        ok = (nd > 0) & vis & np.any(L3 != 0, -1)
        val = diff_f(wi_d, wo, metal, alb, fres) * nd[:, None] * L3 * (power(DIFF_PDF, lp) / DIFF_PDF)[:, None]
        E_D = np.where(ok[:, None], val, 0.0).mean(0)
    total = E_L + P_L * (E_S + P_S * E_D)
    return total, {"E_L": E_L, "P_L": P_L, "E_S": E_S, "P_S": P_S, "E_D": E_D}


def fresnel_dielectric(cos_i, eta_i, eta_t):
    """Unpolarised Fresnel reflectance of a smooth interface (textbook)."""
    sin_t = eta_i / eta_t * np.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t >= 1:
        return 1.0
    cos_t = np.sqrt(1 - sin_t * sin_t)
    rs = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t)
    rp = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t)
    return 0.5 * (rs * rs + rp * rp)


def sheet_expectation(cos_i, eta, max_depth, env_value):
    """Expected radiance of a camera ray meeting a dielectric sheet (parallel
    faces) under a constant environment: escape after the first reflection,
    or after entering and j - 2 internal reflections (vertex j), each vertex
    with the same reflectance F; the logic's Russian roulette at every path
    length > 3 keeps a path with probability 1 - max(0.05, 1 - beta.y) = 0.95
    (beta stays 1) and does not rescale it."""
    F = fresnel_dielectric(cos_i, 1.0, eta)
    e = F
    for j in range(2, max_depth + 1):
        e += (1 - F) * F ** (j - 2) * (1 - F) * 0.95 ** max(0, j - 3)
    return e * np.asarray(env_value, np.float64), F


# ---------------------------------------------------------------- scenes --
def banded_env(w=32, h=16):
    """Four elevation bands (rows 0-3 nadir ... 12-15 zenith), the bright
    band brighter over half the azimuths, so rows and columns are both
    non-uniform."""
    env = np.zeros((h, w, 3), np.float32)
    bands = [(0.15, 0.2, 0.25), (0.4, 0.45, 0.5), (2.5, 2.0, 1.4), (0.9, 1.1, 1.6)]
    for k, c in enumerate(bands):
        env[k * h // 4:(k + 1) * h // 4] = c
    env[2 * h // 4:3 * h // 4, w // 4:3 * w // 4] *= 1.8
    return env


MATERIALS = {  # type: (albedo / f0, roughness, metallic)
    0: ((0.7, 0.5, 0.3), 0.45, 0.3),
    1: ((0.6, 0.6, 0.5), 1.0, 0.0),
    2: ((0.9, 0.7, 0.5), 0.35, 1.0),
}
WALL = (1.2, 500.0, 500.0)


def direct_scene(scenes, mtype, wall=False, size=2):
    """Ground quad of the material under test (+ optionally a wall at
    x = 1.2 facing it) under the banded environment, seen at 30 degrees
    elevation by a narrow camera: every pixel's camera ray hits the ground
    near (-0.6, 0, 0.6).  Depth 1."""
    alb, rough, metal = MATERIALS[mtype]
    parts = [(*scenes.ground_quad(0.0, 50.0), 0)]
    mats = [scenes.material(mtype, alb, roughness=rough, metallic=metal)]
    if wall:
        wx, wy, wz = WALL
        p = np.array([[[wx, -1, -wz], [wx, wy, -wz], [wx, wy, wz]], [[wx, -1, -wz], [wx, wy, wz], [wx, -1, wz]]],
                     np.float64)
        nrm = np.tile(np.array([-1.0, 0, 0]), (2, 3, 1))
        uv = np.zeros((2, 3, 2))
        p, nrm, uv = scenes._orient(p, nrm, uv, np.tile([-1.0, 0, 0], (2, 1)))
        parts.append((p.astype(np.float32), nrm.astype(np.float32), uv.astype(np.float32), 1))
        mats.append(scenes.material(1, (0.5, 0.5, 0.5)))
    pos, nrm, uv, mid = scenes._concat(parts)
    target = np.array([-0.6, 0.0, 0.6])
    eye = target + np.array([0.2, 1.3, 2.2])
    sd = scenes.SceneData(pos, nrm, uv, mid, np.stack(mats), banded_env(),
                          scenes.inv_view_proj(tuple(eye), tuple(target), 6.0, 1.0), size, size, 1, 1, 0,
                          name="direct", cam_eye=tuple(eye), cam_target=tuple(target), cam_fovy=6.0)
    return sd, Blockers(50.0, WALL if wall else None)


def sheet_scene(scenes, eta=4.0, size=2, max_depth=12, env_value=0.7):
    """Dielectric sheet between y = 0 and y = 0.1 (IOR eta, outward faces),
    constant environment, camera rays meeting it at about 45 degrees."""
    top = scenes.ground_quad(0.1, 200.0)
    bp, bn, bu = scenes.ground_quad(0.0, 200.0)
    bp, bn, bu = scenes._orient(bp.astype(np.float64).copy(), -bn.astype(np.float64), bu.astype(np.float64).copy(),
                                np.tile([0.0, -1.0, 0.0], (2, 1)))
    parts = [(*top, 0), (bp.astype(np.float32), bn.astype(np.float32), bu.astype(np.float32), 0)]
    pos, nrm, uv, mid = scenes._concat(parts)
    env = np.full((4, 8, 3), env_value, np.float32)
    target = np.array([0.0, 0.1, 0.3])
    eye = target + np.array([0.0, 1.5, 1.5])
    mats = np.stack([scenes.material(3, (1, 1, 1), ior=eta)])
    sd = scenes.SceneData(pos, nrm, uv, mid, mats, env,
                          scenes.inv_view_proj(tuple(eye), tuple(target), 4.0, 1.0), size, size, 1, max_depth, 0,
                          name="sheet", cam_eye=tuple(eye), cam_target=tuple(target), cam_fovy=4.0)
    return sd


def pixel_ray(ps, px, py):
    """Camera ray (origin, direction) of a pixel, sample 0 (the direction
    does not depend on the sample)."""
    r = ps.gen_ray(px, py, 0).astype(np.float64)
    return r[:3], r[3:]


def direct_expectations(O, scenes, mtype, wall=False, n_quad=768):
    """[(px, py, E, quadrature error bound)] of every pixel of direct_scene."""
    sd, blk = direct_scene(scenes, mtype, wall)
    ps = O.PreparedScene(sd)
    env = Env(sd.env, O)
    alb, rough, metal = MATERIALS[mtype]
    out = []
    for py in range(sd.height):
        for px in range(sd.width):
            o, d = pixel_ray(ps, px, py)
            P = o + d * (-o[1] / d[1])
            E, _ = direct_expectation(env, blk, P, -d, mtype, rough, metal, alb, n_quad=n_quad)
            E2, _ = direct_expectation(env, blk, P, -d, mtype, rough, metal, alb, n_quad=n_quad // 2)
            out.append((px, py, E, np.abs(E - E2)))
    return sd, out


def sheet_expectations(O, scenes, eta=4.0, max_depth=12, env_value=0.7):
    sd = sheet_scene(scenes, eta, max_depth=max_depth, env_value=env_value)
    ps = O.PreparedScene(sd)
    out = []
    for py in range(sd.height):
        for px in range(sd.width):
            _, d = pixel_ray(ps, px, py)
            E, _ = sheet_expectation(abs(d[1]), eta, max_depth, [env_value] * 3)
            out.append((px, py, E, np.zeros(3)))
    return sd, out


def compare(films, spp, expect, sigmas=4.0):
    """films: per-batch films [H][W][4] of spp samples each (independent
    seeds).  Returns (worst |z|, rows of (px, py, channel, mean, E, sigma))
    where z = (mean - E) / (sigma + quadrature bound / sigmas)."""
    imgs = np.stack([f[..., :3].astype(np.float64) / spp for f in films])
    m = imgs.mean(0)
    s = imgs.std(0, ddof=1) / np.sqrt(len(films))
    worst, rows = 0.0, []
    for px, py, E, q in expect:
        for c in range(3):
            sig = s[py, px, c] + q[c] / sigmas
            z = (m[py, px, c] - E[c]) / max(sig, 1e-30)
            worst = max(worst, abs(z))
            rows.append((px, py, c, m[py, px, c], E[c], sig, z))
    return worst, rows
