"""Pins of camera-ray generation (closed-form pinhole unprojection) and of the
counter RNG (range, uniformity, independence across keys)."""
import numpy as np

from firefly_amd import scenes


def test_rng_range_and_uniformity(oracle):
    from scipy.stats import chi2
    L = oracle.lib()
    vals = []
    for pix in range(200):
        base = L.or_rng_base(7, pix, 3)
        for b in range(5):
            for d in range(20):
                vals.append(L.or_rng(base, b, d))
    v = np.array(vals)
    assert v.min() >= 0.0 and v.max() < 1.0
    counts, _ = np.histogram(v, bins=50, range=(0, 1))
    exp = len(v) / 50
    assert ((counts - exp) ** 2 / exp).sum() < chi2.ppf(0.999, 49)
    a = v.reshape(200, 100)
    c = np.corrcoef(a[:, :-1].ravel(), a[:, 1:].ravel())[0, 1]
    assert abs(c) < 0.02


def test_rng_keys_distinct(oracle):
    L = oracle.lib()
    seen = set()
    for seed in range(3):
        for pix in range(50):
            for s in range(4):
                seen.add(L.or_rng_base(seed, pix, s))
    assert len(seen) == 3 * 50 * 4


def test_pinhole_camera_closed_form(oracle):
    w, h = 65, 33
    eye, center = np.array([0.3, 1.2, 3.5]), np.array([0.0, 0.2, 0.0])
    sd = scenes.spheres_scene(width=w, height=h)
    sd.inv_vp = scenes.inv_view_proj(eye, center, 45.0, w / h, znear=0.1, zfar=100.0)
    ps = oracle.PreparedScene(sd)
    fwd = (center - eye) / np.linalg.norm(center - eye)
    r = ps.gen_ray(w // 2, h // 2, 0)  # NDC (0, 0): the optical axis
    np.testing.assert_allclose(r[3:], fwd, atol=2e-6)
    np.testing.assert_allclose(r[:3], eye + 0.1 * fwd, atol=0.0052)  # near plane + <= 0.005 jitter
    # top-centre pixel: angle to the axis = atan(tan(fovy/2) * (1 - 1/h))
    r = ps.gen_ray(w // 2, 0, 0)
    ang = np.arccos(np.clip(r[3:] @ fwd, -1, 1))
    np.testing.assert_allclose(ang, np.arctan(np.tan(np.radians(22.5)) * (1 - 1 / h)), rtol=1e-5)
    # row 0 is the top of the image (NDC y = +1)
    up = np.cross(np.cross(fwd, [0, 1, 0]), fwd)
    assert r[3:] @ up > 0


def test_concentric_disk_closed_form(oracle):
    """ConcentricSampleDisk (Shirley and Chiu's concentric map,
    CRT/dRandom.cu:104-126): with o = 2u - 1, the square of half-side c goes
    onto the circle of radius c, the polar angle runs linearly along each
    edge (pi/4 * oy/ox where |ox| > |oy|, pi/2 - pi/4 * ox/oy otherwise,
    plus pi on the negative half), and equal areas go to equal areas."""
    from scipy.stats import chi2
    rng = np.random.default_rng(3)
    u = rng.random((200000, 2)).astype(np.float32)
    d = oracle.concentric(u).astype(np.float64)
    o = 2.0 * u.astype(np.float64) - 1.0
    c = np.max(np.abs(o), axis=1)
    np.testing.assert_allclose(np.hypot(d[:, 0], d[:, 1]), c, rtol=2e-6, atol=1e-7)
    xwedge = np.abs(o[:, 0]) > np.abs(o[:, 1])
    with np.errstate(divide="ignore", invalid="ignore"):
        expect = np.where(xwedge, np.pi / 4 * o[:, 1] / o[:, 0] + np.where(o[:, 0] < 0, np.pi, 0.0),
                          np.pi / 2 - np.pi / 4 * o[:, 0] / o[:, 1] + np.where(o[:, 1] < 0, np.pi, 0.0))
    diff = np.angle(np.exp(1j * (np.arctan2(d[:, 1], d[:, 0]) - expect)))
    far = c > 1e-3
    assert np.max(np.abs(diff[far])) < 2e-6
    # uniform on the disk: 8 equal-area rings x 16 sectors
    ring = np.minimum((c ** 2 * 8).astype(int), 7)
    sector = ((np.arctan2(d[:, 1], d[:, 0]) + np.pi) / (2 * np.pi) * 16).astype(int) % 16
    counts = np.bincount(ring * 16 + sector, minlength=128)
    exp = len(u) / 128
    assert ((counts - exp) ** 2 / exp).sum() < chi2.ppf(0.999, 127)
    # the centre, the edge midpoints and the corners
    pts = oracle.concentric(np.array([[0.5, 0.5], [1, 0.5], [0.5, 1], [0, 0.5], [1, 1], [0, 0]], np.float32))
    np.testing.assert_allclose(pts, [[0, 0], [1, 0], [0, 1], [-1, 0], [0.5 ** 0.5, 0.5 ** 0.5],
                                     [-0.5 ** 0.5, -0.5 ** 0.5]], atol=1e-6)


def test_thin_lens_rays_focus_on_the_focal_plane(oracle):
    """dCamera::gen_ray's lens branch (CRT/dCamera.cu:28-34): the origin
    moves by a point of the lens disk (radius lens_radius) in world x and y,
    and the ray passes through the point the pinhole ray (same jitter)
    reaches at the focal distance; lens points are uniform on the disk."""
    import dataclasses
    w, h, R, f = 33, 17, 0.2, 2.5
    sd = scenes.spheres_scene(width=w, height=h)
    pin = oracle.PreparedScene(sd)
    lens = oracle.PreparedScene(dataclasses.replace(sd, lens_radius=R, focal=f))
    offs = []
    for px, py in ((0, 0), (w // 2, h // 2), (w - 1, 3)):
        for s in range(400):
            r0 = pin.gen_ray(px, py, s).astype(np.float64)
            r1 = lens.gen_ray(px, py, s).astype(np.float64)
            off = r1[:3] - r0[:3]
            assert off[2] == 0.0
            assert np.hypot(off[0], off[1]) <= R * (1 + 1e-5)
            focus = r0[:3] + r0[3:] * f
            to_focus = focus - r1[:3]
            miss = np.linalg.norm(np.cross(to_focus, r1[3:])) / np.linalg.norm(to_focus)
            assert miss < 2e-6 and to_focus @ r1[3:] > 0
            assert abs(np.linalg.norm(r1[3:]) - 1) < 1e-6
            offs.append(off[:2] / R)
    offs = np.array(offs)
    n = len(offs)
    assert np.all(np.abs(offs.mean(0)) < 4 * 0.5 / np.sqrt(n))    # sd of a coordinate: 1/2
    r2 = (offs ** 2).sum(1)                                          # uniform disk: E r^2 = 1/2, var 1/12
    assert abs(r2.mean() - 0.5) < 4 * (1 / 12) ** 0.5 / np.sqrt(n)
