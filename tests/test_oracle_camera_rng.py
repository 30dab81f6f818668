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
