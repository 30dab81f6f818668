"""Procedural inputs: sizes, winding (front faces seen from outside for the
culling test), unit normals, environment maps and camera matrices."""
import numpy as np

from firefly_amd import scenes


def _face_normals(pos):
    return np.cross(pos[:, 1] - pos[:, 0], pos[:, 2] - pos[:, 0])


def test_spheres_scene_sizes_and_winding():
    sd = scenes.spheres_scene()
    assert sd.n_tris == 2 + 8 * 32 * (2 * (16 - 2) + 2) == 7682
    assert sd.env.shape == (256, 512, 3)
    assert sd.materials.shape == (9, 16)
    fn = _face_normals(sd.positions.astype(np.float64))
    sn = sd.normals.mean(1)
    assert np.all(np.einsum("ij,ij->i", fn, sn) > 0)
    np.testing.assert_allclose(np.linalg.norm(sd.normals, axis=-1), 1, atol=1e-6)


def test_icosphere_counts_and_orientation():
    pos, nrm, uv, region = scenes.displaced_icosphere(3, (0, 0, 0), 1.0, 0.05, 0)
    assert len(pos) == 20 * 4 ** 3 and region.max() == 19
    fn = _face_normals(pos.astype(np.float64))
    assert np.mean(np.einsum("ij,ij->i", fn, pos.mean(1)) > 0) > 0.999
    r = np.linalg.norm(pos.reshape(-1, 3), axis=1)
    assert 0.94 < r.min() and r.max() < 1.06


def test_sky_and_sun_env():
    sky = scenes.sky_env(512, 256)
    assert np.isclose(sky.max(), 0.8, atol=5e-3) and np.isclose(sky.min(), 0.5, atol=5e-3)
    sun = scenes.sun_env(256, 128, seed=0, sigma_deg=2.0)
    assert sun.max() > 1000 and np.isclose(np.median(sun), 0.5)


def test_camera_center_ray_points_at_target():
    inv = scenes.inv_view_proj((0, 1.2, 3.5), (0, 0, 0), 45.0, 1.0).reshape(4, 4).T
    near = inv @ [0, 0, -1, 1]
    far = inv @ [0, 0, 1, 1]
    d = far[:3] / far[3] - near[:3] / near[3]
    d /= np.linalg.norm(d)
    t = -np.array([0, 1.2, 3.5]) / np.linalg.norm([0, 1.2, 3.5])
    np.testing.assert_allclose(d, t, atol=1e-5)
