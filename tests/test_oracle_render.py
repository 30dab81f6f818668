"""Pins of the oracle's full path logic by closed forms and invariants:
escaped camera rays return the environment radiance exactly, results are
independent of thread count, and film layering only regroups the sums."""
import numpy as np

from firefly_amd import scenes


def _far_scene(w=24, h=16):
    sd = scenes.spheres_scene(width=w, height=h)
    # one small triangle behind the camera: every camera ray escapes
    sd.positions = np.array([[[0, 0, 50], [1, 0, 50], [0, 1, 50]]], np.float32)
    sd.normals = np.tile(np.array([0, 0, 1], np.float32), (1, 3, 1))
    sd.uvs = np.zeros((1, 3, 2), np.float32)
    sd.material_ids = np.zeros(1, np.int32)
    sd.materials = sd.materials[:1]
    sd.env = np.full((8, 16, 3), 0.625, np.float32)
    return sd


def test_escaped_primary_rays_return_environment(oracle):
    sd = _far_scene()
    film, st = oracle.PreparedScene(sd).render(spp=3, max_depth=4, n_threads=2)
    assert np.all(film[..., 3] == 3)
    np.testing.assert_allclose(film[..., :3] / 3, 0.625, rtol=1e-6)
    assert st.rays_extension == 0 and st.rays_shadow == 0 and st.rays_primary == 24 * 16 * 3


def test_render_independent_of_threads(oracle):
    sd = scenes.spheres_scene(width=40, height=24)
    ps = oracle.PreparedScene(sd)
    a, _ = ps.render(spp=4, n_threads=1)
    b, _ = ps.render(spp=4, n_threads=7)
    assert np.array_equal(a, b)


def test_film_layers_only_regroup_sums(oracle):
    sd = scenes.spheres_scene(width=32, height=20)
    ps = oracle.PreparedScene(sd)
    a, _ = ps.render(spp=8, spp_batch=1, n_threads=4)
    b, _ = ps.render(spp=8, spp_batch=4, n_threads=4)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


def test_depth_one_is_direct_lighting_only(oracle):
    """max_depth 1: the path ends at its second logic step, so a pixel holds
    the escaped-ray radiance or the first vertex's NEE (never deeper light)."""
    sd = scenes.spheres_scene(width=32, height=20)
    ps = oracle.PreparedScene(sd)
    f1, s1 = ps.render(spp=4, max_depth=1, n_threads=4)
    f3, s3 = ps.render(spp=4, max_depth=3, n_threads=4)
    assert s1.rays_extension <= s1.rays_primary
    assert s3.rays_extension > s1.rays_extension
    assert np.isfinite(f1).all() and np.isfinite(f3).all()
