"""Pins of the environment light: bilinear closed forms, the reference
upper_bound quirk table, the FP64 CDF of a constant map in closed form, and a
chi^2 test of texel sampling against the distribution the reference's
search implies (row picked with the next row's probability, Q8)."""
import json
import os

import numpy as np
import pytest

from firefly_amd import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "upper_bound_cases.json")


def test_upper_bound_golden(oracle):
    for c in json.load(open(GOLDEN))["cases"]:
        lst = [float(x) for x in c["list"]]
        assert oracle.upper_bound(lst, c["val"]) == c["expect"], c


def test_bilinear_closed_forms(oracle):
    w, h = 8, 4
    tex = np.zeros((h, w, 1), np.float32)
    tex[..., 0] = np.arange(w, dtype=np.float32)[None, :]
    # texel centre returns the texel
    for i in range(w):
        assert oracle.bilinear(tex, (i + 0.5) / w, 0.5 / h)[0] == i
    # interior corner between texels i-1 and i averages them
    assert oracle.bilinear(tex, 3 / w, 1 / h)[0] == 2.5
    # wrap: corner at u = 0 averages the last and the first column
    assert oracle.bilinear(tex, 0.0, 1 / h)[0] == 3.5
    assert oracle.bilinear(tex, 1.0, 1 / h)[0] == 3.5
    const = np.full((h, w, 3), 0.7, np.float32)
    np.testing.assert_allclose(oracle.bilinear(const, 0.123, 0.77), 0.7, rtol=1e-6)


@pytest.mark.parametrize("w,h", [(64, 32), (16, 8)])
def test_env_cdf_constant_map_closed_form(oracle, w, h):
    c = 0.8
    my, mp, conds, pdf, denom = oracle.env_build(np.full((h, w, 3), c, np.float32))
    # The corner sample's bilinear weights are 0.5 only to within rounding, so
    # each texel's luminance equals that of c to within ~1 ulp (rtol 1e-6).
    lum = np.float64(np.float32(0.299 * c + 0.587 * c + 0.114 * c))
    s = np.sin(np.float32(np.pi) * (np.arange(h, dtype=np.float32) / np.float32(h))).astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(denom, lum * w * s.sum(), rtol=1e-6)
    np.testing.assert_allclose(mp, s / s.sum(), rtol=1e-6, atol=1e-300)
    np.testing.assert_allclose(my, np.cumsum(s) / s.sum(), rtol=1e-6)
    assert my[0] == 0.0 and abs(my[-1] - 1) < 1e-12
    assert np.all(np.isnan(conds[0]))  # row 0 weighs sin(0) = 0: 0/0
    np.testing.assert_allclose(conds[1:], np.tile((np.arange(w) + 1) / w, (h - 1, 1)), rtol=1e-6)
    np.testing.assert_allclose(pdf, (s / (w * s.sum()) * w * h / (2 * np.pi ** 2))[:, None] * np.ones(w), rtol=2e-6)


def test_env_cdf_rows_monotone(oracle):
    env = scenes.sun_env(128, 64, seed=3, sigma_deg=4.0, peak=50.0)
    my, mp, conds, pdf, denom = oracle.env_build(env)
    assert np.all(np.diff(my) >= 0) and abs(my[-1] - 1) < 1e-12
    c = conds[1:]
    assert np.all(np.diff(c, axis=1) >= 0)
    np.testing.assert_allclose(c[:, -1], 1.0, rtol=1e-12)


def test_env_sampling_chi2(oracle):
    """Texel frequencies match P(x=i, y=k) = marginal_p[k+1] * (conds[k][i] -
    conds[k][i-1]), the law implied by upper_bound(marginal_y) - 1."""
    from scipy.stats import chi2
    w, h = 16, 8
    rng = np.random.default_rng(5)
    env = rng.uniform(0.2, 3.0, (h, w, 3)).astype(np.float32)
    my, mp, conds, pdf, _ = oracle.env_build(env)
    N = 200000
    e = rng.random((N, 2), dtype=np.float32)
    import ctypes as C
    L = oracle.lib()
    myp, cp = np.ascontiguousarray(my), np.ascontiguousarray(conds)
    x, y, d = C.c_int32(), C.c_int32(), np.zeros(3, np.float32)
    counts = np.zeros((h, w))
    for ex, ey in e:
        L.or_env_sample_dir(myp.ctypes.data, cp.ctypes.data, w, h, float(ex), float(ey), C.byref(x), C.byref(y),
                            d.ctypes.data)
        counts[y.value, x.value] += 1
    row_p = np.append(mp[1:], 0.0)
    col_p = np.diff(np.concatenate([np.zeros((h, 1)), np.nan_to_num(conds)], 1), axis=1)
    nan_rows = np.isnan(conds).all(1)  # row 0 (weight sin(0) = 0): the search returns column 0
    col_p[nan_rows] = 0.0
    col_p[nan_rows, 0] = 1.0
    exp = N * row_p[:, None] * col_p
    keep = exp > 5
    stat = ((counts[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert counts[~keep].sum() < 0.001 * N
    assert stat < chi2.ppf(0.999, keep.sum() - 1), stat


def test_env_pdf_lookup(oracle):
    """get_pdf reads texel ((int)(u(w-1)), (int)(v(h-1))) and divides by
    sin(acos(wi.z)) (CRT/dEnvironmentLight.cu:99-102)."""
    w, h = 32, 16
    pdf = np.arange(w * h, dtype=np.float32).reshape(h, w) + 1
    import ctypes as C
    L = oracle.lib()
    for wi in ([0.0, 1.0, 0.0], [0.6, 0.0, 0.8], [-0.36, 0.48, 0.8]):
        wi = np.array(wi, np.float32)
        got = L.or_env_pdf(pdf.ctypes.data, w, h, wi.ctypes.data)
        u = np.arctan2(wi[2], wi[0]) / (2 * np.pi) + 0.5
        v = np.arcsin(wi[1]) / np.pi + 0.5
        ref = pdf[int(np.float32(v) * (h - 1)), int(np.float32(u) * (w - 1))] / np.sin(np.arccos(wi[2]))
        np.testing.assert_allclose(got, ref, rtol=1e-5)


def test_env_lookup_direction_beyond_unit_y(oracle):
    """A normalised float direction can have |y| one ulp above 1 (seen on a
    GGX sample in a batched-request scene); the spherical map clamps the asin
    argument, so the lookup reads the pole row instead of producing NaN."""
    w, h = 32, 16
    rgb = (np.arange(w * h * 3, dtype=np.float32).reshape(h, w, 3) + 1) / (w * h * 3)
    L = oracle.lib()
    for y in (np.nextafter(np.float32(-1.0), np.float32(-2.0)), np.nextafter(np.float32(1.0), np.float32(2.0))):
        wi = np.array([-1.0848e-05, y, 2.73287e-05], np.float32)
        got = np.zeros(3, np.float32)
        L.or_env_L(rgb.ctypes.data, w, h, wi.ctypes.data, got.ctypes.data)
        ref = np.zeros(3, np.float32)
        wc = wi.copy()
        wc[1] = np.float32(np.sign(y))
        L.or_env_L(rgb.ctypes.data, w, h, wc.ctypes.data, ref.ctypes.data)
        assert np.isfinite(got).all()
        np.testing.assert_array_equal(got, ref)


def test_batched_request_scene_renders_finite(oracle):
    """The request whose GGX sample produced |y| > 1 renders without NaN."""
    from oracle import oracle as O
    spec = scenes.batch_request_specs(64, seed=0)[19]
    spec = dict(spec, width=64, height=64)
    sd = scenes.batch_request_scene(spec, scenes.sun_env(256, 128, 3))
    film, _ = O.PreparedScene(sd).render()
    assert np.isfinite(film).all()
