"""Pins of the BSDF lobes by closed forms and integrals: Schlick endpoints,
GGX normalisation, the GGX sampler against its pdf (chi^2), the uniform
hemisphere sampler, and the smooth-dielectric Fresnel equations."""
import ctypes as C

import numpy as np
import pytest

N = np.array([0, 1, 0], np.float32)


def _v(x):
    return np.ascontiguousarray(x, np.float32)


def test_diff_f_normal_incidence(oracle):
    L = oracle.lib()
    a = _v([0.2, 0.5, 0.8])
    out = np.zeros(3, np.float32)
    L.or_diff_f(N.ctypes.data, N.ctypes.data, 0.0, a.ctypes.data, _v([0.04] * 3).ctypes.data, out.ctypes.data)
    # wh = n, F = f0 = 0.04: f = (1 - 0.04) * albedo / pi
    np.testing.assert_allclose(out, 0.96 * a / np.pi, rtol=1e-6)
    L.or_diff_f(N.ctypes.data, N.ctypes.data, 1.0, a.ctypes.data, _v([0.04] * 3).ctypes.data, out.ctypes.data)
    assert np.all(out == 0)  # metallic surfaces have no diffuse lobe


def test_spec_f_normal_incidence_rough(oracle):
    L = oracle.lib()
    out = np.zeros(3, np.float32)
    a = _v([0.3, 0.3, 0.3])
    L.or_spec_f(N.ctypes.data, N.ctypes.data, N.ctypes.data, 1.0, 0.0, a.ctypes.data,
                _v([0.04] * 3).ctypes.data, out.ctypes.data)
    # r = 1: D = 1/pi, k = 1/2 so G = 1, F = f0: f = D*G*F / 4
    np.testing.assert_allclose(out, 0.04 / (4 * np.pi), rtol=1e-6)


@pytest.mark.parametrize("r", [0.3, 0.6, 1.0])
def test_ggx_ndf_normalised(oracle, r):
    L = oracle.lib()
    th = (np.arange(20000) + 0.5) / 20000 * (np.pi / 2)
    d = np.empty_like(th)
    for k, t in enumerate(th):
        h = _v([np.sin(t), np.cos(t), 0])
        d[k] = L.or_ggx_ndf(N.ctypes.data, h.ctypes.data, r)
    integral = 2 * np.pi * np.sum(d * np.cos(th) * np.sin(th)) * (np.pi / 2 / 20000)
    assert abs(integral - 1) < 2e-3, integral


def test_spec_sampler_matches_pdf_chi2(oracle):
    """GGX half-vector sampler (CRT/dMaterial.cu:229-254) draws wi with the
    density spec_get_pdf (CRT/dMaterial.cu:255-265) wherever that pdf is
    positive; half vectors facing away from wo reflect to directions whose
    recomputed half vector is -h, where the pdf evaluates to 0.  Binned chi^2
    of the positive-pdf samples, and the zero-pdf mass equals 1 - integral."""
    from scipy.stats import chi2
    L = oracle.lib()
    r = 0.7
    wo = _v([0.6, 0.8, 0.0])
    rng = np.random.default_rng(1)
    n = 60000
    e = rng.random((n, 2), dtype=np.float32)
    out = np.zeros(3, np.float32)
    ws = np.empty((n, 3), np.float32)
    zero = np.zeros(n, bool)
    for k, (a, b) in enumerate(e):
        L.or_spec_sample(N.ctypes.data, wo.ctypes.data, r, float(a), float(b), out.ctypes.data)
        ws[k] = out
        zero[k] = L.or_spec_pdf(N.ctypes.data, out.ctypes.data, wo.ctypes.data, r) == 0.0
    ws = ws[~zero]
    # bins in (y = cos theta from the normal, phi), expected = n * integral of pdf
    nb_y, nb_p = 8, 8
    yb = np.minimum(((ws[:, 1] + 1) / 2 * nb_y).astype(int), nb_y - 1)
    pb = np.minimum(((np.arctan2(ws[:, 2], ws[:, 0]) + np.pi) / (2 * np.pi) * nb_p).astype(int), nb_p - 1)
    counts = np.zeros((nb_y, nb_p))
    np.add.at(counts, (yb, pb), 1)
    m = 48
    ys = -1 + (np.arange(nb_y * m) + 0.5) * (2 / (nb_y * m))
    ps = -np.pi + (np.arange(nb_p * m) + 0.5) * (2 * np.pi / (nb_p * m))
    pdf = np.zeros((nb_y * m, nb_p * m))
    for i, y in enumerate(ys):
        s = np.sqrt(1 - y * y)
        for j, p in enumerate(ps):
            wi = _v([s * np.cos(p), y, s * np.sin(p)])
            pdf[i, j] = L.or_spec_pdf(N.ctypes.data, wi.ctypes.data, wo.ctypes.data, r)
    cell = (2 / (nb_y * m)) * (2 * np.pi / (nb_p * m))
    exp = (pdf * cell).reshape(nb_y, m, nb_p, m).sum((1, 3)) * n
    keep = exp > 10
    stat = ((counts[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert abs(zero.mean() - (1 - exp.sum() / n)) < 0.004, (zero.mean(), exp.sum() / n)
    assert stat < chi2.ppf(0.999, keep.sum() - 1), stat


def test_diffuse_sampler_uniform_hemisphere(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    out = np.zeros(3, np.float32)
    for n in (N, _v([0.48, 0.6, 0.64])):
        cs = []
        for a, b in rng.random((20000, 2), dtype=np.float32):
            L.or_diff_sample(n.ctypes.data, float(a), float(b), out.ctypes.data)
            cs.append(float(out @ n))
        cs = np.array(cs)
        assert cs.min() >= -1e-6
        assert abs(cs.mean() - 0.5) < 0.01 and abs((cs ** 2).mean() - 1 / 3) < 0.01


def test_fresnel_dielectric(oracle):
    F = oracle.lib().or_fresnel_dielectric
    assert abs(F(1.0, 1.0, 1.5) - 0.04) < 1e-7  # ((n1-n2)/(n1+n2))^2
    assert abs(F(1.0, 1.5, 1.0) - 0.04) < 1e-7
    ct = np.sqrt(1 - 1 / 1.5 ** 2)  # critical angle from inside
    assert F(float(ct) - 1e-3, 1.5, 1.0) == 1.0  # total internal reflection
    assert F(float(ct) + 1e-3, 1.5, 1.0) < 1.0
    tb = np.arctan(1.5)  # Brewster: r_parallel = 0
    ci, ctt = np.cos(tb), np.cos(np.arcsin(np.sin(tb) / 1.5))
    rper = (ci - 1.5 * ctt) / (ci + 1.5 * ctt)
    assert abs(F(float(ci), 1.0, 1.5) - rper ** 2 / 2) < 1e-6
    cosg = np.linspace(0.01, 1, 50)
    vals = [F(float(c), 1.0, 1.5) for c in cosg]
    assert np.all(np.diff(vals) <= 1e-7) and 0 < min(vals) and max(vals) <= 1
