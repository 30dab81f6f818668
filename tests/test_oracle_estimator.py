"""Pins the oracle's composed path estimator (the wf_logic + BRDF_L
composition, CRT/kernel.cu:297-455, CRT/dMaterial.cu:363-469) against
expectations computed without sampling (estimator_expectation.py): the
oracle's mean over 2^20 paths per pixel (32 independent seeds) must lie
within 4 standard errors (plus the quadrature bound) of the expectation,
for each material queue, with a wall blocking part of the sky (Q5 gating),
and for Russian roulette at path lengths > 3 through a dielectric sheet
(Q2).  Each of seven plausible composition mistakes, planted in the oracle,
must fail the same comparison by far."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import estimator_expectation as X
from firefly_amd import scenes

BATCHES, SPP = 32, 32768


def oracle_films(O, sd, batches=BATCHES, spp=SPP):
    ps = O.PreparedScene(sd)
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda b: ps.render(spp=spp, seed=1000 + b, n_threads=1)[0], range(batches)))


@pytest.fixture(scope="module")
def expectations(oracle):
    cache = {}

    def get(kind, *args):
        if (kind, args) not in cache:
            fn = X.direct_expectations if kind == "direct" else X.sheet_expectations
            cache[(kind, args)] = fn(oracle, scenes, *args)
        return cache[(kind, args)]
    return get


def _report(rows, worst):
    lines = [f"px {px} py {py} ch {c}: mean {m:.6f} expected {e:.6f} sigma {s:.2e} z {z:+.2f}"
             for px, py, c, m, e, s, z in rows]
    return f"worst |z| {worst:.2f}\n" + "\n".join(lines)


@pytest.mark.parametrize("mtype", [0, 1, 2], ids=["mix", "diffuse", "conductor"])
def test_direct_lighting_matches_expectation(oracle, expectations, mtype):
    sd, ex = expectations("direct", mtype, False)
    worst, rows = X.compare(oracle_films(oracle, sd), SPP, ex)
    assert worst < 4.0, _report(rows, worst)


def test_partly_blocked_sky_matches_expectation(oracle, expectations):
    """A wall blocks the sky on one side: light samples toward it are
    occluded, and their occlusion drops the BRDF-sampled terms of the same
    vertex too (the light ray's visibility gates the whole NEE sum, Q5)."""
    sd, ex = expectations("direct", 0, True)
    worst, rows = X.compare(oracle_films(oracle, sd), SPP, ex)
    assert worst < 4.0, _report(rows, worst)


def test_russian_roulette_matches_expectation(oracle, expectations):
    """Paths through a dielectric sheet of IOR 4 keep beta = 1 and are
    checked by Russian roulette at every length > 3 (q = 0.05, no
    compensation)."""
    sd, ex = expectations("sheet", 4.0, 12, 0.7)
    worst, rows = X.compare(oracle_films(oracle, sd), SPP, ex)
    assert worst < 4.0, _report(rows, worst)
    F = X.fresnel_dielectric(0.7, 1.0, 4.0)
    assert F > 0.3  # internal reflections are common: Russian roulette matters


@pytest.mark.parametrize("mutation,what", [
    (1, "light-sample MIS weights swapped"),
    (2, "spec BRDF-sampled term without its cosine"),
    (3, "diffuse BRDF-sampled term divided by its pdf twice"),
    (5, "BRDF-sampled terms not gated by the light ray's visibility"),
    (6, "diffuse visibility ray offset along its own direction"),
    (7, "light-sample term without its cosine"),
    (4, "Russian roulette compensated by 1 / (1 - q)"),
])
def test_planted_mistakes_fail(oracle, expectations, mutation, what):
    sd, ex = expectations("sheet", 4.0, 12, 0.7) if mutation == 4 else expectations("direct", 0, False)
    oracle.lib().or_set_mutation(mutation)
    try:
        films = oracle_films(oracle, sd, batches=16)
    finally:
        oracle.lib().or_set_mutation(0)
    worst, rows = X.compare(films, SPP, ex)
    assert worst > 8.0, f"{what}: " + _report(rows, worst)
