"""Pins the oracle's own exp / sin / cos / acospi (pm_oracle.c) against the
mathematics (float64 numpy), and the MWC64X stream against the published
definition (tests/golden/mwc64x_kat.json, Python integers)."""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mwc64x_kat.json")


def ulp_err(got, want):
    got = np.float64(got)
    want = np.float64(want)
    ulp = np.spacing(np.float32(abs(want))).astype(np.float64)
    return abs(got - want) / max(ulp, np.finfo(np.float32).tiny)


def test_expf_accuracy(oracle):
    xs = np.concatenate([np.linspace(-86.9, 88.0, 4001),
                         -np.abs(np.random.default_rng(0).normal(0, 3, 2000)),
                         [0.0, -0.0, 1.0, -1.0, 1e-8, -12.0]]).astype(np.float32)
    worst = max(ulp_err(oracle.expf(x), math.exp(float(x))) for x in xs)
    assert worst <= 2.0, worst


def test_expf_special(oracle):
    assert oracle.expf(0.0) == 1.0
    assert oracle.expf(-100.0) == 0.0
    assert math.isinf(oracle.expf(100.0))
    assert math.isnan(oracle.expf(float("nan")))


def test_sincos_accuracy(oracle):
    xs = np.concatenate([np.linspace(0.0, 2 * math.pi, 5001),
                         np.linspace(-20, 20, 999)]).astype(np.float32)
    for x in xs:
        s, c = oracle.sincosf(x)
        # absolute error against the float64 value of the float32 input
        assert abs(s - math.sin(float(x))) < 2.5e-7, x
        assert abs(c - math.cos(float(x))) < 2.5e-7, x
    # exact quadrant points
    s, c = oracle.sincosf(0.0)
    assert (s, c) == (0.0, 1.0)


def test_acospi_accuracy(oracle):
    xs = np.linspace(-1.0, 1.0, 20001).astype(np.float32)
    for x in xs:
        want = math.acos(float(x)) / math.pi
        assert abs(oracle.acospif(x) - want) < 2e-7, x
    assert oracle.acospif(1.0) == 0.0
    assert abs(oracle.acospif(-1.0) - 1.0) < 1e-7
    assert abs(oracle.acospif(0.0) - 0.5) < 1e-7


def test_mwc64x_matches_published_definition(oracle):
    kat = json.load(open(GOLDEN))
    for e in kat["jump"] + kat["stepped"]:
        assert oracle.mwc_state(e["k"]) == (e["x"], e["c"]), e["k"]
        assert float(oracle.draws(e["k"], 1)[0]) == e["draw"], e["k"]


def test_mwc64x_skip_equals_steps(oracle):
    # Skip(d) must equal d single steps (MWC64X.cl:78-84 vs :109-120)
    for k0 in (0, 7, 14 * 3000, 2 ** 33 + 11):
        a = oracle.draws(k0, 3000)
        b = oracle.draws(k0, 3000, stepped=True)
        assert np.array_equal(a, b)


def test_mwc64x_draw_range(oracle):
    d = oracle.draws(123456, 20000)
    assert d.min() >= 0.0 and d.max() <= 1.0
    assert abs(float(d.mean()) - 0.5) < 0.01


@pytest.mark.parametrize("k", [0, 1, 999, 2 ** 40 + 3])
def test_mwc64x_pow_against_python(oracle, k):
    A, M, BASE = 4294883355, 4294883355 * 2 ** 32 - 1, 4077358422479273989
    v = BASE * pow(A, k, M) % M
    assert oracle.mwc_state(k) == (v // A, v % A)
