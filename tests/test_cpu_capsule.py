"""Capsule (segment-segment) distance of the oracle: closed forms, an
independent brute-force minimiser, symmetry and periodic-image invariance."""
import numpy as np
import pytest

import oracle
from tests.capsule_ref import brute_d2, rod


def d2(a, b, shift=(0.0, 0.0, 0.0), R=1.0):
    return oracle.pair_d2(a, b, shift, R)


def test_perpendicular_skew_crossing():
    # x-axis rod through the origin, y-axis rod crossing above it at height h
    for h in (0.25, 0.0078125, 1e-3):
        a = rod((0, 0, 0), (1, 0, 0), 1.0)
        b = rod((0, 0, h), (0, 1, 0), 1.0)
        v, w = d2(a, b)
        assert w
        assert abs(v - np.float64(np.float32(h)) ** 2) <= 1e-15


def test_collinear_gap():
    for g in (0.5, 0.0625, 0.001953125):
        a = rod((0, 0, 0), (1, 0, 0), 1.0)
        b = rod((1.0 + g, 0, 0), (1, 0, 0), 1.0)
        v, _ = d2(a, b, R=2.0)
        assert abs(v - (np.float32(1.0 + g) - 1.0) ** 2) <= 1e-15


def test_parallel_offset():
    a = rod((0, 0, 0), (1, 0, 0), 1.0)
    b = rod((0.25, 0.125, 0), (1, 0, 0), 1.0)
    v, _ = d2(a, b)
    assert abs(v - 0.125 ** 2) <= 1e-15
    # antiparallel direction, same segment
    b2 = rod((0.25, 0.125, 0), (-1, 0, 0), 1.0)
    assert abs(d2(a, b2)[0] - 0.125 ** 2) <= 1e-15


def test_endpoint_to_interior_and_endpoint_to_endpoint():
    a = rod((0, 0, 0), (1, 0, 0), 1.0)
    b = rod((0.625, 0, 0), (0, 1, 0), 1.0)  # T configuration: end (0.5,0,0) to body
    assert abs(d2(a, b)[0] - 0.125 ** 2) <= 1e-15
    # end-to-end: b leaves a's end along (1,1,0) at distance g
    g = 0.0625
    u = np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
    c = np.array([0.5 + g, 0.0, 0.0]) + 0.5 * u
    b = rod(c, u, 1.0)
    ref = brute_d2(a, b)
    assert abs(d2(a, b)[0] - ref) <= 1e-12


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_pairs_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    worst = 0.0
    for k in range(400):
        c1 = rng.uniform(0, 2, 3)
        u1 = rng.normal(size=3)
        if k % 3 == 0:  # near-parallel pairs
            u2 = u1 / np.linalg.norm(u1) + rng.normal(size=3) * 10.0 ** rng.uniform(-7, -2)
        else:
            u2 = rng.normal(size=3)
        c2 = c1 + rng.normal(size=3) * rng.uniform(0.01, 1.0)
        a = rod(c1, u1, rng.uniform(0.3, 3.0))
        b = rod(c2, u2, rng.uniform(0.3, 3.0))
        v, _ = d2(a, b, R=10.0)
        ref = brute_d2(a, b)
        assert v >= 0.0
        worst = max(worst, abs(v - ref) / max(ref, 1e-6))
        assert abs(v - ref) <= 1e-9 * max(ref, 1e-3) + 1e-13, (k, v, ref)
    assert worst < 1e-6


def test_symmetry_and_image_invariance():
    rng = np.random.default_rng(5)
    for _ in range(300):
        a = rod(rng.uniform(0, 5, 3), rng.normal(size=3), rng.uniform(0.3, 3))
        b = rod(rng.uniform(0, 5, 3), rng.normal(size=3), rng.uniform(0.3, 3))
        sh = np.array([5.0, -5.0, 0.0])
        v1, w1 = d2(a, b, sh, R=0.5)
        v2, w2 = d2(b, a, -sh, R=0.5)
        assert w1 == w2
        assert abs(v1 - v2) <= 1e-12 * max(v1, 1.0)
        # displacing rod b by the shift explicitly gives the same distance
        bb = b.copy()
        bb[:3] = (b[:3].astype(np.float64) + sh).astype(np.float32)
        if np.all(bb[:3].astype(np.float64) == b[:3].astype(np.float64) + sh):
            assert abs(d2(a, bb)[0] - v1) <= 1e-12 * max(v1, 1.0)


def test_reach_flag():
    a = rod((0, 0, 0), (1, 0, 0), 1.0)
    # reach = (l1/2 + l2/2) * 1.000001 + R = 1.010001
    for x, want in ((1.0099, True), (1.0101, False)):
        b = rod((x, 0, 0), (1, 0, 0), 1.0)
        assert d2(a, b, R=0.01)[1] is want
