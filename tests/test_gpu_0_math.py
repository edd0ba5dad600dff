"""Device FP64 building blocks must be bit-identical to the CPU.

The candidate stream, the FP64 re-check and the stored gaps use a correctly
rounded square root, IEEE division and the library's own exp/log; any
difference in the last bit would change rods, decisions or contact gaps.
"""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle

pytestmark = pytest.mark.gpu


def _inputs(n=1 << 20, seed=0):
    rng = np.random.default_rng(seed)
    x = np.concatenate([
        rng.random(n // 4),
        rng.random(n // 4) * 1e-6,
        np.exp(rng.uniform(-700, 700, n // 4)),
        1.0 + (rng.random(n // 4) - 0.5) * 1e-9,
        np.array([1.0, 2.0, 0.25, 4.0, 1e-300, 1e300, 2.0 ** -1000, 5e-324, 1.0 - 2 ** -53]),
    ])
    return x


def test_sqrt_correctly_rounded():
    x = _inputs()
    got = nf.selftest("sqrt", x)
    np.testing.assert_array_equal(got.view(np.uint64), np.sqrt(x).view(np.uint64))


def test_division_ieee():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(1 << 20) * np.exp(rng.uniform(-50, 50, 1 << 20))
    b = rng.standard_normal(1 << 20) * np.exp(rng.uniform(-50, 50, 1 << 20))
    ab = np.empty(2 * a.size)
    ab[0::2], ab[1::2] = a, b
    got = nf.selftest("div", ab)
    np.testing.assert_array_equal(got.view(np.uint64), (a / b).view(np.uint64))


def test_exp_log_match_oracle():
    rng = np.random.default_rng(2)
    xe = np.concatenate([rng.uniform(-50, 50, 1 << 18), rng.uniform(-745, 709, 1 << 16), [0.0, -0.0]])
    xl = np.concatenate([rng.random(1 << 18), np.exp(rng.uniform(-700, 700, 1 << 16)), [1.0, 5e-324]])
    ge, gl = nf.selftest("exp", xe), nf.selftest("log", xl)
    oe = np.array([oracle.lib().orc_exp(float(v)) for v in xe])
    ol = np.array([oracle.lib().orc_log(float(v)) for v in xl])
    np.testing.assert_array_equal(ge.view(np.uint64), oe.view(np.uint64))
    np.testing.assert_array_equal(gl.view(np.uint64), ol.view(np.uint64))


def test_u01_matches_oracle():
    rng = np.random.default_rng(3)
    w = rng.integers(0, 2 ** 63, 1 << 16, dtype=np.uint64)
    got = nf.selftest("u01", w.view(np.float64))
    ref = np.array([oracle.lib().orc_u01(int(v >> np.uint64(32)), int(v & np.uint64(0xFFFFFFFF))) for v in w])
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
