"""Oracle core arithmetic: Philox known answers, uniform conversion, exp/log.

The Philox vectors are the Random123 known-answer vectors for
philox4x32-10 (also reproduced by rocRAND's host implementation).
"""
import math

import numpy as np
import pytest

import oracle

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(oracle.philox(ctr, key)) == want


def test_u01_open_interval_and_resolution():
    L = oracle.lib()
    lo = L.orc_u01(0, 0)
    hi = L.orc_u01(0xFFFFFFFF, 0xFFFFFFFF)
    assert lo == 0.5 * 2.0 ** -52 and 0.0 < lo
    assert hi == 1.0 - 0.5 * 2.0 ** -52 and hi < 1.0
    # 52 random bits: the low 12 bits of the first word are dropped
    assert L.orc_u01(0xFFF, 0) == lo


def test_exp_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-1, 1, 20000), [0.0, 1.0, -1.0]])
    L = oracle.lib()
    worst = max(abs(L.orc_exp(float(x)) / math.exp(x) - 1.0) for x in xs)
    assert worst < 4e-16
    assert L.orc_exp(0.0) == 1.0
    assert L.orc_exp(-800.0) == 0.0 and math.isinf(L.orc_exp(800.0))


def test_log_accuracy():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(0.5, 2, 20000),
                         [1.0, 2.0, 5e-324, 1e308]])
    L = oracle.lib()
    for x in xs:
        ref = math.log(x)
        got = L.orc_log(float(x))
        assert abs(got - ref) <= 3e-16 * max(1.0, abs(ref)), x
    assert L.orc_log(1.0) == 0.0
    assert math.isinf(L.orc_log(0.0))
