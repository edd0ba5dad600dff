"""The FP32 screen and band as the GPU runs them (hardware square root and
reciprocal), fuzzed against the FP64 oracle on the adversarial pairs of
the CPU band test: near-parallel rods, distances within 1e-10..1e-3 of the
cut-off, far-from-origin coordinates with periodic shifts.

- Band: whenever it decides, its reach and cut decisions equal FP64's.
- Screen: "out of reach" only for pairs FP64 finds out of reach; "far" only
  for pairs within reach that FP64 finds not cut; "full test" only for pairs
  within reach.
"""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests.test_cpu_band import make_pairs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("contact", [False, True])
@pytest.mark.parametrize("near_parallel", [False, True])
@pytest.mark.parametrize("box", [5.0, 50.0])
@pytest.mark.parametrize("target_scale", [1.0, 40.0])
def test_device_band_and_screen_never_decide_wrong(contact, near_parallel, box, target_scale):
    rng = np.random.default_rng(100 + int(contact) * 8 + int(near_parallel) * 4 + int(box) + int(target_scale))
    cut = 0.002 if contact else 0.0
    # target_scale 1: distances at the cut; 40: far from it, near the reach
    # band (screen "out" / "far" decisions)
    target = (0.01 + cut) * target_scale
    ra, rb, sh = make_pairs(rng, 20000, box, target, near_parallel)
    reach = 3.0 + 0.01 + cut
    ew = np.float32(2.0 ** -24 * (box + reach) * 1.01)
    r = nf.selftest_pairs(ra, rb, sh, contact, cut, ew)
    _, w64, c64 = oracle.pair_batch(ra, rb, sh, contact, cut)
    dec, w32, scr = r["decided"], r["within"], r["screen"]
    decided = dec >= 0
    assert np.all(w32[decided] == w64[decided])
    assert np.all((dec[decided] == 1) == c64[decided])
    assert np.all(~w64[scr == 0])
    assert np.all(w64[scr == 1] & ~c64[scr == 1])
    assert np.all(w64[scr == 2])
    if target_scale == 1.0:
        assert decided.mean() > 0.01
    else:
        assert (scr == 1).mean() > 0.5
