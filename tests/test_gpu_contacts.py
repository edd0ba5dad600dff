"""GPU contact detection and histograms against the oracle."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


def _packing(p):
    rods, _, _ = oracle.generate(p, threads=16)
    return rods


@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
@pytest.mark.parametrize("case", ["cfg1", "small_cfg2", "tiny_box"])
def test_contacts_parity(case, prec):
    p = getattr(cases, case)(prec)
    rods = _packing(p)
    gap = p.contact_gap
    o = oracle.contacts(p, rods, gap, threads=16)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        c = pk.contacts(gap)
        st = pk.contact_stats()
        g64 = pk.contact_gap64(c.n)
    np.testing.assert_array_equal(c.I, o["I"])
    np.testing.assert_array_equal(c.col, o["J"])
    np.testing.assert_array_equal(c.image, o["image"])
    np.testing.assert_array_equal(c.gap.view(np.uint32), o["gap"].view(np.uint32))
    np.testing.assert_array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64))
    assert st["pair_tests"] == o["pair_tests"]


def test_histograms_parity():
    p = cases.small_cfg2(nf.NF_PREC_FP32_RECHECK)
    rods = _packing(p)
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    ho = oracle.histograms(p, rods, o["I"], o["J"], bins=64, centre_axis=0, pol_axis=(0, 0, 1),
                           phi_bins=32, theta_bins=32)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        pk.contacts(p.contact_gap)
        h = pk.histograms(bins=64, centre_axis=0, pol_axis=(0, 0, 1), phi_bins=32, theta_bins=32)
    for k in ("centre", "polar", "mutual", "phitheta"):
        np.testing.assert_array_equal(h[k], ho[k])
    assert h["centre"].sum() == len(rods)
    assert h["mutual"].sum() == len(o["I"])
