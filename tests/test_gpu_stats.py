"""GPU centre-distance statistics and van der Waals energy against the oracle, bit for bit."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case,bins,width", [("small_iso", 16, 0.05), ("tiny_box", 30, 0.1),
                                             ("small_cfg2", 12, 0.1)])
def test_centre_distance_parity(case, bins, width):
    kw = {"n_rods": 3000} if case != "tiny_box" else {}
    p = getattr(cases, case)(nf.NF_PREC_FP64, **kw)
    rods, _, _ = oracle.generate(p, threads=16)
    o = oracle.centre_distance(p, rods, bins, width)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        g = pk.centre_distance(bins, width)
    np.testing.assert_array_equal(g["counts"], o["counts"])
    assert (g["pairs"], g["sum_d_q24"], g["sum_d2_q16"]) == (o["pairs"], o["sum_d_q24"],
                                                             o["sum_d2_q16"])
    assert g["pairs"] > 0


@pytest.mark.parametrize("case,gap", [("small_iso", 0.01), ("tiny_box", 0.05), ("cfg1", 0.002)])
def test_vdw_parity(case, gap):
    p = getattr(cases, case)(nf.NF_PREC_FP32_RECHECK, contact_gap=gap)
    rods, _, _ = oracle.generate(p, threads=16)
    c = oracle.contacts(p, rods, gap, threads=16)
    eo, no = oracle.vdw(p, rods, c["I"], c["J"], c["image"], gap)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        pk.contacts(gap)
        eg, ng = pk.vdw()
    assert ng == no and ng > 0
    assert np.float64(eg).view(np.uint64) == np.float64(eo).view(np.uint64)
