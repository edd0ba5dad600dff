"""Cell lists of strongly uneven packings (the counting sort that bins rods
into cells for the contact engines and the statistics).

The packing is a dense cluster (a 3,000-rod packing of a 3 um cube placed
in one corner of a 24 um box) plus a sparse background (a second packing
scaled eightfold, without the rods near the cluster), so that a few buckets
of cells hold about half the rods and most cells are empty.  The scaled
background cannot overlap itself: two rods of fixed directions overlap for
a convex set of centre offsets that contains 0, so an offset outside it
stays outside when multiplied by 8.  Contact lists (both engines) and the
centre-distance statistics on a grid of 13.8 M cells (buckets of 4096
cells, more buckets than the histogram pass aims for) equal the oracle's.
"""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu

BIG = 24.0


def _uneven():
    dense, _, _ = oracle.generate(cases.small_iso(nf.NF_PREC_FP64), threads=16)
    dense = dense.copy()
    dense[:, 0:3] += np.float32(0.5)  # centres in [0.5, 3.5), rods within [0, 4]
    sparse, _, _ = oracle.generate(cases.small_iso(nf.NF_PREC_FP64, seed=19), threads=16)
    sparse = sparse.copy()
    sparse[:, 0:3] *= np.float32(BIG / 3.0)
    c = sparse[:, 0:3].astype(np.float64)
    c = np.where(c > BIG - 2.0, c - BIG, c)  # periodic: just below the box end is near 0
    near = np.all((c > -1.2) & (c < 5.2), axis=1)
    rods = np.concatenate([dense, sparse[~near]])
    p = nf.normalize(nf.default_params(box=(BIG, BIG, BIG), n_rods=len(rods), len=1.0, radius=0.005, seed=7,
                                       contact_gap=0.002, precision=nf.NF_PREC_FP64))
    return p, np.ascontiguousarray(rods)


@pytest.mark.parametrize("engine", [nf.NF_ENGINE_FOOTPRINT, nf.NF_ENGINE_CELLS])
@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
def test_contacts_of_an_uneven_packing(engine, prec):
    p, rods = _uneven()
    p.precision = prec
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    with nf.Packing(p, {nf.NF_OPT_CONTACT_ENGINE: engine}) as pk:
        pk.set_rods(rods)
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
        st = pk.contact_stats()
    print(len(rods), "rods", len(o["I"]), "contacts", flush=True)
    assert len(o["I"]) > 100
    np.testing.assert_array_equal(c.I, o["I"])
    np.testing.assert_array_equal(c.col, o["J"])
    np.testing.assert_array_equal(c.image, o["image"])
    np.testing.assert_array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64))
    if engine == nf.NF_ENGINE_CELLS:
        assert st["pair_tests"] == o["pair_tests"]


def test_centre_distance_on_a_fine_grid():
    p, rods = _uneven()
    bins, width = 10, 0.01
    o = oracle.centre_distance(p, rods, bins, width)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        g = pk.centre_distance(bins, width)
    np.testing.assert_array_equal(g["counts"], o["counts"])
    assert (g["pairs"], g["sum_d_q24"], g["sum_d2_q16"]) == (o["pairs"], o["sum_d_q24"], o["sum_d2_q16"])
    assert g["pairs"] > 0
