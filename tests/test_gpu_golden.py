"""The device path reproduces the committed golden fixtures bit for bit:
generation from the seeded Philox stream, contacts, histograms, network."""
import numpy as np
import pytest

import nanofiller_amd as nf
from tests import golden_cases as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", G.NAMES)
def test_device_matches_golden(name):
    p, meta, g = G.load(name)
    hs = meta["hist"]
    with nf.Packing(p) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
        # the cell engine counts the oracle's pairs within centre reach
        pk.set_option(nf.NF_OPT_CONTACT_ENGINE, nf.NF_ENGINE_CELLS)
        cc = pk.contacts(p.contact_gap)
        cst = pk.contact_stats()
        pk.set_option(nf.NF_OPT_CONTACT_ENGINE, nf.NF_ENGINE_FOOTPRINT)
        c = pk.contacts(p.contact_gap)
        h = pk.histograms(bins=hs["bins"], centre_axis=hs["centre_axis"],
                          pol_axis=tuple(hs["pol_axis"]), phi_bins=hs["phi_bins"],
                          theta_bins=hs["theta_bins"])
        info, net = pk.network()
    np.testing.assert_array_equal(rods.view(np.uint32), g["rods_bits"])
    assert st.n_accepted == meta["stats"]["n_accepted"]
    assert st.rounds == meta["stats"]["rounds"]
    assert st.pair_tests == meta["stats"]["pair_tests"]
    np.testing.assert_array_equal(c.I, g["I"])
    np.testing.assert_array_equal(c.col, g["J"])
    np.testing.assert_array_equal(c.image, g["image"])
    np.testing.assert_array_equal(c.gap.view(np.uint32), g["gap_bits"])
    np.testing.assert_array_equal(g64.view(np.uint64), g["gap64_bits"])
    assert cst["pair_tests"] == meta["pair_tests"]
    np.testing.assert_array_equal(cc.col, c.col)
    np.testing.assert_array_equal(cc.row_ptr, c.row_ptr)
    for k in ("centre", "polar", "mutual", "phitheta"):
        np.testing.assert_array_equal(h[k], g["h_" + k])
    assert info == meta["network"]
    for k in ("labels", "row_ptr", "col"):
        np.testing.assert_array_equal(net[k], g[k])
    np.testing.assert_array_equal(net["val"].view(np.uint64), g["val_bits"])
