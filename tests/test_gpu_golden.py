"""The device path reproduces the committed golden fixtures bit for bit:
generation from the seeded Philox stream, contacts, histograms, network."""
import numpy as np
import pytest

import nanofiller_amd as nf
from tests import golden_cases as G

pytestmark = pytest.mark.gpu

# generations whose pair-test count is compared test every pair within reach
EXACT = {nf.NF_OPT_TEST_ALL_PAIRS: 1}


@pytest.mark.parametrize("name", G.NAMES)
def test_device_matches_golden(name):
    p, meta, g = G.load(name)
    hs = meta["hist"]
    with nf.Packing(p, EXACT) as pk:
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


def test_headline_full_size_matches_oracle_digest():
    """The 500,000-rod headline packing and its contact list at full size,
    bit for bit against the oracle's: counters, and SHA-256 digests of the
    rods and of the contact list (I, J, image, gap32 and gap64 bits), written
    by tests/golden/make_full_size.py."""
    p, meta = G.load_digest("headline_500k")
    with nf.Packing(p, EXACT) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
    s = meta["stats"]
    print("headline:", st.n_accepted, st.rounds, st.pair_tests, c.n, flush=True)
    assert st.n_accepted == s["n_accepted"] == 500000
    assert st.rounds == s["rounds"] and st.pair_tests == s["pair_tests"]
    assert st.candidates == s["candidates"] and st.proposals == s["proposals"]
    assert G.rods_digest(rods) == meta["rods_sha256"]
    assert c.n == meta["contacts"]
    assert G.contacts_digest(c.I, c.col, c.image, c.gap, g64) == meta["contacts_sha256"]


def test_contact_workload_full_size_matches_oracle_digest():
    """The 4,000,000-rod contact workload at full size, bit for bit against
    the oracle's: counters, digests of the rods and of the contact list, and
    the three 64-bin histograms about +x (tests/golden/make_full_size.py)."""
    p, meta = G.load_digest("contacts_4M")
    hs = meta["hist"]
    with nf.Packing(p, EXACT) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
        h = pk.histograms(bins=hs["bins"], centre_axis=hs["centre_axis"], pol_axis=tuple(hs["pol_axis"]))
    s = meta["stats"]
    print("4M:", st.n_accepted, st.rounds, st.pair_tests, c.n, flush=True)
    assert st.n_accepted == s["n_accepted"] == 4000000
    assert st.rounds == s["rounds"] and st.pair_tests == s["pair_tests"]
    assert st.candidates == s["candidates"] and st.proposals == s["proposals"]
    assert G.rods_digest(rods) == meta["rods_sha256"]
    assert c.n == meta["contacts"]
    assert G.contacts_digest(c.I, c.col, c.image, c.gap, g64) == meta["contacts_sha256"]
    for k in ("centre", "polar", "mutual"):
        np.testing.assert_array_equal(h[k], np.asarray(meta["h_" + k], np.int64), err_msg=k)


def test_network_workload_full_size_matches_oracle_digest():
    """The 8,000,000-rod network workload at full size, bit for bit against
    the oracle's: counters, digests of the rods, the contact list and the
    network (labels, CSR and FP64 value bits), and the network's counts
    (tests/golden/make_full_size.py)."""
    p, meta = G.load_digest("network_8M")
    with nf.Packing(p, EXACT) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
        info, net = pk.network()
    s = meta["stats"]
    print("8M:", st.n_accepted, st.rounds, st.pair_tests, c.n, info, flush=True)
    assert st.n_accepted == s["n_accepted"] == 8000000
    assert st.rounds == s["rounds"] and st.pair_tests == s["pair_tests"]
    assert st.candidates == s["candidates"] and st.proposals == s["proposals"]
    assert G.rods_digest(rods) == meta["rods_sha256"]
    assert c.n == meta["contacts"]
    assert G.contacts_digest(c.I, c.col, c.image, c.gap, g64) == meta["contacts_sha256"]
    assert info == meta["network"]
    assert G.network_digest(net["labels"], net["row_ptr"], net["col"], net["val"]) == meta["network_sha256"]
