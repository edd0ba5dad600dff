"""The oracle reproduces the committed golden fixtures bit for bit."""
import numpy as np
import pytest

import oracle
from tests import golden_cases as G


@pytest.mark.parametrize("name", G.NAMES)
def test_oracle_matches_golden(name):
    p, meta, g = G.load(name)
    rods, st, acc = oracle.generate(p, threads=2)
    np.testing.assert_array_equal(rods.view(np.uint32), g["rods_bits"])
    np.testing.assert_array_equal(acc, g["round_accepts"])
    assert st == meta["stats"]
    c = oracle.contacts(p, rods, p.contact_gap)
    for k in ("I", "J", "image"):
        np.testing.assert_array_equal(c[k], g[k])
    np.testing.assert_array_equal(c["gap"].view(np.uint32), g["gap_bits"])
    np.testing.assert_array_equal(c["gap64"].view(np.uint64), g["gap64_bits"])
    assert c["pair_tests"] == meta["pair_tests"]
    hs = meta["hist"]
    h = oracle.histograms(p, rods, c["I"], c["J"], bins=hs["bins"], centre_axis=hs["centre_axis"],
                          pol_axis=tuple(hs["pol_axis"]), phi_bins=hs["phi_bins"],
                          theta_bins=hs["theta_bins"])
    for k in ("centre", "polar", "mutual", "phitheta"):
        np.testing.assert_array_equal(h[k], g["h_" + k])
    info, net = oracle.network(p, rods, c["I"], c["J"], c["gap64"], c["image"])
    assert info == meta["network"]
    for k in ("labels", "row_ptr", "col"):
        np.testing.assert_array_equal(net[k], g[k])
    np.testing.assert_array_equal(net["val"].view(np.uint64), g["val_bits"])


def test_golden_fixtures_are_consistent():
    """Independent sanity of the stored data: counts, ordering, invariants."""
    for name in G.NAMES:
        p, meta, g = G.load(name)
        n = len(g["rods_bits"])
        assert n == meta["stats"]["n_accepted"] == p.n_rods
        assert g["round_accepts"].sum() == n
        I, J = g["I"].astype(np.int64), g["J"].astype(np.int64)
        assert np.all(I < J)
        key = (I * n + J) * (1 << 24) + (g["image"].astype(np.int64) & 0xFFFFFF)
        assert np.all(np.diff(I) >= 0)
        assert len(np.unique(key)) == len(key)
        assert g["h_centre"].sum() == n and g["h_polar"].sum() == n
        assert g["h_mutual"].sum() == len(I)
        assert g["row_ptr"][-1] == meta["network"]["nnz"] == len(g["col"])


def test_digest_helpers_and_full_size_fixture():
    """The digests hash exactly the stated bytes, notice a one-bit change,
    and agree with the small fixtures' arrays; the full-size fixture loads
    with its written-out schedule bound and self-consistent counters."""
    import hashlib

    p, meta, g = G.load("iso_fp64")
    rods = g["rods_bits"].view(np.float32).reshape(-1, 8)
    assert G.rods_digest(rods) == hashlib.sha256(g["rods_bits"].tobytes()).hexdigest()
    flipped = g["rods_bits"].copy()
    flipped[3] ^= 1
    assert G.rods_digest(flipped.view(np.float32).reshape(-1, 8)) != G.rods_digest(rods)
    d = G.contacts_digest(g["I"], g["J"], g["image"], g["gap_bits"].view(np.float32),
                          g["gap64_bits"].view(np.float64))
    raw = b"".join(a.tobytes() for a in (g["I"].astype(np.uint32), g["J"].astype(np.uint32),
                                         g["image"].astype(np.int32), g["gap_bits"], g["gap64_bits"]))
    assert d == hashlib.sha256(raw).hexdigest()
    for name, n in (("headline_500k", 500000), ("contacts_4M", 4000000), ("network_8M", 8000000)):
        p, meta = G.load_digest(name)
        assert p.n_rods == meta["stats"]["n_accepted"] == n
        assert p.max_batch == meta["params"]["max_batch"]
        assert sum(meta["round_accepts"]) == n and len(meta["round_accepts"]) == meta["stats"]["rounds"]
        assert 0 < meta["contacts"] and len(meta["rods_sha256"]) == len(meta["contacts_sha256"]) == 64
        if "h_centre" in meta:
            assert sum(meta["h_centre"]) == sum(meta["h_polar"]) == n
            assert sum(meta["h_mutual"]) == meta["contacts"]
        if "network" in meta:
            assert meta["network"]["n_nodes"] == n + 2 and meta["network"]["n_edges"] <= meta["contacts"]
            assert len(meta["network_sha256"]) == 64
    # the network digest hashes exactly the stated bytes
    p, meta, g = G.load("iso_fp64")
    raw = b"".join(a.tobytes() for a in (g["labels"].astype(np.uint32), g["row_ptr"].astype(np.uint32),
                                         g["col"].astype(np.uint32), g["val_bits"]))
    assert G.network_digest(g["labels"], g["row_ptr"], g["col"], g["val_bits"].view(np.float64)) == \
        hashlib.sha256(raw).hexdigest()
