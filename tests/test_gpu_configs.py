"""The batched, contact/statistics and network workloads at sizes the oracle
checks, bit for bit, plus size-independent properties at full size.

- 256-packing batch mix: 24 instances with the full aspect (20-500), volume
  fraction (0.1-2 %) and vMF kappa (1-20, random axes) ranges, N scaled down;
  each instance's rods, counters and status against oracle.generate.
- 4M-rod contact workload laws (uniform density, vMF(+x, 2), 2 nm gap) at
  64,000 rods in a 12.6 um cube: packing, contact list, gaps and the three
  64-bin histograms about +x.
- 8M-rod network workload laws (16 agglomerates, vMF(+x, 3), electrodes on x,
  lambda = 0.1 nm) on a 40 x 4 x 4 um slab of 80,000 rods: packing, contacts,
  network (labels, CSR, FP64 values) and the node-level network, whose
  screening sieve sees coordinates up to 40 um.
- Full-size 4M-rod packing: histogram totals, contact count = mutual-angle
  total, no overlapping pair, sampled contact gaps against the oracle's FP64
  pair evaluation.
"""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from nanofiller_amd import configs

pytestmark = pytest.mark.gpu


def _gen_compare(p, threads=16):
    ro, so, _ = oracle.generate(p, threads=threads)
    with nf.Packing(p) as pk:
        st = pk.generate()[0]
        rg = pk.rods()
    assert st.n_accepted == so["n_accepted"]
    assert st.rounds == so["rounds"] and st.pair_tests == so["pair_tests"]
    np.testing.assert_array_equal(rg.view(np.uint32), ro.view(np.uint32))
    return ro


def test_batch_mix_matches_oracle():
    ps = configs.cfg3(n_inst=24, n_scale=1.0 / 40.0, max_rounds=30)
    with nf.Packing(ps) as pk:
        sts = pk.generate()
        rods = [pk.rods(i) for i in range(len(ps))]
    statuses = set()
    for i, p in enumerate(ps):
        ro, so, _ = oracle.generate(p, threads=16)
        st = sts[i]
        print(f"inst {i}: N {p.n_rods} box {p.box[0]:.2f} len {p.len:.2f} -> {st.n_accepted} "
              f"{nf.STATUS[st.status]} in {st.rounds} rounds", flush=True)
        assert nf.STATUS[st.status] == nf.STATUS[so["status"]], i
        assert st.n_accepted == so["n_accepted"], i
        assert st.rounds == so["rounds"], i
        assert st.candidates == so["candidates"] and st.proposals == so["proposals"], i
        assert st.pair_tests == so["pair_tests"], i
        assert st.rejected_old == so["rejected_old"] and st.rejected_new == so["rejected_new"], i
        np.testing.assert_array_equal(rods[i].view(np.uint32), ro.view(np.uint32), err_msg=str(i))
        statuses.add(nf.STATUS[st.status])
    # the mix reaches its target for some instances and jams for others
    assert "target" in statuses and len(statuses) > 1


def _cfg4_small(prec=nf.NF_PREC_FP32_RECHECK):
    edge = (64000 / 32.0) ** (1.0 / 3.0)
    return configs.cfg4(prec, box=(edge, edge, edge), n_rods=64000)


@pytest.mark.parametrize("prec", [nf.NF_PREC_FP32_RECHECK, nf.NF_PREC_FP64])
def test_contact_workload_small(prec):
    p = _cfg4_small(prec)
    rods = _gen_compare(p)
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    ho = oracle.histograms(p, rods, o["I"], o["J"], bins=64, centre_axis=0, pol_axis=(1, 0, 0))
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
        h = pk.histograms(bins=64, centre_axis=0, pol_axis=(1, 0, 0))
        pk.set_option(nf.NF_OPT_CONTACT_ENGINE, nf.NF_ENGINE_CELLS)
        cc = pk.contacts(p.contact_gap)
        st = pk.contact_stats()
    np.testing.assert_array_equal(cc.col, c.col)
    np.testing.assert_array_equal(cc.row_ptr, c.row_ptr)
    np.testing.assert_array_equal(c.I, o["I"])
    np.testing.assert_array_equal(c.col, o["J"])
    np.testing.assert_array_equal(c.image, o["image"])
    np.testing.assert_array_equal(c.gap.view(np.uint32), o["gap"].view(np.uint32))
    np.testing.assert_array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64))
    assert st["pair_tests"] == o["pair_tests"]
    for k in ("centre", "polar", "mutual"):
        np.testing.assert_array_equal(h[k], ho[k], err_msg=k)
    assert h["centre"].sum() == len(rods) and h["mutual"].sum() == c.n
    # vMF(+x, 2): polar angles to +x lean towards 0
    assert h["polar"][:32].sum() > 2 * h["polar"][32:].sum()


def _cfg5_slab():
    return configs.cfg5(box=(40.0, 4.0, 4.0), n_rods=80000)


def test_network_workload_slab():
    p = _cfg5_slab()
    rods = _gen_compare(p)
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    oi, on = oracle.network(p, rods, o["I"], o["J"], o["gap64"], o["image"])
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        c = pk.contacts(p.contact_gap)
        info, net = pk.network()
        ninfo, g = pk.node_network()
    np.testing.assert_array_equal(c.I, o["I"])
    np.testing.assert_array_equal(c.col, o["J"])
    np.testing.assert_array_equal(c.image, o["image"])
    assert info == oi
    for k in ("labels", "row_ptr", "col"):
        np.testing.assert_array_equal(net[k], on[k], err_msg=k)
    np.testing.assert_array_equal(net["val"].view(np.uint64), on["val"].view(np.uint64))
    assert oi["n_source"] > 0 and oi["n_sink"] > 0
    nd = oracle.node_network(p, rods, c.I, c.col, c.image, threads=16)
    assert ninfo["n_screened"] == int(np.sum(nd["flag"] == 1))
    assert ninfo["n_points"] == nd["n_points"]
    for k in ("node_rod", "seg_a", "seg_b", "edge_a", "edge_b", "edge_contact", "elec_node", "elec_side"):
        np.testing.assert_array_equal(g[k], nd[k], err_msg=k)
    for k in ("node_t", "seg_L", "elec_gap"):
        assert np.array_equal(g[k].view(np.uint64), np.asarray(nd[k], np.float64).view(np.uint64)), k


def test_contact_workload_full_size_properties():
    """4,000,000 rods: properties that hold at any size."""
    p = configs.cfg4()
    with nf.Packing(p) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
        h = pk.histograms(bins=64, centre_axis=0, pol_axis=(1, 0, 0))
        pk.set_option(nf.NF_OPT_CONTACT_ENGINE, nf.NF_ENGINE_CELLS)
        cc = pk.contacts(p.contact_gap)
        g64c = pk.contact_gap64(cc.n)
    # the footprint and cell engines find the same contact list
    np.testing.assert_array_equal(cc.row_ptr, c.row_ptr)
    np.testing.assert_array_equal(cc.col, c.col)
    np.testing.assert_array_equal(cc.image, c.image)
    np.testing.assert_array_equal(g64c.view(np.uint64), g64.view(np.uint64))
    assert st.n_accepted == 4000000 and nf.STATUS[st.status] == "target"
    assert h["centre"].sum() == 4000000 and h["polar"].sum() == 4000000
    assert h["mutual"].sum() == c.n
    # every overlapping pair would be a contact with a negative gap
    assert c.n > 0 and float(g64.min()) >= 0.0 and float(g64.max()) <= p.contact_gap
    I = c.I
    assert np.all(c.col > I)
    # sampled contacts: the oracle's FP64 evaluation of the same pair and image
    rng = np.random.default_rng(0)
    idx = rng.choice(c.n, size=min(2000, c.n), replace=False)
    box = np.array([p.box[0], p.box[1], p.box[2]])
    for k in idx:
        im = int(c.image[k])
        sh = np.array([((im >> (8 * a)) & 0xFF) for a in range(3)], np.int64)
        sh = np.where(sh > 127, sh - 256, sh) * box
        ra, rb = rods[I[k]], rods[c.col[k]]
        d2, _ = oracle.pair_d2(ra, rb, sh, float(ra[7]) + float(rb[7]) + p.contact_gap)
        gap = np.sqrt(d2) - (float(ra[7]) + float(rb[7]))
        assert abs(gap - g64[k]) <= 1e-12, (k, gap, g64[k])
