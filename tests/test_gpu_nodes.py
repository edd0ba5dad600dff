"""Node-level segment network on the GPU against the oracle on the same
contact list: nodes, merged positions, segments, kept edges and electrode
nodes bit for bit; screening counts equal."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu

CASES = {
    # a wide contact gap: third rods often sit between contacting pairs
    "iso_wide_gap_x": lambda: cases.small_iso(nf.NF_PREC_FP64, contact_gap=0.05, tunnel_lambda=0.01),
    "tiny_box_images_z": lambda: cases.tiny_box(nf.NF_PREC_FP64, contact_gap=0.2, elec_axis=2),
    "dense_long_rods_y": lambda: nf.normalize(nf.default_params(box=(3.0, 3.0, 3.0), n_rods=6000,
                                                                len=1.2, radius=0.005, seed=9,
                                                                contact_gap=0.01, elec_axis=1)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_node_network_matches_oracle(name, tmp_path):
    p = CASES[name]()
    with nf.Packing(p) as pk:
        pk.generate()
        rods = pk.rods()
        c = pk.contacts(p.contact_gap)
        info, g = pk.node_network()
        pk.write_node_coo(str(tmp_path) + "/")
    o = oracle.node_network(p, rods, c.I, c.col, c.image, threads=16)
    print(name, info, flush=True)
    assert info["n_screened"] == int(np.sum(o["flag"] == 1))
    assert info["n_points"] == o["n_points"]
    for k in ("node_rod", "seg_a", "seg_b", "edge_a", "edge_b", "edge_contact", "elec_node", "elec_side"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    for k in ("node_t", "seg_L", "elec_gap"):
        assert np.array_equal(g[k].view(np.uint64), np.asarray(o[k], np.float64).view(np.uint64)), k
    np.testing.assert_array_equal(g["seg_A"], rods[g["node_rod"][g["seg_a"]], 7].astype(np.float64))
    lines = (tmp_path / "segment.LA.coo").read_text().splitlines()
    assert lines[0] == "#\tI\tJ\tData1\tData2" and len(lines) == 1 + info["n_segments"]
    assert len((tmp_path / "Nodes.tbl.dat").read_text().splitlines()) == 1 + info["n_nodes"]
    assert len((tmp_path / "internal.DS.coo").read_text().splitlines()) == 1 + info["n_edges"]


def test_node_network_needs_contacts():
    p = cases.small_iso(nf.NF_PREC_FP64)
    with nf.Packing(p) as pk:
        pk.generate()
        with pytest.raises(nf.NFError) as e:
            pk.node_network()
    assert e.value.code == nf.NF_E_STATE
