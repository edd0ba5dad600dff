"""Edge cases of the later entry points: empty contact lists, empty state
files and a one-point density map."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import state_format

pytestmark = pytest.mark.gpu


def _sparse(n=20):
    # a few short rods in a big box: no contacts, no electrode couplings
    return nf.normalize(nf.default_params(box=(50.0, 50.0, 50.0), n_rods=n, len=0.2, radius=0.005,
                                          seed=5, contact_gap=0.002, elec_axis=0,
                                          precision=nf.NF_PREC_FP64))


def test_network_layers_without_contacts():
    p = _sparse()
    with nf.Packing(p) as pk:
        pk.generate()
        c = pk.contacts(p.contact_gap)
        info, _ = pk.network()
        ci, v = pk.conductance(potentials=True)
        ninfo, g = pk.node_network()
    assert c.n == 0 and info["n_edges"] == 0
    assert ci["unknowns"] == 0 and ci["conductance"] == 0.0 and ci["converged"] == 1
    assert np.all(np.isnan(v) | (v == 0.0) | (v == 1.0))
    assert ninfo["n_edges"] == 0 and ninfo["n_screened"] == 0
    assert len(g["edge_a"]) == 0 and len(g["seg_a"]) == ninfo["n_segments"]


def test_empty_state_file_round_trip(tmp_path):
    path = str(tmp_path / "empty.bin")
    state_format.write(path, (2.0, 2.0, 2.0), (2, 2, 2), 1.0, 0.005, {})
    p = nf.normalize(nf.default_params(box=(2.0, 2.0, 2.0), n_rods=10, len=0.5, radius=0.005))
    with nf.Packing(p) as pk:
        assert pk.load_state(path) == 0
        assert pk.rods().shape == (0, 8)
        pk.save_state(str(tmp_path / "again.bin"))
    s = state_format.read(str(tmp_path / "again.bin"))
    assert s["n_rods"] == 0 and s["slots"] == 0


def test_one_point_map_is_a_constant_density():
    """A 1x1x1 map of value c accepts a candidate iff u < c, everywhere."""
    p = nf.normalize(nf.default_params(box=(4.0, 4.0, 4.0), n_rods=1500, len=0.5, radius=0.005, seed=13,
                                       den_mesh=np.full((1, 1, 1), 0.4, np.float32),
                                       precision=nf.NF_PREC_FP32_RECHECK))
    with nf.Packing(p) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
    ro, so, _ = oracle.generate(p, threads=16)
    assert np.array_equal(rods.view(np.uint32), ro.view(np.uint32))
    assert st.proposals == so["proposals"]
    # binomial acceptance of the candidates at 0.4
    frac = st.proposals / st.candidates
    assert abs(frac - 0.4) < 5 * np.sqrt(0.24 / st.candidates)
