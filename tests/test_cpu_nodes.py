"""Node-level segment network of the oracle, pinned by hand-built geometry:
contact points at known axial positions, segment lengths dt / a, merging
of points closer than one radius (mean position), screening by a third rod
inside / just outside the contact's virtual capsule, and electrode nodes at
the rod end facing the plane."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle

A = 0.005


def _p(box=(4.0, 4.0, 4.0), gap=0.004, axis=0, n=10):
    return nf.normalize(nf.default_params(box=box, n_rods=n, len=1.0, radius=A, contact_gap=gap,
                                          elec_axis=axis, precision=nf.NF_PREC_FP64))


def _rod(c, d, length, a=A):
    d = np.asarray(d, float)
    d = d / np.linalg.norm(d)
    return np.array([*c, *d, length, a], np.float32)


def _net(p, rods):
    o = oracle.contacts(p, rods, p.contact_gap)
    return o, oracle.node_network(p, rods, o["I"], o["J"], o["image"])


def _f(x):
    return float(np.float32(x))


def test_two_crossings_give_one_segment():
    g = 0.002
    z0 = 2.0
    zc = _f(z0 + 2 * A + g)
    rods = np.stack([_rod((1.5, 2.0, z0), (1, 0, 0), 2.0),
                     _rod((1.2, 2.0, zc), (0, 1, 0), 1.0),
                     _rod((1.9, 2.0, zc), (0, 1, 0), 1.0)])
    p = _p(axis=2)
    o, r = _net(p, rods)
    assert len(o["I"]) == 2 and list(r["flag"]) == [0, 0]
    # nodes: rod 0 at t = 1.2 - 1.5 and 1.9 - 1.5, rods 1 and 2 at their centres
    np.testing.assert_array_equal(r["node_rod"], [0, 0, 1, 2])
    np.testing.assert_allclose(r["node_t"], [_f(1.2) - _f(1.5), _f(1.9) - _f(1.5), 0.0, 0.0], atol=1e-12)
    np.testing.assert_array_equal(r["seg_a"], [0])
    np.testing.assert_array_equal(r["seg_b"], [1])
    np.testing.assert_allclose(r["seg_L"], [(_f(1.9) - _f(1.2)) / _f(A)], rtol=1e-12)
    np.testing.assert_array_equal(np.sort(np.stack([r["edge_a"], r["edge_b"]], 1), 0), [[0, 2], [1, 3]])


def test_points_within_one_radius_merge():
    g = 0.002
    zc = _f(2.0 + 2 * A + g)
    x2 = 1.2 + 0.8 * A  # second crossing 0.8 a further along rod 0
    zb = _f(2.0 - 2 * A - g)
    # one crossing above rod 0, the other below it, 0.8 a further along
    rods = np.stack([_rod((1.5, 2.0, 2.0), (1, 0, 0), 2.0),
                     _rod((1.2, 2.0, zc), (0, 1, 0), 1.0),
                     _rod((x2, 2.0, zb), (0, 1, 0), 1.0)])
    p = _p(axis=2)
    o, r = _net(p, rods)
    assert len(o["I"]) == 2 and list(r["flag"]) == [0, 0]
    on0 = r["node_rod"] == 0
    assert on0.sum() == 1
    t0 = r["node_t"][on0][0]
    assert t0 == pytest.approx(0.5 * ((_f(1.2) - _f(1.5)) + (_f(x2) - _f(1.5))), abs=1e-9)
    assert not np.any(r["node_rod"][r["seg_a"]] == 0)


@pytest.mark.parametrize("dx, screened", [(1.5 * A, True), (2.5 * A, False)])
def test_third_rod_screens_the_contact(dx, screened):
    g = 0.002
    z0 = 2.0
    zc = _f(z0 + 2 * A + g)
    # D runs along y through the gap region, dx away from the contact segment
    rods = np.stack([_rod((1.5, 2.0, z0), (1, 0, 0), 2.0),
                     _rod((1.2, 2.0, zc), (0, 1, 0), 1.0),
                     _rod((1.2 + dx, 2.0, _f(z0 + A + g / 2)), (0, 1, 0), 0.6)])
    p = _p(axis=2, gap=0.01)
    o, r = _net(p, rods)
    k = [q for q in range(len(o["I"])) if (o["I"][q], o["J"][q]) == (0, 1)]
    assert len(k) == 1
    assert r["flag"][k[0]] == (1 if screened else 0)


def test_electrode_nodes_at_rod_ends():
    p = _p(box=(2.0, 2.0, 2.0), gap=0.004, axis=0)
    # rod along +x touching the source plane, rod along -x touching the sink plane
    rods = np.stack([_rod((0.5 + A + 0.001, 1.0, 1.0), (1, 0, 0), 1.0),
                     _rod((1.5 - A - 0.001, 1.5, 1.0), (-1, 0, 0), 1.0)])
    o, r = _net(p, rods)
    np.testing.assert_array_equal(r["elec_side"], [0, 1])
    np.testing.assert_array_equal(r["node_rod"][r["elec_node"]], [0, 1])
    # the source node of rod 0 is its lower end (t = -l/2); rod 1 points along -x,
    # so its end at the sink plane is also t = -l/2
    np.testing.assert_allclose(r["node_t"][r["elec_node"]], [-0.5, -0.5])
    np.testing.assert_allclose(r["elec_gap"], [0.001, 0.001], atol=1e-6)


def test_random_packing_invariants():
    p = nf.normalize(nf.default_params(box=(3.0, 3.0, 3.0), n_rods=3000, len=0.8, radius=A, seed=7,
                                       contact_gap=0.02, elec_axis=1, precision=nf.NF_PREC_FP64))
    rods, _, _ = oracle.generate(p, threads=8)
    o, r = _net(p, rods)
    nn = len(r["node_rod"])
    # nodes sorted by rod then t, t within the rod
    assert np.all(np.diff(r["node_rod"].astype(np.int64)) >= 0)
    hl = rods[r["node_rod"], 6].astype(np.float64) * 0.5
    assert np.all(np.abs(r["node_t"]) <= hl + 1e-12)
    # segments join consecutive nodes of one rod with L >= 1 (merged below one radius)
    assert np.all(r["seg_b"] == r["seg_a"] + 1)
    assert np.all(r["node_rod"][r["seg_a"]] == r["node_rod"][r["seg_b"]])
    assert np.all(r["seg_L"] > 1.0)
    # per rod: segments = nodes - 1
    rods_with = np.unique(r["node_rod"])
    assert len(r["seg_a"]) == nn - len(rods_with)
    # every node carries at least one edge or electrode coupling
    used = np.zeros(nn, bool)
    used[r["edge_a"]] = used[r["edge_b"]] = True
    used[r["elec_node"]] = True
    assert used.all()
    # kept edges join different rods
    assert np.all(r["node_rod"][r["edge_a"]] != r["node_rod"][r["edge_b"]])
