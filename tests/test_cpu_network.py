"""Oracle network extraction pinned against independent computations.

* series and parallel chains of rods between the electrodes: the Kirchhoff
  matrix, solved with scipy, gives the textbook series/parallel conductance;
* random packings: components against scipy's connected_components, labels
  = smallest id, electrode couplings and conductances recomputed with numpy
  (np.exp), Laplacian structure (symmetry, zero row sums, sorted columns).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components
from scipy.sparse.linalg import spsolve

import nanofiller_amd as nf
import oracle
from tests import cases


def _params(box, gap, lam, axis=0, g0=1.0):
    p = nf.normalize(nf.default_params(box=box, n_rods=1, len=0.9, radius=0.01, contact_gap=gap,
                                       elec_axis=axis, tunnel_lambda=lam, tunnel_g0=g0))
    return p


def _rod(c, d, length, a):
    d = np.asarray(d, float)
    d = d / np.linalg.norm(d)
    return np.array([*c, *d, length, a], np.float32)


def _network(p, rods):
    o = oracle.contacts(p, rods, p.contact_gap)
    return oracle.network(p, rods, o["I"], o["J"], o["gap64"], o["image"])


def _effective_conductance(n, net):
    """Source at potential 1, sink at 0: current into the source."""
    m = sp.csr_matrix((net["val"], net["col"], net["row_ptr"].astype(np.int64)), shape=(n + 2, n + 2))
    inner = np.arange(n)
    a_ii = m[inner][:, inner]
    rhs = -m[inner][:, [n]].toarray().ravel()
    v = spsolve(a_ii.tocsc(), rhs) if n > 1 else np.array([rhs[0] / a_ii.toarray()[0, 0]])
    vfull = np.concatenate([v, [1.0, 0.0]])
    return (m[[n]] @ vfull)[0]


def _g(gap, lam, g0=1.0):
    return g0 * np.exp(-max(gap, 0.0) / lam)


def _f(x):
    """A float32 rod field as the double the oracle sees."""
    return float(np.float32(x))


def _end_gaps(centres, lengths, a, L):
    """Gaps along one axis-aligned line of rods: electrode, rod-rod..., electrode."""
    c = [_f(x) for x in centres]
    h = [_f(x) * 0.5 for x in lengths]
    a = _f(a)
    out = [(c[0] - h[0]) - a]
    for k in range(len(c) - 1):
        out.append((c[k + 1] - h[k + 1]) - (c[k] + h[k]) - 2 * a)
    out.append((L - (c[-1] + h[-1])) - a)
    return out


def test_series_chain():
    """Three rods end to end along x between the electrodes: 1/G = sum 1/g."""
    a, lam, gap = 0.01, 0.02, 0.1
    p = _params((3.0, 1.0, 1.0), gap, lam)
    rods = np.stack([_rod((0.5 + k, 0.5, 0.5), (1, 0, 0), 0.9, a) for k in range(3)])
    info, net = _network(p, rods)
    # rod gaps: electrode 0.05 - a, rod-rod 0.1 - 2a; the wrap-around pair
    # (rod 2 -> image of rod 0) crosses the electrode axis and is excluded
    gs = _end_gaps([0.5, 1.5, 2.5], [0.9] * 3, a, 3.0)
    np.testing.assert_allclose(gs, [0.05 - a, 0.1 - 2 * a, 0.1 - 2 * a, 0.05 - a], atol=1e-7)
    expect = 1.0 / sum(1.0 / _g(x, lam) for x in gs)
    assert info["n_edges"] == 2 and info["n_components"] == 1 and info["percolating"] == 1
    assert info["n_source"] == 1 and info["n_sink"] == 1
    assert _effective_conductance(3, net) == pytest.approx(expect, rel=1e-9)


def test_parallel_chains():
    """Two independent chains: conductances add; components stay separate."""
    a, lam, gap = 0.01, 0.02, 0.1
    p = _params((3.0, 1.0, 1.0), gap, lam)
    r1 = [_rod((0.5 + k, 0.25, 0.5), (1, 0, 0), 0.9, a) for k in range(3)]
    # second chain: two longer rods with a wider gap
    r2 = [_rod((0.76, 0.75, 0.5), (1, 0, 0), 1.4, a), _rod((2.24, 0.75, 0.5), (1, 0, 0), 1.4, a)]
    rods = np.stack(r1 + r2)
    info, net = _network(p, rods)
    g1 = 1.0 / sum(1.0 / _g(x, lam) for x in _end_gaps([0.5, 1.5, 2.5], [0.9] * 3, a, 3.0))
    g2 = 1.0 / sum(1.0 / _g(x, lam) for x in _end_gaps([0.76, 2.24], [1.4] * 2, a, 3.0))
    assert info["n_components"] == 2 and info["percolating"] == 2
    np.testing.assert_array_equal(net["labels"], [0, 0, 0, 3, 3])
    assert _effective_conductance(5, net) == pytest.approx(g1 + g2, rel=1e-9)


def test_electrode_axis_and_g0():
    """Chain along y with elec_axis 1 and g0 = 2.5: G scales with g0."""
    a, lam, gap = 0.01, 0.05, 0.1
    p = _params((1.0, 2.0, 1.0), gap, lam, axis=1, g0=2.5)
    rods = np.stack([_rod((0.5, 0.5 + k, 0.5), (0, 1, 0), 0.9, a) for k in range(2)])
    info, net = _network(p, rods)
    gs = _end_gaps([0.5, 1.5], [0.9] * 2, a, 2.0)
    expect = 1.0 / sum(1.0 / _g(x, lam, 2.5) for x in gs)
    assert info["percolating"] == 1
    assert _effective_conductance(2, net) == pytest.approx(expect, rel=1e-9)


@pytest.mark.parametrize("axis", [0, 2])
def test_random_packing_structure(axis):
    p = cases.small_iso(nf.NF_PREC_FP64, n_rods=2000, contact_gap=0.05, elec_axis=axis,
                        tunnel_lambda=0.01)
    rods, _, _ = oracle.generate(p, threads=4)
    o = oracle.contacts(p, rods, p.contact_gap, threads=4)
    info, net = oracle.network(p, rods, o["I"], o["J"], o["gap64"], o["image"])
    n = len(rods)
    img = o["image"]
    comp = ((img >> (8 * axis)) & 0xFF).astype(np.uint8).view(np.int8)
    keep = comp == 0
    I, J, g64 = o["I"][keep], o["J"][keep], o["gap64"][keep]
    assert info["n_edges"] == keep.sum()
    # components
    adj = sp.coo_matrix((np.ones(len(I)), (I, J)), shape=(n, n))
    ncomp, lab = connected_components(adj, directed=False)
    assert info["n_components"] == ncomp
    mins = np.full(ncomp, n)
    np.minimum.at(mins, lab, np.arange(n))
    np.testing.assert_array_equal(net["labels"], mins[lab])
    assert info["largest"] == np.bincount(lab).max()
    # electrode couplings recomputed in numpy
    L = p.box[axis]
    ext = np.abs(rods[:, 6].astype(np.float64) * 0.5 * rods[:, 3 + axis].astype(np.float64))
    pos = rods[:, axis].astype(np.float64)
    gsrc = (pos - ext) - rods[:, 7].astype(np.float64)
    gsnk = (L - (pos + ext)) - rods[:, 7].astype(np.float64)
    src = np.nonzero(gsrc <= p.contact_gap)[0]
    snk = np.nonzero(gsnk <= p.contact_gap)[0]
    assert info["n_source"] == len(src) and info["n_sink"] == len(snk)
    perc = len(set(lab[src]) & set(lab[snk]))
    assert info["percolating"] == perc
    # matrix
    rp = net["row_ptr"].astype(np.int64)
    m = sp.csr_matrix((net["val"], net["col"], rp), shape=(n + 2, n + 2))
    assert info["nnz"] == m.nnz == rp[-1]
    for r in range(n + 2):
        cols = net["col"][rp[r]:rp[r + 1]]
        assert np.all(np.diff(cols.astype(np.int64)) > 0)
    dense_off = m - sp.diags(m.diagonal())
    assert abs(dense_off - dense_off.T).max() <= 1e-15 * abs(dense_off).max()
    rs = np.asarray(m.sum(axis=1)).ravel()
    assert np.all(np.abs(rs) <= 1e-13 * m.diagonal())
    deg = np.diff(rp) - 1  # off-diagonal entries per row
    np.testing.assert_array_equal(m.diagonal() > 0, deg > 0)
    assert dense_off.max() <= 0.0
    # conductances: g0 exp(-gap / lambda) against numpy's exp, merged over images
    want = sp.coo_matrix((np.exp(-np.maximum(g64, 0) / p.tunnel_lambda), (I, J)), shape=(n + 2, n + 2))
    want = want + want.T
    e_src = sp.coo_matrix((np.exp(-np.maximum(gsrc[src], 0) / p.tunnel_lambda), (src, np.full(len(src), n))),
                          shape=(n + 2, n + 2))
    e_snk = sp.coo_matrix((np.exp(-np.maximum(gsnk[snk], 0) / p.tunnel_lambda),
                           (snk, np.full(len(snk), n + 1))), shape=(n + 2, n + 2))
    want = (want + e_src + e_src.T + e_snk + e_snk.T).tocsr()
    diff = (-dense_off - want)
    assert abs(diff).max() <= 1e-14 * abs(want).max()


def test_empty_contacts():
    p = _params((2.0, 2.0, 2.0), 0.01, 0.01)
    rods = np.stack([_rod((1.0, 1.0, 1.0), (0, 0, 1), 0.5, 0.01)])
    info, net = _network(p, rods)
    assert info == {"n_nodes": 3, "n_components": 1, "largest": 1, "percolating": 0, "nnz": 3,
                    "n_edges": 0, "n_source": 0, "n_sink": 0}
    np.testing.assert_array_equal(net["row_ptr"], [0, 1, 2, 3])
    np.testing.assert_array_equal(net["val"], [0.0, 0.0, 0.0])
