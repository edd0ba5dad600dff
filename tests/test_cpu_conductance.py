"""The oracle's conductance solve pinned against closed forms: series and
parallel chains of rods (textbook 1/G = sum 1/g, G = sum G_k), current
conservation, and the potentials of components touching one electrode."""
import numpy as np
import pytest

import oracle
from tests.test_cpu_network import _end_gaps, _g, _network, _params, _rod


def test_series_chain_closed_form():
    a, lam, gap = 0.01, 0.02, 0.1
    p = _params((3.0, 1.0, 1.0), gap, lam)
    rods = np.stack([_rod((0.5 + k, 0.5, 0.5), (1, 0, 0), 0.9, a) for k in range(3)])
    _, net = _network(p, rods)
    gs = _end_gaps([0.5, 1.5, 2.5], [0.9] * 3, a, 3.0)
    expect = 1.0 / sum(1.0 / _g(x, lam) for x in gs)
    r = oracle.conductance(net)
    assert r["unknowns"] == 3
    assert r["conductance"] == pytest.approx(expect, rel=1e-12)
    assert r["current_sink"] == pytest.approx(expect, rel=1e-12)
    # potentials drop along the chain in proportion to the series resistances
    res = [1.0 / _g(x, lam) for x in gs]
    drop = np.cumsum(res)[:-1] / sum(res)
    np.testing.assert_allclose(r["potentials"], 1.0 - drop, rtol=1e-12)


def test_parallel_chains_add():
    a, lam, gap = 0.01, 0.02, 0.1
    p = _params((3.0, 1.0, 1.0), gap, lam)
    r1 = [_rod((0.5 + k, 0.25, 0.5), (1, 0, 0), 0.9, a) for k in range(3)]
    r2 = [_rod((0.76, 0.75, 0.5), (1, 0, 0), 1.4, a), _rod((2.24, 0.75, 0.5), (1, 0, 0), 1.4, a)]
    _, net = _network(p, np.stack(r1 + r2))
    g1 = 1.0 / sum(1.0 / _g(x, lam) for x in _end_gaps([0.5, 1.5, 2.5], [0.9] * 3, a, 3.0))
    g2 = 1.0 / sum(1.0 / _g(x, lam) for x in _end_gaps([0.76, 2.24], [1.4] * 2, a, 3.0))
    r = oracle.conductance(net)
    assert r["unknowns"] == 5
    assert r["conductance"] == pytest.approx(g1 + g2, rel=1e-12)


def test_one_sided_and_floating_components():
    """A chain touching only the source sits at 1, one touching only the sink
    at 0, a floating rod is NaN; no current flows."""
    a, lam, gap = 0.01, 0.02, 0.1
    p = _params((3.0, 1.0, 1.0), gap, lam)
    rods = np.stack([_rod((0.5, 0.2, 0.5), (1, 0, 0), 0.9, a),     # source only
                     _rod((2.5, 0.5, 0.5), (1, 0, 0), 0.9, a),     # sink only
                     _rod((1.5, 0.8, 0.5), (0, 1, 0), 0.1, a)])    # floating
    _, net = _network(p, rods)
    r = oracle.conductance(net)
    assert r["unknowns"] == 0 and r["conductance"] == 0.0
    v = r["potentials"]
    assert v[0] == 1.0 and v[1] == 0.0 and np.isnan(v[2])
