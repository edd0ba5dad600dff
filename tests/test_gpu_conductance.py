"""GPU conductance solve (Jacobi-PCG in FP64) against closed forms and the
oracle's direct solve on the same network.  Tolerance: relative 1e-8 on the
conductance and absolute 1e-8 on potentials in [0, 1], for a CG residual
||r|| / ||b|| <= 1e-12 on these well-conditioned networks."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases
from tests.test_cpu_network import _end_gaps, _g, _rod

pytestmark = pytest.mark.gpu


def _p(box, gap, lam, n, axis=0):
    return nf.normalize(nf.default_params(box=box, n_rods=n, len=0.9, radius=0.01, contact_gap=gap,
                                          elec_axis=axis, tunnel_lambda=lam, precision=nf.NF_PREC_FP64))


def _solve(p, rods=None, tol=1e-12):
    with nf.Packing(p) as pk:
        if rods is None:
            pk.generate()
        else:
            pk.set_rods(rods)
        pk.contacts(p.contact_gap)
        info, net = pk.network()
        ci, v = pk.conductance(tol=tol, max_iter=50000, potentials=True)
    return info, net, ci, v


def test_series_and_parallel_closed_forms():
    a, lam, gap = 0.01, 0.02, 0.1
    r1 = [_rod((0.5 + k, 0.25, 0.5), (1, 0, 0), 0.9, a) for k in range(3)]
    r2 = [_rod((0.76, 0.75, 0.5), (1, 0, 0), 1.4, a), _rod((2.24, 0.75, 0.5), (1, 0, 0), 1.4, a)]
    rods = np.stack(r1 + r2)
    _, _, ci, v = _solve(_p((3.0, 1.0, 1.0), gap, lam, 5), rods)
    g1 = 1.0 / sum(1.0 / _g(x, lam) for x in _end_gaps([0.5, 1.5, 2.5], [0.9] * 3, a, 3.0))
    g2 = 1.0 / sum(1.0 / _g(x, lam) for x in _end_gaps([0.76, 2.24], [1.4] * 2, a, 3.0))
    assert ci["converged"] == 1 and ci["unknowns"] == 5
    assert ci["conductance"] == pytest.approx(g1 + g2, rel=1e-10)
    assert ci["current_sink"] == pytest.approx(g1 + g2, rel=1e-10)
    res = [1.0 / _g(x, lam) for x in _end_gaps([0.5, 1.5, 2.5], [0.9] * 3, a, 3.0)]
    np.testing.assert_allclose(v[:3], 1.0 - np.cumsum(res)[:-1] / sum(res), atol=1e-10)


CASES = {
    "small_iso_percolating": lambda: cases.small_iso(nf.NF_PREC_FP64, contact_gap=0.05,
                                                     tunnel_lambda=0.01),
    "slab_x": lambda: nf.normalize(nf.default_params(box=(1.0, 6.0, 6.0), n_rods=20000, len=0.3,
                                                     radius=0.005, seed=5, contact_gap=0.004,
                                                     elec_axis=0, tunnel_lambda=0.002)),
    "dense_long_rods_z": lambda: nf.normalize(nf.default_params(box=(3.0, 3.0, 3.0), n_rods=6000,
                                                                len=1.2, radius=0.005, seed=9,
                                                                contact_gap=0.01, elec_axis=2,
                                                                tunnel_lambda=0.004)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_against_oracle(name):
    p = CASES[name]()
    info, net, ci, v = _solve(p)
    o = oracle.conductance(net)
    print(name, info["percolating"], ci, flush=True)
    assert ci["unknowns"] == o["unknowns"]
    assert ci["converged"] == 1
    if o["unknowns"] == 0:
        assert ci["conductance"] == 0.0
    else:
        assert ci["conductance"] == pytest.approx(o["conductance"], rel=1e-8)
        assert ci["current_sink"] == pytest.approx(o["conductance"], rel=1e-8)
    ov = o["potentials"]
    np.testing.assert_array_equal(np.isnan(v), np.isnan(ov))
    ok = ~np.isnan(ov)
    np.testing.assert_allclose(v[ok], ov[ok], atol=1e-8)


def test_no_percolation_and_state_errors():
    p = _p((3.0, 1.0, 1.0), 0.1, 0.02, 2)
    rods = np.stack([_rod((0.5, 0.2, 0.5), (1, 0, 0), 0.9, 0.01), _rod((2.5, 0.5, 0.5), (1, 0, 0), 0.9, 0.01)])
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        with pytest.raises(nf.NFError) as e:
            pk.conductance()
        assert e.value.code == nf.NF_E_STATE
        pk.contacts(p.contact_gap)
        pk.network()
        ci, v = pk.conductance(potentials=True)
    assert ci["unknowns"] == 0 and ci["conductance"] == 0.0 and ci["converged"] == 1
    np.testing.assert_array_equal(v, [1.0, 0.0])


def test_solve_is_reproducible():
    p = CASES["small_iso_percolating"]()
    with nf.Packing(p) as pk:
        pk.generate()
        pk.contacts(p.contact_gap)
        pk.network()
        a, va = pk.conductance(tol=1e-11, potentials=True)
        b, vb = pk.conductance(tol=1e-11, potentials=True)
    assert a["iterations"] == b["iterations"] and a["conductance"] == b["conductance"]
    np.testing.assert_array_equal(va, vb)
