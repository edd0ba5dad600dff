"""GPU network extraction (components + FP64 Kirchhoff CSR) against the oracle, bit for bit."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


def _slab(prec=nf.NF_PREC_FP64):
    """Thin slab along the electrode axis: > 1024 rods couple to each electrode,
    so the electrode diagonals use several 1024-entry blocks."""
    return nf.normalize(nf.default_params(box=(1.0, 6.0, 6.0), n_rods=20000, len=0.3, radius=0.005,
                                          seed=5, contact_gap=0.004, elec_axis=0,
                                          tunnel_lambda=0.002, precision=prec))


CASES = {
    "small_iso_percolating": lambda: cases.small_iso(nf.NF_PREC_FP64, contact_gap=0.05,
                                                     tunnel_lambda=0.01),
    "tiny_box_images_z": lambda: cases.tiny_box(nf.NF_PREC_FP64, contact_gap=0.2, elec_axis=2),
    "cfg1_fp32": lambda: cases.cfg1(nf.NF_PREC_FP32_RECHECK, elec_axis=1),
    "slab_many_couplings": _slab,
}


def _compare(p, rods, gpu_info, gpu_net):
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    oi, on = oracle.network(p, rods, o["I"], o["J"], o["gap64"], o["image"])
    assert gpu_info == oi
    for k in ("labels", "row_ptr", "col"):
        np.testing.assert_array_equal(gpu_net[k], on[k], err_msg=k)
    np.testing.assert_array_equal(gpu_net["val"].view(np.uint64), on["val"].view(np.uint64))
    return oi


@pytest.mark.parametrize("case", list(CASES))
def test_network_parity(case):
    p = CASES[case]()
    rods, _, _ = oracle.generate(p, threads=16)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        pk.contacts(p.contact_gap)
        info, net = pk.network()
    oi = _compare(p, rods, info, net)
    if case == "slab_many_couplings":
        assert oi["n_source"] > 1024 and oi["n_sink"] > 1024


def test_network_after_gpu_generation():
    """Generate, contacts and network all on the device; oracle on the same rods."""
    p = cases.small_cfg2(nf.NF_PREC_FP32_RECHECK, contact_gap=0.02, tunnel_lambda=0.005)
    with nf.Packing(p) as pk:
        pk.generate()
        rods = pk.rods()
        pk.contacts(p.contact_gap)
        info, net = pk.network()
        # a second call on the same contacts gives the same matrix
        info2, net2 = pk.network()
    assert info2 == info
    np.testing.assert_array_equal(net2["val"].view(np.uint64), net["val"].view(np.uint64))
    _compare(p, rods, info, net)


def test_network_without_contacts_fails():
    p = cases.small_iso()
    with nf.Packing(p) as pk:
        with pytest.raises(nf.NFError):
            pk.network()
