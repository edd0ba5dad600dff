"""Dead-end pruning on the GPU (nf_reduce) against the oracle: kept ids,
round count, the pruned packing bit for bit, and idempotence (pruning the
pruned packing again removes nothing)."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from nanofiller_amd import configs
from tests import cases

pytestmark = pytest.mark.gpu

CASES = {
    "iso_x": lambda: cases.small_iso(nf.NF_PREC_FP64, n_rods=2500, contact_gap=0.03, elec_axis=0),
    "tiny_box_z": lambda: cases.tiny_box(nf.NF_PREC_FP64, contact_gap=0.2, elec_axis=2),
    "network_slab_x": lambda: configs.cfg5(box=(40.0, 4.0, 4.0), n_rods=80000, contact_gap=0.02),
}


@pytest.mark.parametrize("name", list(CASES))
def test_reduce_matches_oracle(name):
    p = CASES[name]()
    rods, _, _ = oracle.generate(p, threads=16)
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    keep, rounds = oracle.reduce(p, rods, o["I"], o["J"], o["image"], p.contact_gap, 2)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        pk.contacts(p.contact_gap)
        info, ids = pk.reduce(2)
        pruned = pk.rods()
        c2 = pk.contacts(p.contact_gap)
        info2, ids2 = pk.reduce(2)
    print(name, info, flush=True)
    assert info["n_before"] == len(rods) and info["n_after"] == int(keep.sum())
    assert info["rounds"] == rounds
    np.testing.assert_array_equal(ids, np.flatnonzero(keep).astype(np.uint32))
    np.testing.assert_array_equal(pruned.view(np.uint32), rods[keep].view(np.uint32))
    # the contact list of the pruned packing is the oracle's on the kept rods
    o2 = oracle.contacts(p, rods[keep], p.contact_gap, threads=16)
    np.testing.assert_array_equal(c2.I, o2["I"])
    np.testing.assert_array_equal(c2.col, o2["J"])
    assert info2["n_after"] == info2["n_before"] == info["n_after"] and info2["rounds"] == 1
    np.testing.assert_array_equal(ids2, np.arange(info["n_after"], dtype=np.uint32))


def test_reduce_needs_contacts():
    p = cases.small_iso(nf.NF_PREC_FP64)
    with nf.Packing(p) as pk:
        pk.generate()
        with pytest.raises(nf.NFError) as e:
            pk.reduce(2)
    assert e.value.code == nf.NF_E_STATE
