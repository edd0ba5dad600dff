"""Post-processing of a batch, one instance at a time (NF_OPT_POST_INSTANCE):
contacts, histograms, network and the VTK writer of each instance of a
batched generation against the oracle on that instance's rods; switching
the instance drops the previous results."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


def test_each_instance_of_a_batch(tmp_path):
    prec = nf.NF_PREC_FP32_RECHECK
    ps = [cases.small_iso(prec, contact_gap=0.02, elec_axis=1), cases.small_cfg2(prec),
          cases.tiny_box(prec, elec_axis=2)]
    with nf.Packing(ps) as pk:
        pk.generate()
        for i, p in enumerate(ps):
            pk.select(i)
            rods = pk.rods(i)
            c = pk.contacts(p.contact_gap)
            g64 = pk.contact_gap64(c.n)
            h = pk.histograms(bins=32, centre_axis=1, pol_axis=(0, 0, 1))
            info, net = pk.network()
            pk.write_vtk(str(tmp_path / f"i{i}.vtk"))
            o = oracle.contacts(p, rods, p.contact_gap, threads=16)
            np.testing.assert_array_equal(c.I, o["I"], err_msg=str(i))
            np.testing.assert_array_equal(c.col, o["J"])
            np.testing.assert_array_equal(c.image, o["image"])
            np.testing.assert_array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64))
            ho = oracle.histograms(p, rods, o["I"], o["J"], bins=32, centre_axis=1, pol_axis=(0, 0, 1))
            for k in ("centre", "polar", "mutual"):
                np.testing.assert_array_equal(h[k], ho[k], err_msg=f"{i} {k}")
            oi, on = oracle.network(p, rods, o["I"], o["J"], o["gap64"], o["image"])
            assert info == oi, i
            np.testing.assert_array_equal(net["val"].view(np.uint64), on["val"].view(np.uint64))
            lines = (tmp_path / f"i{i}.vtk").read_text().splitlines()
            assert any(l.startswith(f"POINTS {len(rods)} ") for l in lines)
        # the network of instance 2 belongs to it: switching drops it
        pk.select(0)
        with pytest.raises(nf.NFError) as e:
            pk.network()
        assert e.value.code == nf.NF_E_STATE
        with pytest.raises(nf.NFError):
            pk.select(3)
