"""expand and milestone state dumps on the GPU.

- nf_expand along x, y and z: the expanded packing bit for bit against
  oracle.expand with the buffer B the library chose, and B = ceil((l_max +
  gap) / E) E with E the generation cell edge (read from a saved state's
  header); contacts of the expanded packing against the oracle's.
- Milestones (VOLUME_PERCENT_LIST): the dumps are named by the fraction
  reached, ascend, hold prefixes of the final packing (rods are accepted in
  id order), and the final packing equals a run without milestones.
"""
import math
import os

import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("direction", [1, 2, 4])
def test_expand_matches_oracle(direction, tmp_path):
    p = cases.small_iso(nf.NF_PREC_FP32_RECHECK, box=(3.0, 2.5, 2.0), n_rods=2500, contact_gap=0.01)
    k = direction >> 1
    with nf.Packing(p) as pk:
        pk.generate()
        rods = pk.rods()
        pk.save_state(str(tmp_path / "s.bin"))
        with pk.expand(direction) as ex:
            q = ex.params()
            out = ex.rods()
            c = ex.contacts(q.contact_gap)
            g64 = ex.contact_gap64(c.n)
    hd = nf.read_state_header(str(tmp_path / "s.bin"))
    E = p.box[k] / hd["ext"][k]
    lmax = float(rods[:, 6].max())
    B = (q.box[k] - p.box[k]) / 2.0
    assert abs(B - math.ceil((lmax + p.contact_gap) / E) * E) < 1e-9
    for a in range(3):
        if a != k:
            assert q.box[a] == p.box[a]
    want = oracle.expand(rods, k, p.box[k], B)
    np.testing.assert_array_equal(out.view(np.uint32), want.view(np.uint32))
    o = oracle.contacts(q, want, q.contact_gap, threads=16)
    np.testing.assert_array_equal(c.I, o["I"])
    np.testing.assert_array_equal(c.col, o["J"])
    np.testing.assert_array_equal(c.image, o["image"])
    np.testing.assert_array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64))


def test_expand_rejects_a_buffer_wider_than_the_box():
    p = cases.tiny_box(nf.NF_PREC_FP64)
    with nf.Packing(p) as pk:
        pk.generate()
        with pytest.raises(nf.NFError) as e:
            pk.expand(1)
    assert e.value.code == nf.NF_E_INVALID


def test_milestones(tmp_path):
    p = cases.cfg1(nf.NF_PREC_FP32_RECHECK)
    with nf.Packing(p) as pk:
        pk.generate()
        final = pk.rods()
    ms = [0.3, 0.6, 0.9, 1.2]
    with nf.Packing(p) as pk:
        pk.set_milestones(ms, str(tmp_path))
        st = pk.generate()[0]
        rods = pk.rods()
    np.testing.assert_array_equal(rods.view(np.uint32), final.view(np.uint32))
    files = sorted(os.listdir(tmp_path))
    assert files, "no milestone written"
    vol = math.pi * p.radius ** 2 * p.len / (p.box[0] * p.box[1] * p.box[2]) * 100.0
    last_n, last_vf = 0, 0.0
    final_keys = {r.tobytes() for r in final}
    for f in sorted(files, key=lambda s: float(s[6:-4])):
        vf = float(f[6:-4])
        assert f == "state_%4.2f.bin" % vf
        hd = nf.read_state_header(str(tmp_path / f))
        n = hd["n_rods"]
        assert n > last_n and vf > last_vf and abs(n * vol - vf) < 0.006
        # the dump holds the first n accepted rods (as a set: the file is in slot order)
        with nf.Packing(nf.normalize(nf.default_params(box=tuple(p.box), n_rods=max(n, 1), len=p.len,
                                                       radius=p.radius))) as q:
            got = q.load_state(str(tmp_path / f))
            r = q.rods()
        assert got == n
        assert {x.tobytes() for x in r} == {x.tobytes() for x in final[:n]}
        assert {x.tobytes() for x in r} <= final_keys
        last_n, last_vf = n, vf
    # every milestone at or below the final fraction was passed by some dump
    assert last_vf >= max(m for m in ms if m <= st.n_accepted * vol)
