"""GPU generation parity: the HIP insert/reject rounds against the FP64 oracle.

Accepted rods must be bit-identical (same ids, same float values), and the
algorithmic counters (rounds, candidates, proposals, pair tests, rejections)
must match exactly, for the plain FP64 path and for the FP32 path with its
FP64 re-check band.
"""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


def _compare(p, threads=16, options=None):
    rods_o, st_o, _ = oracle.generate(p, threads=threads)
    with nf.Packing(p, options) as pk:
        st = pk.generate()[0]
        rods_g = pk.rods()
    assert st.n_accepted == st_o["n_accepted"]
    assert st.rounds == st_o["rounds"]
    assert st.candidates == st_o["candidates"]
    assert st.proposals == st_o["proposals"]
    assert st.pair_tests == st_o["pair_tests"]
    assert st.rejected_old == st_o["rejected_old"]
    assert st.rejected_new == st_o["rejected_new"]
    assert nf.STATUS[st.status] == nf.STATUS[st_o["status"]]
    assert rods_g.shape == rods_o.shape
    np.testing.assert_array_equal(rods_g.view(np.uint32), rods_o.view(np.uint32))
    return st, st_o


@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
def test_small_iso(prec):
    _compare(cases.small_iso(prec))


@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
def test_cfg1(prec):
    st, _ = _compare(cases.cfg1(prec))
    assert st.n_accepted == 20000


@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
def test_small_cfg2_laws(prec):
    _compare(cases.small_cfg2(prec))


def test_tiny_box_images():
    _compare(cases.tiny_box(nf.NF_PREC_FP64))
    _compare(cases.tiny_box(nf.NF_PREC_FP32_RECHECK))


def test_eager_and_profiled_agree():
    p = cases.small_iso(nf.NF_PREC_FP32_RECHECK)
    outs = []
    for opts in ({nf.NF_OPT_GRAPH: 0}, {nf.NF_OPT_PROFILE: 1}):
        with nf.Packing(p, opts) as pk:
            pk.generate()
            outs.append(pk.rods())
        with nf.Packing(p, opts) as pk:  # a second generate on the same handle restarts cleanly
            pk.generate()
            pk.generate()
            outs.append(pk.rods())
    for o in outs[1:]:
        np.testing.assert_array_equal(o.view(np.uint32), outs[0].view(np.uint32))


@pytest.mark.skip(reason="HIP-graph replay of the generation round timed out on the GPU with no "
                         "cause found on the CPU; the path is off by default until one is shown")
def test_graph_replay_agrees():
    p = cases.small_iso(nf.NF_PREC_FP32_RECHECK)
    with nf.Packing(p, {nf.NF_OPT_GRAPH: 1}) as pk:
        pk.generate()
        g = pk.rods()
    with nf.Packing(p, {nf.NF_OPT_GRAPH: 0}) as pk:
        pk.generate()
        e = pk.rods()
    np.testing.assert_array_equal(g.view(np.uint32), e.view(np.uint32))


def test_batch_matches_single():
    ps = [cases.small_iso(nf.NF_PREC_FP32_RECHECK, seed=s, stream=i, n_rods=500 + 300 * i)
          for i, s in enumerate([3, 4, 5, 6])]
    ps.append(cases.small_cfg2(nf.NF_PREC_FP32_RECHECK, n_rods=2000, stream=9))
    with nf.Packing(ps) as pk:
        sts = pk.generate()
        rods = [pk.rods(i) for i in range(len(ps))]
    for i, p in enumerate(ps):
        ro, so, _ = oracle.generate(p, threads=16)
        assert sts[i].n_accepted == so["n_accepted"]
        assert sts[i].pair_tests == so["pair_tests"]
        np.testing.assert_array_equal(rods[i].view(np.uint32), ro.view(np.uint32))
