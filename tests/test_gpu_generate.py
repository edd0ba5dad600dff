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

# generations whose pair-test count is compared test every pair within reach
EXACT = {nf.NF_OPT_TEST_ALL_PAIRS: 1}


def _compare(p, threads=16, options=None):
    rods_o, st_o, _ = oracle.generate(p, threads=threads)
    with nf.Packing(p, {**EXACT, **(options or {})}) as pk:
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


def test_graph_replay_agrees():
    """HIP-graph replay of the round (NF_OPT_GRAPH) gives the eager packing."""
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
    with nf.Packing(ps, EXACT) as pk:
        sts = pk.generate()
        rods = [pk.rods(i) for i in range(len(ps))]
    for i, p in enumerate(ps):
        ro, so, _ = oracle.generate(p, threads=16)
        assert sts[i].n_accepted == so["n_accepted"]
        assert sts[i].pair_tests == so["pair_tests"]
        np.testing.assert_array_equal(rods[i].view(np.uint32), ro.view(np.uint32))


@pytest.mark.parametrize("cell_size", [-0.5, -1.4, 0.6])
def test_cell_size_does_not_change_packing(cell_size):
    """Accept decisions depend on the candidate stream only: any cell edge
    (what the autotuner picks) gives the same packing and pair-test count."""
    p = cases.small_cfg2(nf.NF_PREC_FP32_RECHECK, n_rods=3000)
    q = p.copy()
    q.cell_size = cell_size
    with nf.Packing(p, EXACT) as pk:
        s0 = pk.generate()[0]
        r0 = pk.rods()
    with nf.Packing(q, EXACT) as pk:
        s1 = pk.generate()[0]
        r1 = pk.rods()
    np.testing.assert_array_equal(r0.view(np.uint32), r1.view(np.uint32))
    assert s0.pair_tests == s1.pair_tests and s0.rounds == s1.rounds


def test_work_schedule_does_not_change_packing():
    """A launch with few cells per wave splits dense cells over two work
    items and takes each queue's costliest cells first; one with many cells
    takes one item per cell in index order.  Both give the same packing and
    pair-test count: the dense packing alone (about 180 cells) against the
    same packing in a batch with a sparse 40 um companion (about 800k cells)."""
    p = cases.small_cfg2(nf.NF_PREC_FP32_RECHECK, n_rods=6000)
    big = nf.normalize(nf.default_params(box=(40, 40, 40), n_rods=1000, len=0.5, radius=0.005, seed=8,
                                         stream=3, precision=nf.NF_PREC_FP32_RECHECK))
    with nf.Packing(p, EXACT) as pk:
        s0 = pk.generate()[0]
        r0 = pk.rods()
    with nf.Packing([p, big], EXACT) as pk:
        s1 = pk.generate()[0]
        r1 = pk.rods(0)
    np.testing.assert_array_equal(r0.view(np.uint32), r1.view(np.uint32))
    assert s0.pair_tests == s1.pair_tests and s0.rounds == s1.rounds
    assert s0.rejected_old == s1.rejected_old and s0.rejected_new == s1.rejected_new


def test_long_rods_in_small_box_default_cells():
    """Rods almost as long as the box: the default cell edge keeps several
    cells per axis; the packing still equals the oracle's."""
    p = nf.normalize(nf.default_params(box=(9.2, 9.2, 9.2), n_rods=2500, len=4.9, radius=0.005,
                                       seed=3, stream=247, precision=nf.NF_PREC_FP32_RECHECK))
    ro, so, _ = oracle.generate(p, threads=16)
    with nf.Packing(p, EXACT) as pk:
        st = pk.generate()[0]
        rg = pk.rods()
    np.testing.assert_array_equal(rg.view(np.uint32), ro.view(np.uint32))
    assert st.pair_tests == so["pair_tests"]


@pytest.mark.parametrize("scale", [20, 60])
def test_slot_overflow_regenerates_the_same_packing(scale):
    """Cell slots planned far too small (2% or 6% of the plan) overflow;
    the packing is generated again with twice the slots until they suffice,
    identical to the default run, alone and in a batch.  The grown slots are
    kept for the next generation."""
    ps = [cases.small_cfg2(nf.NF_PREC_FP32_RECHECK), cases.small_iso(nf.NF_PREC_FP32_RECHECK)]
    refs = []
    for p in ps:
        with nf.Packing(p, EXACT) as pk:
            st = pk.generate()[0]
            refs.append((st, pk.rods()))
        assert st.slot_growths == 0
    with nf.Packing(ps[0], {**EXACT, nf.NF_OPT_SLOT_SCALE: scale}) as pk:
        st2 = pk.generate()[0]
        rods = pk.rods()
        st3 = pk.generate()[0]
    assert st2.slot_growths > 0 and st3.slot_growths == 0
    for s in (st2, st3):
        assert nf.STATUS[s.status] == nf.STATUS[refs[0][0].status]
        assert s.n_accepted == refs[0][0].n_accepted and s.pair_tests == refs[0][0].pair_tests
    np.testing.assert_array_equal(rods.view(np.uint32), refs[0][1].view(np.uint32))
    with nf.Packing(ps, {nf.NF_OPT_SLOT_SCALE: scale}) as pk:
        sts = pk.generate()
        for i in range(2):
            assert sts[i].n_accepted == refs[i][0].n_accepted and sts[i].rounds == refs[i][0].rounds
            np.testing.assert_array_equal(pk.rods(i).view(np.uint32), refs[i][1].view(np.uint32))
    assert sts[0].slot_growths > 0


@pytest.mark.parametrize("tiles", [{nf.NF_OPT_WAVES_OLD: 4, nf.NF_OPT_WAVES_NEW: 8},
                                   {nf.NF_OPT_PART_SHIFT: 0, nf.NF_OPT_SPLIT_MIN: 8},
                                   {nf.NF_OPT_PART_SHIFT: 3, nf.NF_OPT_SPLIT_MIN: 16, nf.NF_OPT_WAVES_OLD: 12}])
def test_tile_shape_does_not_change_packing(tiles):
    """The launch shape the autotuner picks (waves per CU, dense-cell split)
    changes speed only: the same packing and counters as the oracle."""
    p = cases.small_cfg2(nf.NF_PREC_FP32_RECHECK)
    ro, so, _ = oracle.generate(p, threads=16)
    with nf.Packing(p, {**EXACT, **tiles}) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
    assert st.n_accepted == so["n_accepted"] and st.pair_tests == so["pair_tests"]
    np.testing.assert_array_equal(rods.view(np.uint32), ro.view(np.uint32))


@pytest.mark.parametrize("case", ["small_cfg2_fp32", "cfg1_fp64", "jamming_batch"])
def test_cut_proposals_leave_the_walk_with_the_same_packing(case):
    """By default a proposal leaves its stencil walk once an accepted rod
    cuts it; with NF_OPT_TEST_ALL_PAIRS every pair within reach is tested.
    Both give the oracle's packing and counters, except the pair-test count,
    which the default can only lower.  The jamming batch (the 256-packing
    mix scaled down, 30-round cap) is where most proposals are cut."""
    from nanofiller_amd import configs

    if case == "small_cfg2_fp32":
        ps = [cases.small_cfg2(nf.NF_PREC_FP32_RECHECK)]
    elif case == "cfg1_fp64":
        ps = [cases.cfg1(nf.NF_PREC_FP64)]
    else:
        ps = configs.cfg3(n_inst=12, n_scale=1.0 / 40.0, max_rounds=30)
    out = {}
    for mode, opts in (("default", {}), ("exact", EXACT)):
        with nf.Packing(ps if len(ps) > 1 else ps[0], opts) as pk:
            sts = pk.generate()
            out[mode] = (sts, [pk.rods(i) for i in range(len(ps))])
    (sd, rd), (se, re_) = out["default"], out["exact"]
    cut = 0
    for i, p in enumerate(ps):
        ro, so, _ = oracle.generate(p, threads=16)
        for st, rods in ((sd[i], rd[i]), (se[i], re_[i])):
            np.testing.assert_array_equal(rods.view(np.uint32), ro.view(np.uint32), err_msg=str(i))
            assert nf.STATUS[st.status] == nf.STATUS[so["status"]], i
            assert st.rounds == so["rounds"] and st.candidates == so["candidates"], i
            assert st.rejected_old == so["rejected_old"] and st.rejected_new == so["rejected_new"], i
        assert se[i].pair_tests == so["pair_tests"], i
        assert 0 < sd[i].pair_tests <= so["pair_tests"], i
        cut += so["pair_tests"] - sd[i].pair_tests
    print(case, "pair tests skipped by the early exit:", cut, flush=True)
