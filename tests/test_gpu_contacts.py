"""GPU contact detection and histograms against the oracle."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


def _packing(p):
    rods, _, _ = oracle.generate(p, threads=16)
    return rods


@pytest.mark.parametrize("engine", [nf.NF_ENGINE_FOOTPRINT, nf.NF_ENGINE_CELLS])
@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
@pytest.mark.parametrize("case", ["cfg1", "small_cfg2", "tiny_box"])
def test_contacts_parity(case, prec, engine):
    p = getattr(cases, case)(prec)
    rods = _packing(p)
    gap = p.contact_gap
    o = oracle.contacts(p, rods, gap, threads=16)
    with nf.Packing(p, {nf.NF_OPT_CONTACT_ENGINE: engine}) as pk:
        pk.set_rods(rods)
        c = pk.contacts(gap)
        st = pk.contact_stats()
        g64 = pk.contact_gap64(c.n)
    np.testing.assert_array_equal(c.I, o["I"])
    np.testing.assert_array_equal(c.col, o["J"])
    np.testing.assert_array_equal(c.image, o["image"])
    np.testing.assert_array_equal(c.gap.view(np.uint32), o["gap"].view(np.uint32))
    np.testing.assert_array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64))
    if engine == nf.NF_ENGINE_CELLS:
        assert st["pair_tests"] == o["pair_tests"]  # pairs within centre reach


@pytest.mark.parametrize("fine,occ,engine", [(100, 0, nf.NF_ENGINE_FOOTPRINT), (250, 0, nf.NF_ENGINE_FOOTPRINT),
                                              (600, 0, nf.NF_ENGINE_FOOTPRINT), (5000, 0, nf.NF_ENGINE_FOOTPRINT),
                                              (0, 8, nf.NF_ENGINE_FOOTPRINT), (0, 90, nf.NF_ENGINE_FOOTPRINT),
                                              (0, 200, nf.NF_ENGINE_CELLS)])
def test_footprint_cell_sizes_do_not_change_contacts(fine, occ, engine):
    """Any footprint fine cell (a tenth of a rod length to a whole home cell)
    and home cell size gives the same contact list: runs of home cells whose
    crossing rods overflow the survivor table in the agglomerates are redone
    one cell at a time (occupancy 90), and home cells so large that one cell
    overflows fall back to the cell engine (200), with the same list."""
    p = cases.small_cfg2(nf.NF_PREC_FP32_RECHECK)
    rods = _packing(p)
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    opts = {nf.NF_OPT_CONTACT_ENGINE: nf.NF_ENGINE_FOOTPRINT, nf.NF_OPT_FOOTPRINT_FINE: fine,
            nf.NF_OPT_FOOTPRINT_OCC: occ}
    with nf.Packing(p, opts) as pk:
        pk.set_rods(rods)
        c = pk.contacts(p.contact_gap)
        g64 = pk.contact_gap64(c.n)
        used = pk.contact_engine()
    print(fine, occ, used, flush=True)
    assert used["engine"] == engine
    if occ == 90:
        assert used["split_units"] > 0
    np.testing.assert_array_equal(c.I, o["I"])
    np.testing.assert_array_equal(c.col, o["J"])
    np.testing.assert_array_equal(c.image, o["image"])
    np.testing.assert_array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64))


def test_histograms_parity():
    p = cases.small_cfg2(nf.NF_PREC_FP32_RECHECK)
    rods = _packing(p)
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    ho = oracle.histograms(p, rods, o["I"], o["J"], bins=64, centre_axis=0, pol_axis=(0, 0, 1),
                           phi_bins=32, theta_bins=32)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        pk.contacts(p.contact_gap)
        h = pk.histograms(bins=64, centre_axis=0, pol_axis=(0, 0, 1), phi_bins=32, theta_bins=32)
    for k in ("centre", "polar", "mutual", "phitheta"):
        np.testing.assert_array_equal(h[k], ho[k])
    assert h["centre"].sum() == len(rods)
    assert h["mutual"].sum() == len(o["I"])


@pytest.mark.parametrize("case", ["cfg1", "small_cfg2", "tiny_box"])
def test_contact_area_parity(case, tmp_path):
    """Contact areas S on the device equal the oracle's bit for bit, and the
    COO writer's S column is their float rounding."""
    p = getattr(cases, case)(nf.NF_PREC_FP32_RECHECK, elec_axis=2)
    rods = _packing(p)
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        c = pk.contacts(p.contact_gap)
        S = pk.contact_area(c.n)
        pk.write_coo(str(tmp_path) + "/")
    So = oracle.contact_areas(p, rods, c.I, c.col, c.image, p.contact_gap)
    np.testing.assert_array_equal(S.view(np.uint64), So.view(np.uint64))
    assert c.n > 0 and np.all(S > 0.0)
    # internal.srt.coo lists the contacts whose image does not cross the z axis
    keep = ((c.image >> 16) & 0xFF) == 0
    rows = np.loadtxt(tmp_path / "internal.srt.coo", comments="#", ndmin=2)
    assert rows.shape[0] == int(keep.sum())
    # default stream formatting: 6 significant digits of the float value
    want = np.array([float("%g" % float(np.float32(v))) for v in S[keep]])
    np.testing.assert_array_equal(rows[:, 3], want)


def test_contact_area_branches():
    """Hand-built pairs hitting each branch of the area model (crossing
    ellipse, parallel strip, end-cap disc) and a periodic image: device and
    oracle agree bit for bit, and the branches give their closed forms."""
    import math

    a = 0.005
    D = 2 * a + 0.001
    rods = np.array([
        # perpendicular crossing at the centres
        [2.0, 2.0, 2.0, 1, 0, 0, 1.0, a], [2.0, 2.0, 2.0 + D, 0, 1, 0, 1.0, a],
        # side-by-side parallel, overlapping over 0.8
        [5.0, 2.0, 2.0, 1, 0, 0, 1.0, a], [5.2, 2.0 + D, 2.0, 1, 0, 0, 1.0, a],
        # end cap on a side
        [2.0, 5.0, 5.0, 1, 0, 0, 1.0, a], [2.0, 5.0, 5.4 + 2 * a + 0.001, 0, 0, 1, 0.8, a],
        # crossing through the periodic boundary in x
        [7.8, 5.0, 2.0, 1, 0, 0, 1.0, a], [0.1, 5.0, 2.0 + D, 0, 0.6, 0.8, 1.0, a],
    ], np.float32)
    p = nf.normalize(nf.default_params(box=(8, 8, 8), n_rods=len(rods), len=1.0, radius=a,
                                       contact_gap=0.002, precision=nf.NF_PREC_FP64))
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        c = pk.contacts(p.contact_gap)
        S = pk.contact_area(c.n)
    assert c.n == 4
    So = oracle.contact_areas(p, rods, c.I, c.col, c.image, p.contact_gap)
    np.testing.assert_array_equal(S.view(np.uint64), So.view(np.uint64))
    af = float(np.float32(a))
    h = max(float(np.float32(2.0 + D)) - 2.0 - af - 0.002, 0.0)
    dr = af * math.sqrt(1 - (h / af) ** 2)
    assert S[0] == pytest.approx(math.sqrt(2) * math.pi * dr * dr, rel=1e-9)
    assert S[1] == pytest.approx(2 * dr * (1.0 - (float(np.float32(5.2)) - 5.0)), rel=1e-6)
    assert S[2] == pytest.approx(math.pi * af * af, rel=1e-12)
    assert int(c.image[3]) != 0 and S[3] > math.pi * 1e-6
