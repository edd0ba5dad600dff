"""Oracle statistics beyond the histograms: deterministic atan, the
centre-centre distance statistics (statKernel_CENT) and the van der Waals
model (statKernel_VdW_naive), pinned independently:

* atan against libm;
* centre distances against a vectorised numpy recount (same FP64 operations,
  so bit-exact), and a cubic lattice whose neighbour shells are known;
* the VdW model on hand-built geometries (crossing rods, a T contact at a rod
  end, gaps beyond the cut-off, parallel rods), linearity in (A, B), and the
  attractive sign of the dispersion term.
"""
import math

import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases


def test_atan_against_libm():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(-50, 50, 20000),
                         10.0 ** rng.uniform(-12, 12, 20000), [0.0, 1.0, -1.0, 1e300, -1e-300]])
    for x in xs:
        a, b = oracle.atan(float(x)), math.atan(float(x))
        assert abs(a - b) <= 1e-15 * abs(b) + 1e-300, x  # a few ulp


def _numpy_centre(p, rods, bins, width):
    r = rods.astype(np.float64)
    n = len(r)
    rmax = bins * width
    e2 = (np.arange(bins + 1) * width) ** 2
    nim = [int(math.ceil(rmax / p.box[d])) + 1 for d in range(3)]
    counts = np.zeros(bins, np.int64)
    sq = s2q = cnt = 0
    i, j = np.triu_indices(n, 1)
    for iz in range(-nim[2], nim[2] + 1):
        for iy in range(-nim[1], nim[1] + 1):
            for ix in range(-nim[0], nim[0] + 1):
                dx = (r[i, 0] - r[j, 0]) - ix * p.box[0]
                dy = (r[i, 1] - r[j, 1]) - iy * p.box[1]
                dz = (r[i, 2] - r[j, 2]) - iz * p.box[2]
                d2 = dx * dx + dy * dy + dz * dz
                m = d2 < e2[bins]
                d2 = d2[m]
                b = np.searchsorted(e2, d2, side="right") - 1
                np.add.at(counts, b, 1)
                cnt += len(d2)
                sq += int(np.floor(np.sqrt(d2) * 16777216.0 + 0.5).astype(np.uint64).sum())
                s2q += int(np.floor(d2 * 65536.0 + 0.5).astype(np.uint64).sum())
    return counts, cnt, sq, s2q


def test_centre_distance_matches_numpy_recount():
    p = cases.small_iso(nf.NF_PREC_FP64, n_rods=600)
    rods, _, _ = oracle.generate(p, threads=4)
    bins, width = 24, 0.05
    o = oracle.centre_distance(p, rods, bins, width)
    counts, cnt, sq, s2q = _numpy_centre(p, rods, bins, width)
    np.testing.assert_array_equal(o["counts"], counts)
    assert o["pairs"] == cnt and o["sum_d_q24"] == sq and o["sum_d2_q16"] == s2q


def test_centre_distance_cubic_lattice():
    """Simple cubic lattice, spacing s, 4^3 sites in a periodic box of 4 s:
    shells at s (6 neighbours), s sqrt 2 (12), s sqrt 3 (8), 2 s (6)."""
    s = 0.5
    g = np.arange(4) * s + 0.25
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    n = 64
    rods = np.zeros((n, 8), np.float32)
    rods[:, 0], rods[:, 1], rods[:, 2] = X.ravel(), Y.ravel(), Z.ravel()
    rods[:, 5], rods[:, 6], rods[:, 7] = 1.0, 0.1, 0.005
    p = nf.normalize(nf.default_params(box=(2, 2, 2), n_rods=n, len=0.1, radius=0.005))
    width = 0.05
    o = oracle.centre_distance(p, rods, 42, width)  # up to 2.1 = 4.2 s
    c = o["counts"]
    k = lambda d: int(math.floor(d / width + 1e-9))  # bin of distance d
    assert c[k(s)] == n * 6 // 2
    assert c[k(s * math.sqrt(2))] == n * 12 // 2
    assert c[k(s * math.sqrt(3))] == n * 8 // 2
    assert c[k(2 * s)] == n * 6 // 2
    assert c.sum() == o["pairs"]


A, B = 15.2, 24100.0


def _rod(c, u, length, a=0.005):
    u = np.asarray(u, float)
    u = u / np.linalg.norm(u)
    return np.array([*c, *u, length, a], np.float32)


def test_vdw_crossing_rods_closed_form():
    """Perpendicular rods crossing at their centres with gap g: interior
    closest points, sin(gamma) = 1, excA = 4 a1 a2, theta = atan(a / (g + a
    - dlim)); energy from the curve-to-surface integral."""
    a, dlim = 0.005, 0.002
    for g in (0.0005, 0.001, 0.0019):
        r1 = _rod((1, 1, 1), (1, 0, 0), 1.0)
        z2 = np.float32(1 + 2 * a + g)
        r2 = _rod((1, 1, z2), (0, 1, 0), 1.0)
        e = oracle.vdw_pair(r1, r2, (0, 0, 0), dlim, A, B)
        d = float(z2) - 1.0 - 2 * float(np.float32(a))
        aa = float(np.float32(a))
        th = math.atan(aa / ((d + aa) - dlim))
        exc = 4 * aa * aa
        dA, dl, R = d * 1e4, dlim * 1e4, aa * 1e4
        q = 1 - th * dA / dl / dl / 2
        c2s = (-A * math.pi * R / 2 * th * (dA ** -4 - dl ** -4 / q)
               + B * 2 * math.pi * R / 5 * th * (dA ** -10 - dl ** -10 / q ** 4))
        assert e == pytest.approx(exc / th / aa * c2s, rel=1e-9)


def test_vdw_geometry_rules():
    a, dlim = 0.005, 0.002
    r1 = _rod((1, 1, 1), (1, 0, 0), 1.0)
    # T contact at the end of rod 2: closest point of rod 2 not interior
    r2 = _rod((1, 1, 1 + 2 * a + 0.001 + 0.5), (0, 0, 1), 1.0)
    assert oracle.vdw_pair(r1, r2, (0, 0, 0), dlim, A, B) is None
    # crossing but gap beyond the cut-off
    r3 = _rod((1, 1, 1 + 2 * a + 0.0021), (0, 1, 0), 1.0)
    assert oracle.vdw_pair(r1, r3, (0, 0, 0), dlim, A, B) is None
    # the periodic shift is applied to rod 2
    r4 = _rod((1, 1, 1 + 2 * a + 0.001 - 5.0), (0, 1, 0), 1.0)
    assert oracle.vdw_pair(r1, r4, (0, 0, 5.0), dlim, A, B) is not None
    # parallel rods side by side: the parallel excluded area 2 a l
    r5 = _rod((1, 1 + 2 * a + 0.001, 1), (1, 0, 0), 1.0)
    assert oracle.vdw_pair(r1, r5, (0, 0, 0), dlim, A, B) is not None


def test_vdw_linear_in_constants_and_attractive():
    p = cases.small_iso(nf.NF_PREC_FP64, n_rods=3000, contact_gap=0.01)
    rods, _, _ = oracle.generate(p, threads=4)
    c = oracle.contacts(p, rods, p.contact_gap, threads=4)
    eA, nA = oracle.vdw(p, rods, c["I"], c["J"], c["image"], p.contact_gap, 1.0, 0.0)
    eB, nB = oracle.vdw(p, rods, c["I"], c["J"], c["image"], p.contact_gap, 0.0, 1.0)
    e, n = oracle.vdw(p, rods, c["I"], c["J"], c["image"], p.contact_gap, A, B)
    assert n == nA == nB and 0 < n <= len(c["I"])
    assert e == pytest.approx(A * eA + B * eB, rel=1e-10)
    assert eA < 0.0  # dispersion alone attracts
