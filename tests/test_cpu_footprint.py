"""Footprint contact pass (nf_seg.h) on the host, against the oracle.

- Coverage: every point within a_max + gap/2 of a rod's axis that lies in a
  home cell makes the rod cross that cell (seg_clip) and lies in a fine
  cell of the rod's fine mask there (seg_fine_mask), sampled with the FP64
  cell of the point.
- Near masks: every point within a_max + gap/2 of a rod's axis lies in a
  cell of its binned cell's 3 x 3 x 3 block that the mask (seg_near_mask)
  sets, or the mask is SEG_LONG.
- Contacts: the host emulation of the pass (every rod's capsule tested
  against the home cells it can reach, either per cell or per run of 4
  cells along x as the kernel's work units do, or by its near mask as the
  kernel does for rods that fit their block, fine masks, every pair
  sharing a fine cell through the same seg_test the GPU kernel runs) gives
  the oracle's contact list, FP64 gaps bit for bit, for several packings,
  home-cell edges (one cell per box edge included) and fine subdivisions,
  in both precisions.  Each contact is reported exactly once.
"""
import ctypes as C
import os

import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libnf_band.so")


@pytest.fixture(scope="module")
def H():
    L = C.CDLL(LIB)
    L.nfh_fine_mask.restype = C.c_uint64
    L.nfh_fine_mask.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.POINTER(C.c_int)]
    L.nfh_footprint_contacts.restype = C.c_int64
    L.nfh_footprint_contacts.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_double, C.c_int,
                                         C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_uint64)]
    L.nfh_near_mask.restype = C.c_uint32
    L.nfh_near_mask.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float]
    return L


@pytest.mark.parametrize("edge,f,length", [(1.0, 4, 1.0), (0.5, 2, 1.0), (0.3, 1, 2.0), (0.8, 3, 0.4)])
def test_fine_mask_covers_the_capsule(H, edge, f, length):
    rng = np.random.default_rng(5)
    a, gap = 0.005, 0.002
    r = np.float32((a + gap / 2) * 1.0001 + 1e-5)
    rtrue = a + gap / 2
    e = np.full(3, edge / f, np.float32)
    hits = 0
    for trial in range(400):
        lo = (rng.integers(0, 20, 3) * edge).astype(np.float32)
        d = rng.normal(size=3)
        if trial % 10 == 0:
            d = np.eye(3)[trial % 3] + 1e-9 * rng.normal(size=3)  # axis-aligned
        d /= np.linalg.norm(d)
        # a rod passing near the home box
        c = lo + rng.uniform(-0.3, 1.3, 3) * edge
        rod = np.array([*c, *d, length, a], np.float32)
        cross = C.c_int()
        m = H.nfh_fine_mask(rod.ctypes.data, lo.ctypes.data, e.ctypes.data, f, r, C.byref(cross))
        P = rod[:3].astype(np.float64)
        D = rod[3:6].astype(np.float64)
        hl = float(np.float32(0.5 * length))
        s = rng.uniform(-1, 1, 300) * hl
        v = rng.normal(size=(300, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        q = P[None] + s[:, None] * D[None] + v * rng.uniform(0, rtrue, 300)[:, None]
        for pt in q:
            k = np.floor((pt - lo.astype(np.float64)) / e.astype(np.float64)).astype(np.int64)
            if np.any(k < 0) or np.any(k >= f):
                continue  # outside the home box
            hits += 1
            assert cross.value == 1
            bit = int(k[0] + f * (k[1] + f * k[2]))
            assert (m >> bit) & 1, (trial, pt, k)
    assert hits > 100


@pytest.mark.parametrize("h,length", [(1.0, 0.6), (0.5, 0.45), (0.3, 0.1), (1.0, 2.5)])
def test_near_mask_covers_the_capsule(H, h, length):
    """Every point within a_max + gap/2 of the axis of a rod binned in cell
    cc lies in a cell of cc's 3 x 3 x 3 block whose bit the near mask sets,
    unless the mask is SEG_LONG; a rod that fits its block gets a mask, one
    whose capsule leaves the block gets SEG_LONG."""
    rng = np.random.default_rng(11)
    a, gap = 0.005, 0.002
    rtrue = a + gap / 2
    r = np.float32(rtrue * 1.0001 + 1e-5)
    mx = np.float32(4e-6)
    hh = np.full(3, h, np.float32)
    hits = masks = longs = 0
    for trial in range(300):
        cc = rng.integers(1, 30, 3).astype(np.int32)
        d = rng.normal(size=3)
        if trial % 10 == 0:
            d = np.eye(3)[trial % 3] + 1e-9 * rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = (cc + rng.uniform(0.0, 1.0, 3)) * h
        rod = np.array([*c, *d, length, a], np.float32)
        m = H.nfh_near_mask(rod.ctypes.data, cc.ctypes.data, hh.ctypes.data, r, mx)
        P = rod[:3].astype(np.float64)
        D = rod[3:6].astype(np.float64)
        hl = float(np.float32(0.5 * length))
        ext = np.abs(D) * hl + rtrue
        lo, hi = (cc - 1) * h, (cc + 2) * h
        clearly_in = np.all((P - ext >= lo + 1e-4) & (P + ext <= hi - 1e-4))
        clearly_out = np.any((P - ext < lo - 1e-4) | (P + ext > hi + 1e-4))
        if m == 1 << 31:
            longs += 1
            assert not clearly_in, (trial, P, D)
            continue
        assert not clearly_out, (trial, P, D)
        masks += 1
        assert m < (1 << 27)
        s = rng.uniform(-1, 1, 300) * hl
        v = rng.normal(size=(300, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        q = P[None] + s[:, None] * D[None] + v * rng.uniform(0, rtrue, 300)[:, None]
        k = np.floor(q / h).astype(np.int64) - cc[None].astype(np.int64)
        assert np.all(np.abs(k) <= 1), trial
        bits = (k[:, 0] + 1) + 3 * (k[:, 1] + 1) + 9 * (k[:, 2] + 1)
        assert np.all((m >> bits) & 1), (trial, m)
        hits += len(q)
    assert masks > 20 and hits > 5000
    if length > h:
        assert longs > 20


def test_near_mask_of_a_short_central_rod(H):
    """A short rod at the centre of its cell touches that cell alone (bit
    13); moved next to the cell's upper x face it also touches the x + 1
    neighbour (bit 14)."""
    cc = np.array([4, 5, 6], np.int32)
    hh = np.full(3, 1.0, np.float32)
    rod = np.array([4.5, 5.5, 6.5, 0, 0, 1, 0.2, 0.005], np.float32)
    assert H.nfh_near_mask(rod.ctypes.data, cc.ctypes.data, hh.ctypes.data, 0.006, 4e-6) == 1 << 13
    rod[0] = 4.997
    assert H.nfh_near_mask(rod.ctypes.data, cc.ctypes.data, hh.ctypes.data, 0.006, 4e-6) == (1 << 13) | (1 << 14)


def _emulate(H, p, rods, edge, f, prec, row, near):
    n = len(rods)
    r = np.ascontiguousarray(rods, np.float32)
    box = np.array([p.box[0], p.box[1], p.box[2]], np.float64)
    ms = C.c_uint64()
    z = np.zeros(1, np.uint32)
    cap = H.nfh_footprint_contacts(r.ctypes.data, n, box.ctypes.data, edge, f, p.contact_gap, prec, row, near,
                                   z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data, 0,
                                   C.byref(ms))
    cap = max(cap, 1)
    I = np.zeros(cap, np.uint32)
    J = np.zeros(cap, np.uint32)
    im = np.zeros(cap, np.int32)
    g = np.zeros(cap, np.float64)
    k = H.nfh_footprint_contacts(r.ctypes.data, n, box.ctypes.data, edge, f, p.contact_gap, prec, row, near,
                                 I.ctypes.data, J.ctypes.data, im.ctypes.data, g.ctypes.data, cap,
                                 C.byref(ms))
    assert k <= cap
    # the oracle's order: (I, J, image) with the image compared as signed
    # (x, y, z) shifts
    sb = [((im[:k] >> (8 * a)) & 0xFF).astype(np.int8) for a in range(3)]
    order = np.lexsort((sb[2], sb[1], sb[0], J[:k], I[:k]))
    return I[:k][order], J[:k][order], im[:k][order], g[:k][order], ms.value


CASES = {
    "cfg1": (lambda: cases.cfg1(nf.NF_PREC_FP64), [(0.5, 2), (1.25, 4)]),
    "small_cfg2": (lambda: cases.small_cfg2(nf.NF_PREC_FP32_RECHECK), [(0.6, 2), (0.3, 1), (1.25, 4)]),
    "tiny_box": (lambda: cases.tiny_box(nf.NF_PREC_FP64), [(0.5, 2), (2.0, 3)]),
    # rods longer than the box: a rod reaches its own periodic images
    "tiny_box_wide_gap": (lambda: cases.tiny_box(nf.NF_PREC_FP64, n_rods=40, contact_gap=0.05),
                          [(1.5, 2), (0.5, 2)]),
    "wide_gap": (lambda: cases.small_iso(nf.NF_PREC_FP64, contact_gap=0.05), [(0.5, 2), (1.0, 4)]),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
@pytest.mark.parametrize("row,near", [(1, 0), (4, 0), (4, 1)])
def test_emulated_contacts_match_oracle(H, name, prec, row, near):
    make, grids = CASES[name]
    p = make()
    rods, _, _ = oracle.generate(p, threads=8)
    o = oracle.contacts(p, rods, p.contact_gap, threads=8)
    assert len(o["I"]) > 0
    for edge, f in grids:
        I, J, im, g, ms = _emulate(H, p, rods, edge, f, prec, row, near)
        assert ms > 0
        np.testing.assert_array_equal(I, o["I"], err_msg=f"edge {edge} f {f}")
        np.testing.assert_array_equal(J, o["J"])
        np.testing.assert_array_equal(im, o["image"])
        np.testing.assert_array_equal(g.view(np.uint64), o["gap64"].view(np.uint64))
