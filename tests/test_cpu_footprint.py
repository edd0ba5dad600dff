"""Footprint contact pass (nf_seg.h) on the host, against the oracle.

- Coverage: every point within a_max + gap/2 of a rod's axis that lies in a
  home cell makes the rod cross that cell (seg_clip) and lies in a fine
  cell of the rod's fine mask there (seg_fine_mask), sampled with the FP64
  cell of the point.
- Contacts: the host emulation of the pass (every rod's capsule tested
  against the home cells it can reach, either per cell or per run of 4
  cells along x as the kernel's work units do, fine masks, every pair
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
                                         C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int64, C.POINTER(C.c_uint64)]
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


def _emulate(H, p, rods, edge, f, prec, row):
    n = len(rods)
    r = np.ascontiguousarray(rods, np.float32)
    box = np.array([p.box[0], p.box[1], p.box[2]], np.float64)
    ms = C.c_uint64()
    z = np.zeros(1, np.uint32)
    cap = H.nfh_footprint_contacts(r.ctypes.data, n, box.ctypes.data, edge, f, p.contact_gap, prec, row,
                                   z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data, 0,
                                   C.byref(ms))
    cap = max(cap, 1)
    I = np.zeros(cap, np.uint32)
    J = np.zeros(cap, np.uint32)
    im = np.zeros(cap, np.int32)
    g = np.zeros(cap, np.float64)
    k = H.nfh_footprint_contacts(r.ctypes.data, n, box.ctypes.data, edge, f, p.contact_gap, prec, row,
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
    "small_cfg2": (lambda: cases.small_cfg2(nf.NF_PREC_FP32_RECHECK), [(0.6, 2), (0.3, 1)]),
    "tiny_box": (lambda: cases.tiny_box(nf.NF_PREC_FP64), [(0.5, 2), (2.0, 3)]),
    # rods longer than the box: a rod reaches its own periodic images
    "tiny_box_wide_gap": (lambda: cases.tiny_box(nf.NF_PREC_FP64, n_rods=40, contact_gap=0.05),
                          [(1.5, 2), (0.5, 2)]),
    "wide_gap": (lambda: cases.small_iso(nf.NF_PREC_FP64, contact_gap=0.05), [(0.5, 2)]),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("prec", [nf.NF_PREC_FP64, nf.NF_PREC_FP32_RECHECK])
@pytest.mark.parametrize("row", [1, 4])
def test_emulated_contacts_match_oracle(H, name, prec, row):
    make, grids = CASES[name]
    p = make()
    rods, _, _ = oracle.generate(p, threads=8)
    o = oracle.contacts(p, rods, p.contact_gap, threads=8)
    assert len(o["I"]) > 0
    for edge, f in grids:
        I, J, im, g, ms = _emulate(H, p, rods, edge, f, prec, row)
        assert ms > 0
        np.testing.assert_array_equal(I, o["I"], err_msg=f"edge {edge} f {f}")
        np.testing.assert_array_equal(J, o["J"])
        np.testing.assert_array_equal(im, o["image"])
        np.testing.assert_array_equal(g.view(np.uint64), o["gap64"].view(np.uint64))
