"""The FP32 capsule predicate with its error band, fuzzed against the FP64
oracle on the CPU (host build of the library's own predicate).

Whenever the FP32 path decides (does not ask for an FP64 re-check), its
reach and cut decisions must equal the FP64 oracle's.  Cases concentrate
where FP32 is weakest: near-parallel rods, distances within 1e-10..1e-3 of
the cut-off, far-from-origin coordinates with periodic shifts.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libnf_band.so")


@pytest.fixture(scope="module")
def band():
    if not os.path.exists(LIB):
        pytest.skip("band harness not built (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    L.nfh_pair_f32_batch.restype = None
    L.nfh_pair_f32_batch.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_double, C.c_double, C.c_float, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
    return L


def f32_batch(L, ra, rb, sh, contact, cut, ew):
    n = len(ra)
    dec = np.zeros(n, np.int32)
    wi = np.zeros(n, np.int32)
    d2 = np.zeros(n, np.float32)
    ra = np.ascontiguousarray(ra, np.float32)
    rb = np.ascontiguousarray(rb, np.float32)
    sh = np.ascontiguousarray(sh, np.float64)
    L.nfh_pair_f32_batch(n, ra.ctypes.data, rb.ctypes.data, sh.ctypes.data, int(contact), cut, 1.0,
                         ew, dec.ctypes.data, wi.ctypes.data, d2.ctypes.data)
    return dec, wi.astype(bool), d2


def make_pairs(rng, n, box, target, near_parallel):
    """Pairs whose segment distance is close to `target` (before float rounding)."""
    ra, rb, sh = [], [], []
    for _ in range(n):
        l1, l2 = rng.uniform(0.3, 3.0, 2)
        u1 = rng.normal(size=3)
        u1 /= np.linalg.norm(u1)
        if near_parallel:
            u2 = u1 + rng.normal(size=3) * 10.0 ** rng.uniform(-8, -2)
        else:
            u2 = rng.normal(size=3)
        u2 /= np.linalg.norm(u2)
        nrm = np.cross(u1, u2)
        if np.linalg.norm(nrm) < 1e-12:
            nrm = np.cross(u1, [1.0, 0.0, 0.0])
        nrm /= np.linalg.norm(nrm)
        s = rng.uniform(-0.4, 0.4) * l1
        t = rng.uniform(-0.4, 0.4) * l2
        d = target * (1.0 + rng.choice([-1, 1]) * 10.0 ** rng.uniform(-10, -3))
        c1 = rng.uniform(0, box, 3)
        c2 = c1 + s * u1 + d * nrm - t * u2
        img = rng.integers(-1, 2, 3)
        c2w = c2 - img * box  # wrapped copy, the shift brings it back
        ra.append([*c1, *u1, l1, 0.005])
        rb.append([*c2w, *u2, l2, 0.005])
        sh.append(img * box)
    return np.array(ra, np.float32), np.array(rb, np.float32), np.array(sh, np.float64)


@pytest.mark.parametrize("contact", [False, True])
@pytest.mark.parametrize("near_parallel", [False, True])
@pytest.mark.parametrize("box", [5.0, 50.0])
def test_band_never_decides_wrong(band, contact, near_parallel, box):
    rng = np.random.default_rng(int(contact) * 4 + int(near_parallel) * 2 + int(box))
    cut = 0.002 if contact else 0.0
    target = 0.01 + cut
    ra, rb, sh = make_pairs(rng, 20000, box, target, near_parallel)
    reach = 3.0 + 0.01 + cut
    ew = np.float32(2.0 ** -24 * (box + reach) * 1.01)
    dec, w32, _ = f32_batch(band, ra, rb, sh, contact, cut, ew)
    _, w64, c64 = oracle.pair_batch(ra, rb, sh, contact, cut)
    decided = dec >= 0
    assert np.all(w32[decided] == w64[decided])
    assert np.all((dec[decided] == 1) == c64[decided])
    # adversarial set (distances within 1e-10..1e-3 relative of the cut):
    # the band is a few 1e-2 of R^2 wide, so most are re-checked, not all
    assert decided.mean() > 0.01


def test_band_recheck_rate_on_random_pairs(band):
    rng = np.random.default_rng(11)
    n = 50000
    box = 20.0
    c1 = rng.uniform(0, box, (n, 3))
    c2 = c1 + rng.normal(size=(n, 3)) * 0.8
    u1 = rng.normal(size=(n, 3))
    u2 = rng.normal(size=(n, 3))
    u1 /= np.linalg.norm(u1, axis=1)[:, None]
    u2 /= np.linalg.norm(u2, axis=1)[:, None]
    L = rng.uniform(0.3, 3.0, (n, 2))
    ra = np.column_stack([c1, u1, L[:, 0], np.full(n, 0.005)]).astype(np.float32)
    rb = np.column_stack([c2, u2, L[:, 1], np.full(n, 0.005)]).astype(np.float32)
    sh = np.zeros((n, 3))
    ew = np.float32(2.0 ** -24 * (box + 3.1) * 1.01)
    dec, w32, _ = f32_batch(band, ra, rb, sh, False, 0.0, ew)
    _, w64, c64 = oracle.pair_batch(ra, rb, sh, False, 0.0)
    decided = dec >= 0
    assert np.all((dec[decided] == 1) == c64[decided])
    assert np.all(w32[decided] == w64[decided])
    assert (~decided).mean() < 1e-3


def screen_batch(L, ra, rb, sh, contact, cut, ew):
    L.nfh_pair_screen_batch.restype = None
    L.nfh_pair_screen_batch.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_double, C.c_double, C.c_float, C.c_void_p]
    n = len(ra)
    out = np.zeros(n, np.int32)
    ra = np.ascontiguousarray(ra, np.float32)
    rb = np.ascontiguousarray(rb, np.float32)
    sh = np.ascontiguousarray(sh, np.float64)
    L.nfh_pair_screen_batch(n, ra.ctypes.data, rb.ctypes.data, sh.ctypes.data, int(contact), cut,
                            1.0, ew, out.ctypes.data)
    return out


SCREEN_UNDECIDED, SCREEN_OUT, SCREEN_FAR, SCREEN_FULL = -1, 0, 1, 2


def _check_screen(scr, w64, c64):
    assert not np.any(w64[scr == SCREEN_OUT])
    assert np.all(w64[scr == SCREEN_FAR]) and not np.any(c64[scr == SCREEN_FAR])
    assert np.all(w64[scr == SCREEN_FULL])


@pytest.mark.parametrize("contact", [False, True])
@pytest.mark.parametrize("near_parallel", [False, True])
@pytest.mark.parametrize("box", [5.0, 50.0])
def test_screen_never_drops_a_cut(band, contact, near_parallel, box):
    """Line-distance screen: 'far' only when FP64 says within and not cut,
    on pairs built at distances within 1e-10..1e-3 of the cut-off."""
    rng = np.random.default_rng(100 + int(contact) * 4 + int(near_parallel) * 2 + int(box))
    cut = 0.002 if contact else 0.0
    ra, rb, sh = make_pairs(rng, 20000, box, 0.01 + cut, near_parallel)
    ew = np.float32(2.0 ** -24 * (box + 3.0 + 0.01 + cut) * 1.01)
    scr = screen_batch(band, ra, rb, sh, contact, cut, ew)
    _, w64, c64 = oracle.pair_batch(ra, rb, sh, contact, cut)
    _check_screen(scr, w64, c64)


def test_screen_rejects_most_random_pairs(band):
    """Random pairs within reach: most are settled by the line distance."""
    rng = np.random.default_rng(12)
    n = 50000
    box = 20.0
    c1 = rng.uniform(0, box, (n, 3))
    c2 = c1 + rng.normal(size=(n, 3)) * 0.8
    u1 = rng.normal(size=(n, 3))
    u2 = rng.normal(size=(n, 3))
    u1 /= np.linalg.norm(u1, axis=1)[:, None]
    u2 /= np.linalg.norm(u2, axis=1)[:, None]
    L = rng.uniform(0.3, 3.0, (n, 2))
    ra = np.column_stack([c1, u1, L[:, 0], np.full(n, 0.005)]).astype(np.float32)
    rb = np.column_stack([c2, u2, L[:, 1], np.full(n, 0.005)]).astype(np.float32)
    sh = np.zeros((n, 3))
    ew = np.float32(2.0 ** -24 * (box + 3.1) * 1.01)
    for contact, cut in ((False, 0.0), (True, 0.002)):
        scr = screen_batch(band, ra, rb, sh, contact, cut, ew)
        _, w64, c64 = oracle.pair_batch(ra, rb, sh, contact, cut)
        _check_screen(scr, w64, c64)
        within = w64.sum()
        assert (scr == SCREEN_FAR).sum() > 0.9 * within
        assert (scr == SCREEN_UNDECIDED).mean() < 1e-3


def test_screen_parallel_and_crossing_closed_forms(band):
    """Closed forms: crossing perpendicular rods at offset h are 'far' only
    for h clearly above R; exactly parallel rods are never 'far'."""
    a = 0.005
    R = 2 * a
    ew = np.float32(2.0 ** -24 * 10 * 1.01)
    hs = np.array([0.5 * R, 0.999 * R, R, 1.001 * R, 1.1 * R, 3 * R])
    ra = np.array([[1, 1, 1, 1, 0, 0, 1.0, a]] * len(hs), np.float32)
    rb = np.array([[1, 1, 1 + h, 0, 1, 0, 1.0, a] for h in hs], np.float32)
    sh = np.zeros((len(hs), 3))
    scr = screen_batch(band, ra, rb, sh, False, 0.0, ew)
    assert list(scr[:3]) == [SCREEN_FULL] * 3
    assert scr[4] == SCREEN_FAR and scr[5] == SCREEN_FAR
    par = np.array([[1, 1, 1 + h, 1, 0, 0, 1.0, a] for h in hs], np.float32)
    scr = screen_batch(band, ra, par, sh, False, 0.0, ew)
    assert not np.any(scr == SCREEN_FAR)
