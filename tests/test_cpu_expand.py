"""oracle.expand (simPrms::expand, src/simPrms.cpp:744-852, copies chosen by
centre) against a hand-computed case and invariants of a generated packing:
the originals keep their ids, shifted by B; a copy exists exactly for each
rod within B of a face, one box edge away; nothing leaves [0, L + 2B)."""
import numpy as np

import oracle
from tests import cases


def test_hand_case():
    # L = 1, B = 0.2: 0.95 is within B of the far face (copy below at 0.15),
    # 0.1 within B of the near face (copy above at 1.3)
    xs = [0.1, 0.5, 0.95]
    r = np.zeros((3, 8), np.float32)
    r[:, 0] = xs
    r[:, 1] = [0.3, 0.6, 0.9]
    r[:, 3] = 1.0
    r[:, 6] = 0.2
    r[:, 7] = 0.005
    out = oracle.expand(r, 0, 1.0, 0.2)
    x = r[:, 0].astype(np.float64)  # the stored FP32 centres
    want = np.float32([x[0] + 0.2, x[1] + 0.2, x[2] + 0.2, x[2] + (0.2 - 1.0), x[0] + (0.2 + 1.0)])
    np.testing.assert_array_equal(out[:, 0], want)
    assert abs(float(out[3, 0]) - 0.15) < 1e-6 and abs(float(out[4, 0]) - 1.3) < 1e-6
    np.testing.assert_array_equal(out[:, 1], np.float32([0.3, 0.6, 0.9, 0.9, 0.3]))
    np.testing.assert_array_equal(out[:, 2:], r[[0, 1, 2, 2, 0], 2:])
    # along y with B = 0.05 no centre is within B of a face: no copies
    out = oracle.expand(r, 1, 1.0, 0.05)
    assert out.shape == (3, 8)


def test_invariants_of_a_packing():
    p = cases.cfg1()
    rods, _, _ = oracle.generate(p, threads=8)
    n = len(rods)
    for k in range(3):
        L = p.box[k]
        B = 1.2
        out = oracle.expand(rods, k, L, B)
        x = rods[:, k].astype(np.float64)
        below, above = x >= L - B, x < B
        assert len(out) == n + below.sum() + above.sum()
        # originals in id order, shifted by B (one FP32 rounding)
        np.testing.assert_array_equal(out[:n, k], (x + B).astype(np.float32))
        others = [a for a in range(8) if a != k]
        np.testing.assert_array_equal(out[:n][:, others], rods[:, others])
        nb = below.sum()
        cb, ca = out[n:n + nb], out[n + nb:]
        np.testing.assert_array_equal(cb[:, others], rods[below][:, others])
        np.testing.assert_array_equal(ca[:, others], rods[above][:, others])
        # each copy sits one box edge from its original, within rounding
        assert np.all(np.abs(cb[:, k].astype(np.float64) - (x[below] + B - L)) <= 4e-7 * (L + 2 * B))
        assert np.all(np.abs(ca[:, k].astype(np.float64) - (x[above] + B + L)) <= 4e-7 * (L + 2 * B))
        assert out[:, k].min() >= 0.0 and out[:, k].max() < L + 2 * B
        # the buffers hold every rod of the far and near slabs
        assert 0 < nb < n and 0 < above.sum() < n
