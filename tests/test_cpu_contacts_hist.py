"""Oracle contacts against an all-pairs, all-images brute force; oracle
histograms against numpy recounts (arccos binning)."""
import itertools
import math

import numpy as np

import nanofiller_amd as nf
import oracle


def packing(n=1500, box=3.0, **kw):
    base = dict(box=(box,) * 3, n_rods=n, len=1.0, radius=0.005, seed=12, contact_gap=0.002)
    base.update(kw)
    p = nf.normalize(nf.default_params(**base))
    rods, _, _ = oracle.generate(p, threads=8)
    return p, rods


def test_contacts_brute_force():
    p, rods = packing()
    gap = 0.02  # wider than the default to get many pairs
    o = oracle.contacts(p, rods, gap, threads=4)
    n = len(rods)
    ii, jj = np.triu_indices(n, 1)
    got, tests = [], 0
    for img in itertools.product([-1, 0, 1], repeat=3):
        sh = np.tile(np.array(img, np.float64) * 3.0, (len(ii), 1))
        d2, w, cut = oracle.pair_batch(rods[ii], rods[jj], sh, True, gap)
        tests += int(w.sum())
        for k in np.where(cut)[0]:
            g = math.sqrt(d2[k]) - (float(rods[ii[k], 7]) + float(rods[jj[k], 7]))
            got.append((int(ii[k]), int(jj[k]), img, g))
    got.sort(key=lambda t: (t[0], t[1], t[2]))
    assert len(got) == len(o["I"]) > 100
    assert tests == o["pair_tests"]
    for k, (i, j, img, g) in enumerate(got):
        assert (o["I"][k], o["J"][k]) == (i, j)
        code = o["image"][k]
        assert tuple(int(np.int8((code >> (8 * d)) & 0xFF)) for d in range(3)) == img
        assert o["gap64"][k] == g
        assert o["gap"][k] == np.float32(g)


def test_histograms_recount():
    p, rods = packing(n=1200, ort_law=nf.NF_ORT_VMF, kappa=2.0, axis=(1, 0, 0))
    o = oracle.contacts(p, rods, 0.02, threads=4)
    nb = 64
    h = oracle.histograms(p, rods, o["I"], o["J"], bins=nb, centre_axis=0, pol_axis=(1, 0, 0),
                          phi_bins=16, theta_bins=8)
    x = rods[:, 0].astype(np.float64)
    cb = np.clip((x / 3.0 * nb).astype(int), 0, nb - 1)
    np.testing.assert_array_equal(h["centre"], np.bincount(cb, minlength=nb))
    c = rods[:, 3:6].astype(np.float64)
    cn = np.linalg.norm(c, axis=1)
    th = np.arccos(np.clip(c[:, 0] / cn, -1, 1))
    pb = np.clip(np.floor(th / (np.pi / nb)).astype(int), 0, nb - 1)
    np.testing.assert_array_equal(h["polar"], np.bincount(pb, minlength=nb))
    a, b = c[o["I"]], c[o["J"]]
    g = np.arccos(np.clip(np.abs((a * b).sum(1)) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1)), 0, 1))
    mb = np.clip(np.floor(g / (np.pi / 2 / nb)).astype(int), 0, nb - 1)
    np.testing.assert_array_equal(h["mutual"], np.bincount(mb, minlength=nb))
    # phi-theta: theta from +z over [0, pi], phi over [0, 2 pi)
    t2 = np.clip(np.floor(np.arccos(np.clip(c[:, 2] / cn, -1, 1)) / (np.pi / 8)).astype(int), 0, 7)
    ph = np.mod(np.arctan2(c[:, 1], c[:, 0]), 2 * np.pi)
    p2 = np.clip(np.floor(ph / (2 * np.pi / 16)).astype(int), 0, 15)
    ref = np.zeros((8, 16), np.int64)
    np.add.at(ref, (t2, p2), 1)
    np.testing.assert_array_equal(h["phitheta"], ref)
    assert h["centre"].sum() == len(rods) and h["mutual"].sum() == len(o["I"])
