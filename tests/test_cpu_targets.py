"""Target-law bin probabilities (nanofiller_amd.targets) against the oracle's
candidate stream, which samples those laws before any overlap test, and the
agglomerate centres of the ABI against the oracle's."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from nanofiller_amd import targets


def draw(p, n):
    agg = oracle.agg_centres(p) if p.den_law == nf.NF_DEN_AGGLOM else None
    out = []
    for s in range(n):
        v, r = oracle.candidate(p, 0, s, agg)
        if v == 1:
            out.append(r)
    return np.array(out)


def test_agglomerate_centres_match_the_oracle():
    p = nf.normalize(nf.default_params(box=(20, 20, 20), n_rods=100, len=1.0, radius=0.005, seed=2,
                                       den_law=nf.NF_DEN_AGGLOM, n_agg=8, agg_sigma=1.5, agg_peak=3.0))
    np.testing.assert_array_equal(nf.agglomerate_centres(p), oracle.agg_centres(p))
    q = nf.normalize(nf.default_params(box=(5, 5, 5), n_rods=10, len=1.0, radius=0.005))
    assert nf.agglomerate_centres(q).shape == (0, 3)


def test_density_bins_against_candidates():
    p = nf.normalize(nf.default_params(box=(12, 12, 12), n_rods=100, len=1.0, radius=0.005, seed=5,
                                       den_law=nf.NF_DEN_AGGLOM, n_agg=6, agg_sigma=1.2, agg_peak=3.0))
    r = draw(p, 60000)
    pr = targets.density_bins(p, 1, 32)
    assert abs(pr.sum() - 1.0) < 1e-12 and pr.max() > 1.2 / 32
    h = np.bincount(np.minimum((r[:, 1] / 12.0 * 32).astype(int), 31), minlength=32)
    assert targets.chi2_test(h, pr)["p"] > 1e-4
    # a wrong profile (uniform) is rejected
    assert targets.chi2_test(h, np.full(32, 1 / 32))["p"] < 1e-6


@pytest.mark.parametrize("law,kappa,axis", [(nf.NF_ORT_ISO, 0.0, (0, 0, 1)),
                                            (nf.NF_ORT_VMF, 5.0, (0, 0, 1)),
                                            (nf.NF_ORT_VMF, 2.0, (1, 0, 0))])
def test_polar_bins_against_candidates(law, kappa, axis):
    p = nf.normalize(nf.default_params(box=(10, 10, 10), n_rods=100, len=1.0, radius=0.005, seed=9,
                                       ort_law=law, kappa=kappa or 1.0, axis=axis))
    r = draw(p, 40000)
    t = targets.packing_tests(p, r, 0, axis, bins=64)
    assert t["polar"]["p"] > 1e-4 and t["density"]["p"] > 1e-4
    # merged tail bins: every merged bin expects at least 5 counts
    assert t["polar"]["dof"] <= 63


def test_chi2_merges_small_bins():
    r = targets.chi2_test([0, 1, 100, 100, 2], [0.0005, 0.0005, 0.49, 0.49, 0.019])
    assert r["dof"] == 1 and r["n"] == 203  # bins {0,1,2}, {3,4}


def test_total_variation_distance():
    # exact frequencies: no distance; moving 10% of the mass between two bins: 0.1
    assert targets.chi2_test([25, 25, 50], [0.25, 0.25, 0.5])["tv"] == 0.0
    assert abs(targets.chi2_test([350, 150, 500], [0.25, 0.25, 0.5])["tv"] - 0.1) < 1e-12
