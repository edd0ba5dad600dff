"""Candidate sampler of the oracle against the target distributions it
implements (the reference's target profiles: density map accept test,
orientation law, rod length), by goodness-of-fit tests on its own stream."""
import numpy as np
import pytest
from scipy import stats

import nanofiller_amd as nf
import oracle


def draw(p, n, rnd=0):
    agg = oracle.agg_centres(p) if p.den_law == nf.NF_DEN_AGGLOM else None
    out, vals = [], []
    for s in range(n):
        v, r = oracle.candidate(p, rnd, s, agg)
        vals.append(v)
        if v == 1:
            out.append(r)
    return np.array(out), np.array(vals)


def P(**kw):
    base = dict(box=(10, 10, 10), n_rods=1000, len=1.0, radius=0.005, seed=21)
    base.update(kw)
    return nf.normalize(nf.default_params(**base))


def test_uniform_positions_chi2():
    r, v = draw(P(), 40000)
    assert (v == 1).all()
    h, _ = np.histogramdd(r[:, :3], bins=(8, 8, 8), range=[(0, 10)] * 3)
    assert stats.chisquare(h.ravel()).pvalue > 1e-4
    assert (r[:, :3] >= 0).all() and (r[:, :3] < 10).all()


def test_isotropic_directions():
    r, _ = draw(P(), 40000)
    n = np.linalg.norm(r[:, 3:6].astype(np.float64), axis=1)
    assert np.abs(n - 1).max() < 1e-6
    assert stats.kstest(r[:, 5], "uniform", args=(-1, 2)).pvalue > 1e-4
    phi = np.arctan2(r[:, 4], r[:, 3])
    assert stats.kstest(phi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > 1e-4


@pytest.mark.parametrize("kappa,axis", [(5.0, (0, 0, 1)), (2.0, (1, 0, 0)), (20.0, (1, 2, -2))])
def test_vmf_directions(kappa, axis):
    r, _ = draw(P(ort_law=nf.NF_ORT_VMF, kappa=kappa, axis=axis), 40000)
    mu = np.array(axis, np.float64) / np.linalg.norm(axis)
    w = r[:, 3:6].astype(np.float64) @ mu
    cdf = lambda x: (np.exp(kappa * (x - 1)) - np.exp(-2 * kappa)) / (1 - np.exp(-2 * kappa))
    assert stats.kstest(np.clip(w, -1, 1), cdf).pvalue > 1e-4
    # azimuth about the axis is uniform
    e1, e2 = np.zeros(3), np.zeros(3)
    oracle.lib().orc_vmf_basis(mu.ctypes.data, e1.ctypes.data, e2.ctypes.data)
    ang = np.arctan2(r[:, 3:6] @ e2, r[:, 3:6] @ e1)
    assert stats.kstest(ang, "uniform", args=(-np.pi, 2 * np.pi)).pvalue > 1e-4


def test_lognormal_lengths_clipped():
    p = P(len_law=nf.NF_LEN_LOGNORMAL, len=1.0, len_sigma=0.3, len_min=0.3, len_max=3.0)
    r, _ = draw(p, 40000)
    L = r[:, 6].astype(np.float64)
    assert L.min() >= np.float32(0.3) and L.max() <= np.float32(3.0)
    z = np.log(L[(L > 0.31) & (L < 2.9)])
    # truncated normal on the unclipped range
    a, b = np.log(0.31) / 0.3, np.log(2.9) / 0.3
    assert stats.kstest(z / 0.3, stats.truncnorm(a, b).cdf).pvalue > 1e-4


def test_agglomerate_acceptance_rate():
    p = P(den_law=nf.NF_DEN_AGGLOM, n_agg=8, agg_sigma=1.5, agg_peak=3.0, box=(20, 20, 20))
    r, v = draw(p, 60000)
    acc = (v == 1).mean()
    # expected acceptance E[rho]/peak by independent Monte Carlo over the box
    c = oracle.agg_centres(p)
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 20, (400000, 3))
    d = x[:, None, :] - c[None, :, :]
    d -= 20 * np.round(d / 20)
    d2 = (d ** 2).sum(-1).min(1)
    rho = 1 + 2 * np.exp(-d2 / (2 * 1.5 ** 2))
    expect = rho.mean() / 3.0
    sd = np.sqrt(expect * (1 - expect) / len(v))
    assert abs(acc - expect) < 5 * sd + 2e-3
    # accepted centres are denser near agglomerates
    d = r[:, None, :3].astype(np.float64) - c[None]
    d -= 20 * np.round(d / 20)
    near = ((d ** 2).sum(-1).min(1) < 1.5 ** 2).mean()
    frac_near = (d2 < 1.5 ** 2).mean()
    assert near > 1.5 * frac_near
