"""Bin probabilities of the target laws of a packing and the chi-square test
of a histogram against them.

The generator samples candidate centres from the density law (von Neumann
acceptance) and directions from the orientation law; a packing then keeps
the candidates that do not overlap.  These helpers give the expected share
of each bin under the target law itself:

- centre density along one axis: uniform, or the agglomerate law
  rho(x) = 1 + (peak - 1) exp(-d^2 / (2 sigma^2)) (d the periodic distance to
  the nearest centre, nf_agglomerate_centres) integrated over the other two
  axes by the midpoint rule;
- polar angle to an axis over [0, pi]: isotropic (uniform in cos theta) or
  von Mises-Fisher about that axis, F(c) = (e^{kappa (c - 1)} - e^{-2 kappa})
  / (1 - e^{-2 kappa}).

chi2_test merges neighbouring bins until each expects at least 5 counts.
"""
from __future__ import annotations

import math

import numpy as np

from . import NF_DEN_AGGLOM, NF_DEN_UNIFORM, NF_ORT_ISO, NF_ORT_VMF, agglomerate_centres


def density_bins(p, axis: int, bins: int, sub: int = 8, grid: int = 96) -> np.ndarray:
    """Expected share of the centres in each of `bins` slabs along `axis`."""
    if p.den_law == NF_DEN_UNIFORM:
        return np.full(bins, 1.0 / bins)
    if p.den_law != NF_DEN_AGGLOM:
        raise ValueError("density_bins covers the uniform and agglomerate laws")
    box = np.array([p.box[0], p.box[1], p.box[2]], np.float64)
    cen = agglomerate_centres(p)
    others = [a for a in range(3) if a != axis]
    u = (np.arange(bins * sub) + 0.5) / (bins * sub) * box[axis]
    v = (np.arange(grid) + 0.5) / grid * box[others[0]]
    w = (np.arange(grid) + 0.5) / grid * box[others[1]]
    inv2s2 = 1.0 / (2.0 * p.agg_sigma * p.agg_sigma)
    mean = np.empty(bins * sub)
    V, W = np.meshgrid(v, w, indexing="ij")
    for i, x in enumerate(u):
        d2 = np.full(V.shape, np.inf)
        for c in cen:
            dx = x - c[axis]
            dx -= box[axis] * np.round(dx / box[axis])
            dy = V - c[others[0]]
            dy -= box[others[0]] * np.round(dy / box[others[0]])
            dz = W - c[others[1]]
            dz -= box[others[1]] * np.round(dz / box[others[1]])
            d2 = np.minimum(d2, dx * dx + dy * dy + dz * dz)
        mean[i] = np.mean(1.0 + (p.agg_peak - 1.0) * np.exp(-d2 * inv2s2))
    per_bin = mean.reshape(bins, sub).mean(axis=1)
    return per_bin / per_bin.sum()


def polar_bins(p, axis, bins: int) -> np.ndarray:
    """Expected share of the directions in each polar-angle bin of [0, pi]
    about `axis` (for vMF, `axis` must be the law's mean direction)."""
    c = np.cos(np.arange(bins + 1) * math.pi / bins)
    if p.ort_law == NF_ORT_ISO:
        F = (c + 1.0) / 2.0
    elif p.ort_law == NF_ORT_VMF:
        mu = np.array([p.axis[0], p.axis[1], p.axis[2]], np.float64)
        ax = np.asarray(axis, np.float64)
        if abs(float(np.dot(mu / np.linalg.norm(mu), ax / np.linalg.norm(ax))) - 1.0) > 1e-12:
            raise ValueError("the polar axis must be the vMF mean direction")
        k = p.kappa
        F = (np.exp(k * (c - 1.0)) - math.exp(-2.0 * k)) / (1.0 - math.exp(-2.0 * k))
    else:
        raise ValueError("unknown orientation law")
    return F[:-1] - F[1:]


def chi2_test(counts, probs, min_expected: float = 5.0) -> dict:
    """Pearson chi-square of counts against bin probabilities, neighbouring
    bins merged until each expects min_expected counts; `tv` is the total
    variation distance between the observed bin frequencies and the law
    (half the L1 distance, unmerged bins), an effect size that, unlike p,
    does not grow with the sample: with millions of rods a bias of a few
    tenths of a percent already gives p near 0."""
    from scipy.stats import chi2

    o = np.asarray(counts, np.float64)
    e = np.asarray(probs, np.float64) * o.sum()
    mo, me = [], []
    acc_o = acc_e = 0.0
    for oi, ei in zip(o, e):
        acc_o += oi
        acc_e += ei
        if acc_e >= min_expected:
            mo.append(acc_o)
            me.append(acc_e)
            acc_o = acc_e = 0.0
    if acc_e > 0.0 and me:
        mo[-1] += acc_o
        me[-1] += acc_e
    mo, me = np.array(mo), np.array(me)
    stat = float(np.sum((mo - me) ** 2 / me))
    dof = len(me) - 1
    n = o.sum()
    tv = float(0.5 * np.sum(np.abs(o / n - np.asarray(probs, np.float64)))) if n > 0 else 0.0
    return {"chi2": stat, "dof": dof, "p": float(chi2.sf(stat, dof)) if dof > 0 else 1.0,
            "n": int(n), "tv": tv}


def packing_tests(p, rods: np.ndarray, axis: int, pol_axis, bins: int = 64) -> dict:
    """chi-square of a packing's centre profile along `axis` and of its polar
    angles about `pol_axis` against the target laws of p."""
    L = p.box[axis]
    x = rods[:, axis].astype(np.float64)
    hc = np.bincount(np.minimum((x / L * bins).astype(np.int64), bins - 1), minlength=bins)
    ax = np.asarray(pol_axis, np.float64)
    ax = ax / np.linalg.norm(ax)
    ct = np.clip(rods[:, 3:6].astype(np.float64) @ ax, -1.0, 1.0)
    th = np.arccos(ct)
    hp = np.bincount(np.minimum((th / math.pi * bins).astype(np.int64), bins - 1), minlength=bins)
    return {"density": chi2_test(hc, density_bins(p, axis, bins)),
            "polar": chi2_test(hp, polar_bins(p, ax, bins))}
