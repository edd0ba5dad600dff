"""Contact area model (surf_to_cylinder_capped, inc/intersect/
intersect_cylinder_capped.h:204-293) in the oracle, pinned by closed forms of
each branch, a brute-force scan of the patch radius, and translation
invariance (the reference's absolute-position term is not reproduced)."""
import math

import numpy as np
import pytest

import oracle

A = 0.005     # rod radius (um)
AF = float(np.float32(A))  # as stored in the rod record
RLIM = 0.002  # contact gap


def rod(centre, direction, length, radius=A):
    c = np.asarray(direction, np.float64)
    c = c / np.linalg.norm(c)
    return np.array([*centre, *c, length, radius], np.float32)


def patch_radius(D, a_self, a_other, rlim=RLIM):
    h = max(D - a_other - rlim, 0.0)
    return a_self * math.sqrt(1.0 - (h / a_self) ** 2) if h <= a_self else 0.0


def area(r1, r2, shift=(0.0, 0.0, 0.0), rlim=RLIM):
    return oracle.contact_area(r1, r2, np.asarray(shift, np.float64), rlim)


@pytest.mark.parametrize("gap", [0.0, 0.0005, 0.001, 0.0019])
def test_perpendicular_crossing_is_the_ellipse(gap):
    """Long rods crossing at 90 degrees at their centres: semi-axes dr and
    dr / sin 45, S = pi/2 dr (A+ + A-) = sqrt(2) pi dr^2."""
    D = 2 * A + gap
    r1 = rod((5.0, 5.0, 5.0), (1, 0, 0), 1.0)
    r2 = rod((5.0, 5.0, 5.0 + D), (0, 1, 0), 1.0)
    dr = patch_radius(float(np.float32(5.0 + D)) - 5.0, AF, AF)
    assert area(r1, r2) == pytest.approx(math.sqrt(2.0) * math.pi * dr * dr, rel=1e-12)


@pytest.mark.parametrize("gamma_deg", [20.0, 45.0, 70.0, 120.0, 160.0])
def test_angled_crossing(gamma_deg):
    """Crossing at angle gamma (the model folds c2 onto the acute side):
    A+- = dr / sin(gamma'/2), gamma' = min(gamma, 180 - gamma)."""
    g = math.radians(gamma_deg)
    D = 2 * A + 0.001
    r1 = rod((0.0, 0.0, 0.0), (1, 0, 0), 2.0)
    r2 = rod((0.0, 0.0, D), (math.cos(g), math.sin(g), 0), 2.0)
    dr = patch_radius(float(np.float32(D)), AF, AF)
    gp = min(g, math.pi - g)
    expect = 0.5 * math.pi * dr * (2 * dr / math.sin(gp / 2))
    assert area(r1, r2) == pytest.approx(expect, rel=1e-6)


def test_crossing_near_a_rod_end():
    """Rod 2 crosses rod 1 close to rod 1's +end: A+ is the distance along the
    bisector to that end, (l1/2 - t1) cos(gamma/2); A- the strip extent."""
    D = 2 * A + 0.0005
    t1 = 0.5 - 2.5 * A  # just inside l1/2 - a1
    g = math.radians(30.0)
    r1 = rod((0.0, 0.0, 0.0), (1, 0, 0), 1.0)
    r2 = rod((t1, 0.0, D), (math.cos(g), math.sin(g), 0), 1.0)
    dr = patch_radius(float(np.float32(D)), AF, AF)
    t1f = float(np.float32(t1))
    ap = (0.5 - t1f) * math.cos(g / 2)
    am = dr / math.sin(g / 2)
    assert ap < am  # the end limits the + side
    assert area(r1, r2) == pytest.approx(0.5 * math.pi * dr * (ap + am), rel=1e-6)


@pytest.mark.parametrize("offset", [0.0, 0.2, 0.7])
@pytest.mark.parametrize("anti", [False, True])
def test_parallel_overlap_is_the_strip(offset, anti):
    """Side-by-side parallel (or antiparallel) rods overlapping over
    l - offset: S = 2 dr (l - offset)."""
    D = 2 * A + 0.001
    r1 = rod((1.0, 1.0, 1.0), (1, 0, 0), 1.0)
    r2 = rod((1.0 + offset, 1.0 + D, 1.0), (-1 if anti else 1, 0, 0), 1.0)
    dy = float(np.float32(1.0 + D)) - 1.0
    dr = patch_radius(dy, AF, AF)
    ov = 1.0 - (float(np.float32(1.0 + offset)) - 1.0)
    assert area(r1, r2) == pytest.approx(max(2 * dr * ov, math.pi * dr * dr), rel=1e-9)


def test_end_cap_t_contact_is_the_full_disc():
    """Rod 2 ends on rod 1's side, axes meeting: the line distance is 0, so
    the patch is the whole radius: S = pi min(a1, a2)^2."""
    a2 = 0.004
    r1 = rod((0.0, 0.0, 0.0), (1, 0, 0), 1.0)
    r2 = rod((0.0, 0.0, 0.5 + A + a2 + 0.001), (0, 0, 1), 1.0, radius=a2)
    a2f = float(np.float32(a2))
    assert area(r1, r2) == pytest.approx(math.pi * a2f * a2f, rel=1e-12)


def test_separated_beyond_reach_has_no_area():
    r1 = rod((0.0, 0.0, 0.0), (1, 0, 0), 1.0)
    r2 = rod((0.0, 0.0, 2 * A + RLIM + 2 * A + 1e-3), (0, 1, 0), 1.0)
    assert area(r1, r2) == 0.0


def test_patch_radius_by_brute_force():
    """dr1 is the half-width of rod 1's surface strip within rlim of rod 2's
    surface, for rods crossing at 90 degrees (scan of the surface at x = 0)."""
    for D in (2 * A, 2 * A + 0.0007, 2 * A + 0.0016):
        ys = np.linspace(0.0, A, 200001)
        z = np.sqrt(np.maximum(A * A - ys * ys, 0.0))   # rod 1's surface (axis x)
        dist = (D - z) - A                              # to rod 2's surface (axis y at height D)
        y_max = ys[dist <= RLIM].max()
        assert y_max == pytest.approx(patch_radius(D, A, A), abs=2 * A / 200000)


def test_translation_invariance():
    """Moving both rods by an exactly representable offset, or through a
    periodic image, leaves S unchanged bit for bit."""
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(400):
        c1 = rng.uniform(1, 2, 3)
        d1, d2 = rng.normal(size=3), rng.normal(size=3)
        c2 = c1 + rng.normal(size=3) * 0.006
        r1, r2 = rod(c1, d1, rng.uniform(0.3, 1.5)), rod(c2, d2, rng.uniform(0.3, 1.5))
        s0 = area(r1, r2, rlim=0.004)
        for off in (4.0, 8.0, -1.0):
            q1, q2 = r1.copy(), r2.copy()
            q1[:3] += np.float32(off)
            q2[:3] += np.float32(off)
            exact = (np.all(q1[:3].astype(np.float64) - r1[:3] == off)
                     and np.all(q2[:3].astype(np.float64) - r2[:3] == off))
            if exact:
                assert area(q1, q2, rlim=0.004) == s0
        # rod 2 one box away, brought back by its image shift
        q2 = r2.copy()
        q2[0] += np.float32(8.0)
        if float(q2[0]) - float(r2[0]) == 8.0:
            assert area(r1, q2, shift=(-8.0, 0.0, 0.0), rlim=0.004) == s0
            checked += 1
    assert checked > 20
