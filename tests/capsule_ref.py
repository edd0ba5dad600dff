"""Independent segment-distance references for the tests.

brute_d2 minimises F(s, t) = |w + s h1 - t h2|^2 over [-1, 1]^2 without any
of the oracle's case logic: for a grid of s values it takes the exact
clamped optimum in t, then refines s by golden-section search (the partial
minimum over t is convex in s).
"""
import numpy as np


def _best_t(w, h1, h2, s):
    e = h2 @ h2
    v = w[None, :] + s[:, None] * h1[None, :]
    t = np.clip((v @ h2) / e, -1.0, 1.0)
    r = v - t[:, None] * h2[None, :]
    return np.einsum("ij,ij->i", r, r)


def brute_d2(r1, r2, shift=(0.0, 0.0, 0.0)):
    """r = (x, y, z, cx, cy, cz, l, a) as float32 values; rod 2 displaced by shift."""
    r1 = np.asarray(r1, np.float64)
    r2 = np.asarray(r2, np.float64)
    w = (r1[:3] - r2[:3]) - np.asarray(shift, np.float64)
    h1 = 0.5 * r1[6] * r1[3:6]
    h2 = 0.5 * r2[6] * r2[3:6]
    s = np.linspace(-1.0, 1.0, 4001)
    f = _best_t(w, h1, h2, s)
    k = int(np.argmin(f))
    lo, hi = s[max(k - 1, 0)], s[min(k + 1, len(s) - 1)]
    g = 0.5 * (np.sqrt(5.0) - 1.0)
    a, b = lo, hi
    for _ in range(100):
        c, d = b - g * (b - a), a + g * (b - a)
        fc, fd = _best_t(w, h1, h2, np.array([c, d]))
        if fc < fd:
            b = d
        else:
            a = c
    return float(min(f.min(), _best_t(w, h1, h2, np.array([0.5 * (a + b)]))[0]))


def rod(c, u, length, a=0.005):
    u = np.asarray(u, np.float64)
    u = u / np.linalg.norm(u)
    return np.array([c[0], c[1], c[2], u[0], u[1], u[2], length, a], np.float32)
