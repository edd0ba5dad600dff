"""Dead-end pruning (reduce, src/conductivity.cu:79-138) in the oracle:
hand-built contact graphs with known survivors and round counts, and a
random packing against an independent one-rod-at-a-time peeling (the kept
set of simultaneous rounds is the unique largest set in which every rod has
min_cont contacts, so any removal order reaches it)."""
from collections import deque

import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

BOX = (4.0, 2.0, 2.0)


def _params():
    return nf.normalize(nf.default_params(box=BOX, n_rods=64, len=0.5, radius=0.005,
                                          contact_gap=0.002, elec_axis=0))


def _rods(n, source=(), sink=()):
    """n short rods along y in the middle of the box (no coupling), those in
    `source` / `sink` moved to touch the x = 0 / x = 4 plane (0.001 gap)."""
    r = np.zeros((n, 8), np.float32)
    r[:, 0] = 2.0
    r[:, 1] = 1.0
    r[:, 2] = 1.0
    r[:, 4] = 1.0
    r[:, 6] = 0.5
    r[:, 7] = 0.005
    for i in source:
        r[i, 0] = 0.006
    for i in sink:
        r[i, 0] = 3.994
    return r


def _img(x=0, y=0, z=0):
    return (x & 0xFF) | ((y & 0xFF) << 8) | ((z & 0xFF) << 16)


def _run(edges, n, source=(), sink=(), min_cont=2):
    p = _params()
    I = np.array([e[0] for e in edges], np.uint32)
    J = np.array([e[1] for e in edges], np.uint32)
    im = np.array([e[2] if len(e) > 2 else 0 for e in edges], np.int32)
    keep, rounds = oracle.reduce(p, _rods(n, source, sink), I, J, im, p.contact_gap, min_cont)
    return set(np.flatnonzero(keep).tolist()), rounds


def test_chain_between_the_electrodes_is_kept():
    kept, rounds = _run([(0, 1), (1, 2), (2, 3), (3, 4)], 5, source=[0], sink=[4])
    assert kept == {0, 1, 2, 3, 4} and rounds == 1


def test_dangling_branch_is_peeled_tip_first():
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (2, 5), (5, 6)]
    kept, rounds = _run(edges, 7, source=[0], sink=[4])
    assert kept == {0, 1, 2, 3, 4} and rounds == 3  # 6, then 5, then nothing


def test_loops_stay_and_open_chains_go():
    edges = [(0, 1), (1, 2), (0, 2),      # floating triangle: every rod has 2
             (4, 5), (5, 6),              # open chain, no electrode
             ]
    kept, rounds = _run(edges, 8)         # 3 and 7 are isolated
    assert kept == {0, 1, 2} and rounds == 3  # {3, 4, 6, 7}, then {5}, then nothing


def test_crossing_contacts_do_not_count_and_images_do():
    edges = [(0, 1, _img(x=1)),                  # crosses the electrode axis
             (2, 3, _img()), (2, 3, _img(y=1))]  # one pair, two images: a loop
    kept, _ = _run(edges, 4, source=[0, 1])
    assert kept == {2, 3}


@pytest.mark.parametrize("min_cont,expect", [(0, set(range(7))), (1, {0, 1, 2, 3, 4, 5, 6}),
                                             (3, set())])
def test_threshold(min_cont, expect):
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (2, 5), (5, 6)]
    kept, _ = _run(edges, 7, source=[0], sink=[4], min_cont=min_cont)
    assert kept == expect


def _peel_one_at_a_time(n, edges, coup, min_cont, seed):
    deg = coup.astype(np.int64).copy()
    adj = [[] for _ in range(n)]
    for a, b in edges:
        deg[a] += 1
        deg[b] += 1
        adj[a].append(b)
        adj[b].append(a)
    alive = np.ones(n, bool)
    order = np.random.default_rng(seed).permutation(n)
    q = deque(int(i) for i in order if deg[i] < min_cont)
    while q:
        i = q.popleft()
        if not alive[i]:
            continue
        alive[i] = False
        for j in adj[i]:
            if alive[j]:
                deg[j] -= 1
                if deg[j] < min_cont:
                    q.append(j)
    return alive


@pytest.mark.parametrize("axis", [0, 2])
def test_random_packing_against_sequential_peeling(axis):
    p = cases.small_iso(nf.NF_PREC_FP64, n_rods=2500, contact_gap=0.03, elec_axis=axis)
    rods, _, _ = oracle.generate(p, threads=8)
    c = oracle.contacts(p, rods, p.contact_gap, threads=8)
    keep, rounds = oracle.reduce(p, rods, c["I"], c["J"], c["image"], p.contact_gap, 2)
    imgax = (c["image"] >> (8 * axis)) & 0xFF
    ok = imgax == 0
    edges = list(zip(c["I"][ok].tolist(), c["J"][ok].tolist()))
    L = p.box[axis]
    hl = rods[:, 6].astype(np.float64) * 0.5
    ext = np.abs(hl * rods[:, 3 + axis].astype(np.float64))
    pos = rods[:, axis].astype(np.float64)
    rad = rods[:, 7].astype(np.float64)
    coup = (((pos - ext) - rad) <= p.contact_gap).astype(int) + \
           (((L - (pos + ext)) - rad) <= p.contact_gap).astype(int)
    for seed in (0, 1):
        alive = _peel_one_at_a_time(len(rods), edges, coup, 2, seed)
        np.testing.assert_array_equal(alive, keep)
    # the kept set is a fixed point: every kept rod has 2 contacts among kept rods
    deg = coup.copy()
    for a, b in edges:
        if keep[a] and keep[b]:
            deg[a] += 1
            deg[b] += 1
    assert np.all(deg[keep] >= 2)
    assert 0 < keep.sum() < len(rods) and rounds >= 2
