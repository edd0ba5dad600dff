"""Oracle generation: an independent Python restatement of the round rule
(brute force over all pairs and periodic images, no cell list) must give the
oracle's packing bit-for-bit; the packing has no overlapping pair; results do
not depend on the thread count; termination rules."""
import itertools
import math

import numpy as np
import pytest

import nanofiller_amd as nf
import oracle


def params(**kw):
    base = dict(box=(2.5, 2.5, 2.5), n_rods=300, len=1.0, radius=0.005, seed=3, contact_gap=0.002,
                precision=nf.NF_PREC_FP64, min_batch=64, max_batch=400)
    base.update(kw)
    return nf.normalize(nf.default_params(**base))


def images(box, reach):
    k = [int(math.floor(reach / b)) + 1 for b in box]
    return [np.array(v, np.float64) * np.array(box) for v in
            itertools.product(*[range(-kk, kk + 1) for kk in k])]


def py_generate(p):
    """The round rule restated with brute-force neighbour enumeration."""
    N = p.n_rods
    agg = oracle.agg_centres(p) if p.den_law == nf.NF_DEN_AGGLOM else None
    hl_max = 0.5 * (p.len_max if p.len_law == nf.NF_LEN_LOGNORMAL else p.len)
    reach = 2 * hl_max * 1.000001 + 2 * p.radius + p.sep + 1e-6
    imgs = images(p.box, reach)
    acc = np.zeros((0, 8), np.float32)
    n_acc, rnd, stall, mprev, aprev = 0, 0, 0, 0, 0
    tests = 0
    status = None

    def clamp(m):
        return int(min(max(m, p.min_batch), p.max_batch))

    while status is None:
        D = N - n_acc
        if rnd == 0:
            M = clamp(D + D // 4)
        elif aprev == 0:
            M = p.max_batch
        else:
            M = clamp(-(-(D * mprev * 5) // (aprev * 4)))
        props = []
        for s in range(M):
            v, r = oracle.candidate(p, rnd, s, agg)
            if v == 1:
                props.append((s, r))
        slots = np.array([s for s, _ in props], np.int64)
        R = np.array([r for _, r in props], np.float32).reshape(-1, 8)
        P = len(props)
        dead1 = np.zeros(P, bool)
        if n_acc and P:
            a = np.repeat(R, len(acc), axis=0)
            b = np.tile(acc, (P, 1))
            for sh in imgs:
                shv = np.tile(sh, (len(a), 1))
                _, w, cut = oracle.pair_batch(a, b, shv, False, p.sep, p.soft_frac)
                tests += int(w.sum())
                dead1 |= cut.reshape(P, len(acc)).any(axis=1)
        alive = np.where(~dead1)[0]
        dead2 = np.zeros(P, bool)
        if len(alive) > 1:
            ii, jj = np.meshgrid(alive, alive, indexing="ij")
            m = slots[jj] < slots[ii]
            ii, jj = ii[m], jj[m]
            for sh in imgs:
                shv = np.tile(sh, (len(ii), 1))
                _, w, cut = oracle.pair_batch(R[ii], R[jj], shv, False, p.sep, p.soft_frac)
                tests += int(w.sum())
                np.logical_or.at(dead2, ii[cut], True)
        surv = [k for k in range(P) if not dead1[k] and not dead2[k]]
        take = surv[:D]
        if take:
            acc = np.concatenate([acc, R[take]])
        a_r = len(take)
        n_acc += a_r
        mprev, aprev = M, a_r
        rnd += 1
        stall = stall + 1 if a_r == 0 else 0
        if n_acc == N:
            status = nf.NF_STATUS_TARGET
        elif stall >= p.stall_rounds:
            status = nf.NF_STATUS_STALLED
        elif rnd >= p.max_rounds:
            status = nf.NF_STATUS_MAX_ROUNDS
    return acc, rnd, tests


@pytest.mark.parametrize("kw", [
    {},
    dict(len_law=nf.NF_LEN_LOGNORMAL, len_sigma=0.3, len_min=0.3, len_max=1.5, n_rods=250,
         ort_law=nf.NF_ORT_VMF, kappa=5.0, axis=(0, 0, 1), seed=9),
    dict(den_law=nf.NF_DEN_AGGLOM, n_agg=3, agg_sigma=0.5, agg_peak=3.0, n_rods=200, seed=4),
])
def test_round_rule_restated(kw):
    p = params(**kw)
    rods, st, _ = oracle.generate(p, threads=4)
    ref, rounds, tests = py_generate(p)
    assert st["rounds"] == rounds
    assert st["pair_tests"] == tests
    np.testing.assert_array_equal(rods.view(np.uint32), ref.view(np.uint32))


def test_no_overlapping_pair_brute_force():
    p = params(box=(3, 3, 3), n_rods=2000, seed=5, max_batch=0)
    p = nf.normalize(p)
    rods, st, _ = oracle.generate(p, threads=8)
    assert st["n_accepted"] == 2000
    n = len(rods)
    ii, jj = np.triu_indices(n, 1)
    reach = 1.0 * 1.000001 + 0.01 + 1e-6
    worst = np.inf
    for sh in images(p.box, reach):
        shv = np.tile(sh, (len(ii), 1))
        d2, w, cut = oracle.pair_batch(rods[ii], rods[jj], shv, False, 0.0)
        assert not cut.any()
        if w.any():
            worst = min(worst, d2[w].min())
    assert worst >= 0.01 ** 2 * (1 - 1e-12)


def test_thread_count_does_not_change_packing():
    p = params(box=(3, 3, 3), n_rods=1500, seed=6, max_batch=0)
    p = nf.normalize(p)
    r1, s1, _ = oracle.generate(p, threads=1)
    r8, s8, _ = oracle.generate(p, threads=8)
    assert s1 == s8
    np.testing.assert_array_equal(r1.view(np.uint32), r8.view(np.uint32))


def test_unreachable_target_stalls():
    # far beyond jamming for l/d = 100 in a small box
    p = params(box=(1.2, 1.2, 1.2), n_rods=20000, seed=7, stall_rounds=3, max_rounds=200,
               max_batch=2000)
    _, st, acc = oracle.generate(p, threads=8)
    assert st["status"] == nf.NF_STATUS_STALLED
    assert st["n_accepted"] < 20000
    assert list(acc[-3:]) == [0, 0, 0]


def test_round_cap():
    p = params(box=(3, 3, 3), n_rods=2000, seed=8, max_rounds=2, max_batch=100, min_batch=100)
    _, st, _ = oracle.generate(p, threads=4)
    assert st["rounds"] == 2 and st["status"] == nf.NF_STATUS_MAX_ROUNDS
