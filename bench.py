"""Benchmark: accepted rods/s and pair-tests/s of the MI355X generator.

Headline workload (one GPU): 500,000 rods in a 20 um periodic cube,
log-normal lengths (median 1 um, sigma 0.3, clipped to [0.3, 3] um), radius
5 nm, density = background + 8 Gaussian agglomerates (sigma 1.5 um, peak 3x),
vMF(+z, kappa 5) orientations, hard core 10 nm, Philox seed 2.  A step is one
complete generation of that packing (every insert/reject round, from the
first candidate to the final packing resident in HBM) by the FP32 path with
FP64 re-check band; `value` is accepted rods/s over the timed steps.

The same JSON line carries:
  * fp64_plain: the same packing by the plain FP64 path; identity of the two
    packings and of their contact lists;
  * roofline: the reject-vs-accepted kernel, algorithmic tile bytes per launch
    over its HIP-event time per launch (one extra, profiled generation);
  * batch: 256 independent packings of uneven size (2k-200k rods) in one
    batched FP32 + recheck generation, against the same 256 packings made one
    at a time by the plain FP64 path, with identical packings;
  * contacts: 4,000,000 rods in a 50 um cube: contact list by the footprint
    engine (FP32 + recheck, and FP64) and by the cell-pair engine, three
    64-bin int64 histograms, contact-kernel HBM and VALU-issue rooflines;
  * network: 8,000,000 rods in a 40 um cube: contacts, components and the
    FP64 Kirchhoff matrix;
  * cpu_baseline: the C oracle on a bounded sample of the headline workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs main,batch,contacts,network,conduct]
(legs prof_main and prof_contacts run only the kernels a counter profile is
taken on: one headline generation; one 4M generation and one contact pass
per engine)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import nanofiller_amd as nf  # noqa: E402
from nanofiller_amd import configs  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "accepted rods/s & pair-tests/s, FP32+recheck vs FP64, identical contact lists"
K_OLD = nf.KERNEL_CLASSES.index("reject_old")
K_NEW = nf.KERNEL_CLASSES.index("reject_new")


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_pair_generate_cfg2_500k_hbm_counters.md")
VALU_SUMMARY = os.path.join(ROOT, "profiles", "pair_kernels_valu_counters.md")
# VALU issue roof: a wave64 VALU instruction occupies a 16-lane SIMD for 4
# cycles; 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9


def profiled_traffic():
    """HBM bytes per launch of the reject-vs-accepted kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated
    against a plain copy kernel), or None when the summary is absent."""
    import re

    try:
        with open(TRAFFIC_SUMMARY) as f:
            m = re.search(r"traffic_bytes_per_launch k_pair<0, 1>: (\d+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


CONTACT_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "contacts_4M_hbm_counters.md")


def profiled_contact_traffic():
    """HBM bytes of one 4M-rod footprint contact pass from the committed
    FETCH_SIZE / WRITE_SIZE summary (calibrated), or None when absent."""
    import re

    try:
        with open(CONTACT_TRAFFIC_SUMMARY) as f:
            m = re.search(r"traffic_bytes_per_pass k_seg_cross<1>: (\d+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def profiled_valu(key: str):
    """SQ_INSTS_VALU per launch of one pair kernel on one workload from the
    committed rocprofv3 counter summary (lines 'valu_per_launch <key>: N'),
    or None when absent."""
    import re

    try:
        with open(VALU_SUMMARY) as f:
            m = re.search(r"valu_per_launch " + re.escape(key) + r": (\d+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def valu_roof(key: str, ms_per_launch: float):
    """VALU-issue roofline of a kernel: its counted wave instructions at 4
    cycles each spread over every SIMD, against its measured time."""
    n = profiled_valu(key)
    if n is None or ms_per_launch <= 0:
        return None
    issue_ms = n * VALU_CYCLES / SIMDS / CLOCK_HZ * 1e3
    return {"bound": "valu", "instructions_per_launch": n, "issue_ms": issue_ms,
            "launch_ms": ms_per_launch, "frac": issue_ms / ms_per_launch,
            "source": os.path.relpath(VALU_SUMMARY, ROOT)}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_generation(p, steps: int, warmup: int, options=None):
    with nf.Packing(p, options or {}) as pk:
        for i in range(warmup):
            pk.generate()
        sts = []
        t0 = time.perf_counter()
        for i in range(steps):
            sts.append(pk.generate()[0])
        dt = time.perf_counter() - t0
        rods = pk.rods()
        cont = pk.contacts(p.contact_gap)
    return dt, sts, rods, cont


def profiled_generation(p, tiles=None) -> dict:
    """One generation with HIP events per kernel class: the two pair passes'
    times and streamed bytes (probes and neighbour tiles)."""
    with nf.Packing(p, {nf.NF_OPT_PROFILE: 1, **(tiles or {})}) as pk:
        sp = pk.generate()[0]
    return {"wall_ms": sp.wall_ms, "rounds": sp.rounds, "launches": sp.rounds_launched,
            "kernel_ms": {k: sp.kern_ms[i] for i, k in enumerate(nf.KERNEL_CLASSES)},
            "reject_old_tile_bytes": sp.tile_bytes, "reject_new_tile_bytes": sp.tile_bytes_new,
            "pair_tests": sp.pair_tests}


def same_contacts(a, b) -> bool:
    return (np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col)
            and np.array_equal(a.image, b.image) and np.array_equal(a.gap.view(np.uint32),
                                                                      b.gap.view(np.uint32)))


def leg_main(a, out: dict) -> None:
    p32 = configs.cfg2(nf.NF_PREC_FP32_RECHECK)
    tiles = configs.tile_options("500k")  # launch shape picked by nanofiller_amd/autotune.py
    log(f"headline: FP32 + FP64 re-check (tile options {tiles})")
    dt, sts, rods32, c32 = timed_generation(p32, a.steps, a.warmup, tiles)
    st = sts[-1]
    rods_s = st.n_accepted * a.steps / dt
    ms_step = dt / a.steps * 1e3
    log(f"FP32: {rods_s:.4e} rods/s, {ms_step:.2f} ms/step, {st.rounds} rounds")
    # one profiled generation for the per-kernel times and the roofline
    with nf.Packing(p32, {nf.NF_OPT_PROFILE: 1, **tiles}) as pk:
        sp = pk.generate()[0]
    launches = max(sp.rounds_launched, 1)
    k_old = sp.kern_ms[K_OLD]
    achieved = (sp.tile_bytes / launches) / (k_old / launches * 1e-3) / 1e9 if k_old > 0 else 0.0
    log(f"profiled: reject_old {k_old:.3f} ms over {launches} launches, "
        f"reject_new {sp.kern_ms[K_NEW]:.3f} ms")
    out.update(value=rods_s, ms_per_step=ms_step, n_rods=p32.n_rods)
    out["roofline"] = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(),
        "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT),
        "kernel": "k_pair<reject vs accepted, FP32+recheck>",
        "bytes_per_launch": sp.tile_bytes / launches, "avg_launch_ms": k_old / launches,
        "launches": launches, "valu": valu_roof("k_pair<0, 1> 500k", k_old / launches),
    }
    out["extra"]["fp32_recheck"] = {
        "tile_options": tiles,
        "accepted_rods_per_s": rods_s, "ms_per_step": ms_step,
        "rounds": st.rounds, "candidates": st.candidates, "proposals": st.proposals,
        "pair_tests_evaluated": st.pair_tests, "rechecks": st.rechecks,
        "recheck_fraction": st.rechecks / max(st.pair_tests, 1), "status": nf.STATUS[st.status],
        "profiled_kernel_ms": {k: sp.kern_ms[i] for i, k in enumerate(nf.KERNEL_CLASSES)},
        "profiled_wall_ms": sp.wall_ms,
    }
    # FP32-vs-FP64 distance error: one generation that also evaluates every
    # FP32-decided pair in FP64 (NF_OPT_DIAG) and tests every pair within
    # reach (NF_OPT_TEST_ALL_PAIRS), whose count is the oracle's pair-test
    # count; its packing must not change
    with nf.Packing(p32, {nf.NF_OPT_DIAG: 1, nf.NF_OPT_TEST_ALL_PAIRS: 1, **tiles}) as pk:
        sd = pk.generate()[0]
        rods_diag = pk.rods()
    # pair-tests/s in the oracle's unit: pairs within reach, over the timed steps
    pair_s = sd.pair_tests * a.steps / dt
    out["pair_s"] = pair_s
    out["extra"]["fp32_recheck"]["pair_tests"] = sd.pair_tests
    out["extra"]["fp32_recheck"]["pair_tests_per_s"] = pair_s
    out["extra"]["fp32_recheck"]["max_abs_d32_d64_um"] = sd.max_abs_d32_d64
    out["extra"]["fp32_recheck"]["diag_identical_packing"] = bool(
        rods_diag.shape == rods32.shape and np.array_equal(rods_diag.view(np.uint32), rods32.view(np.uint32)))
    log(f"diag: max |d32 - d64| = {sd.max_abs_d32_d64:.3e} um")
    out["extra"]["chi2"] = {"500k_lognormal_agglomerates_vmf_z5": chi2_entry(p32, rods32, 0, (0, 0, 1))}
    leg_cfg1(a, out)
    if a.no_fp64:
        return
    p64 = configs.cfg2(nf.NF_PREC_FP64)
    log("headline: plain FP64")
    steps64 = max(1, min(a.steps, 3))
    dt64, sts64, rods64, c64 = timed_generation(p64, steps64, 1, tiles)
    rods64_s = sts64[-1].n_accepted * steps64 / dt64
    same_rods = rods64.shape == rods32.shape and np.array_equal(rods64.view(np.uint32),
                                                                  rods32.view(np.uint32))
    same_c = same_contacts(c64, c32)
    out["extra"]["fp64_plain"] = {
        "accepted_rods_per_s": rods64_s,
        "pair_tests_per_s": sum(s.pair_tests for s in sts64) / dt64,
        "ms_per_step": dt64 / steps64 * 1e3,
        "speedup_fp32_recheck_over_fp64": rods_s / rods64_s,
        "identical_packing": bool(same_rods), "identical_contact_lists": bool(same_c),
        "contacts": int(c32.n),
    }
    log(f"FP64: {rods64_s:.4e} rods/s; identical packing {same_rods}, contacts {same_c}")


def chi2_entry(p, rods, axis, pol_axis) -> dict:
    """chi-square of the packing's centre profile along `axis` and polar
    angles about `pol_axis` against the target laws (64 bins)."""
    from nanofiller_amd import targets

    t = targets.packing_tests(p, rods, axis, pol_axis, bins=64)
    return {"rods": int(len(rods)), "density_axis": "xyz"[axis],
            "density": t["density"], "polar_axis": list(pol_axis), "polar": t["polar"]}


def leg_cfg1(a, out: dict) -> None:
    """20,000 rods, 5 um box, 1 um cells, seed 1: the plain FP64 GPU path and
    the FP32 + recheck path on the same packing."""
    res = {}
    rods = {}
    for name, prec in (("fp64_plain", nf.NF_PREC_FP64), ("fp32_recheck", nf.NF_PREC_FP32_RECHECK)):
        p = configs.cfg1(prec)
        dt, sts, r, c = timed_generation(p, a.steps, a.warmup, configs.tile_options("20k"))
        rods[name] = (r, c)
        res[name] = {"accepted_rods_per_s": sts[-1].n_accepted * a.steps / dt,
                     "pair_tests_per_s": sum(s.pair_tests for s in sts) / dt,
                     "ms_per_step": dt / a.steps * 1e3, "rounds": sts[-1].rounds,
                     "pair_tests": sts[-1].pair_tests, "contacts": int(c.n)}
    (r64, c64), (r32, c32) = rods["fp64_plain"], rods["fp32_recheck"]
    res["identical_packing"] = bool(np.array_equal(r64.view(np.uint32), r32.view(np.uint32)))
    res["identical_contact_lists"] = bool(same_contacts(c64, c32))
    res["workload"] = ("20,000 rods, l 1 um, d 10 nm, 5 um periodic cube, uniform, isotropic, seed 1, "
                       "1 um cells, contacts <= 2 nm")
    out["extra"]["cfg1"] = res
    out["extra"].setdefault("chi2", {})["20k_uniform_isotropic"] = chi2_entry(configs.cfg1(), r64, 0,
                                                                             (0, 0, 1))
    log(f"20k rods: FP64 {res['fp64_plain']['accepted_rods_per_s']:.4e} rods/s, "
        f"FP32 {res['fp32_recheck']['accepted_rods_per_s']:.4e} rods/s")


def leg_batch(a, out: dict) -> None:
    """256 packings: batched FP32 + recheck, batched plain FP64 (the batching
    gain alone) and the plain FP64 single-instance path one packing at a time.
    Instances past the jamming limit of random sequential addition end at the
    round cap (configs.CFG3_MAX_ROUNDS) or stall, identically on every path."""
    import collections

    ps32 = configs.cfg3(nf.NF_PREC_FP32_RECHECK)
    ps64 = configs.cfg3(nf.NF_PREC_FP64)
    total = sum(p.n_rods for p in ps32)

    def batched(ps, label):
        log(f"batch: {len(ps)} packings, {total} rods, batched {label}")
        with nf.Packing(ps, configs.tile_options("batch256")) as pk:
            pk.generate()  # warm-up
            t0 = time.perf_counter()
            sts = pk.generate()
            dt = time.perf_counter() - t0
            rods = [pk.rods(i) for i in range(len(ps))]
        acc = sum(s.n_accepted for s in sts)
        log(f"batched {label}: {acc / dt:.4e} rods/s ({dt * 1e3:.1f} ms, "
            f"max rounds {max(s.rounds for s in sts)})")
        return dt, sts, rods

    dt, sts, rods32 = batched(ps32, "FP32 + re-check")
    dt64b, sts64b, rods64b = batched(ps64, "plain FP64")
    acc = sum(s.n_accepted for s in sts)
    same_b = all(r.shape == q.shape and np.array_equal(r.view(np.uint32), q.view(np.uint32))
                 for r, q in zip(rods32, rods64b))
    # the plain FP64 single-instance path, one packing at a time
    with nf.Packing(ps64[0]) as pk:
        pk.generate()  # warm-up
    t_seq = 0.0
    acc64 = 0
    same = True
    for i, p in enumerate(ps64):
        with nf.Packing(p) as pk:
            t0 = time.perf_counter()
            s = pk.generate()[0]
            t_seq += time.perf_counter() - t0
            r = pk.rods()
        acc64 += s.n_accepted
        same = same and r.shape == rods32[i].shape and np.array_equal(r.view(np.uint32),
                                                                      rods32[i].view(np.uint32))
    b_rate, b64_rate, s_rate = acc / dt, acc / dt64b, acc64 / t_seq
    log(f"sequential FP64: {s_rate:.4e} rods/s ({t_seq * 1e3:.1f} ms); identical packings {same}")
    status = collections.Counter(nf.STATUS[s.status] for s in sts)
    out["extra"]["batch"] = {
        "tile_options": configs.tile_options("batch256"),
        "workload": f"{len(ps32)} packings, N log-uniform 2k-200k (sum {total}), aspect 20-500, "
                    "vf 0.1-2%, half isotropic / half vMF kappa 1-20 about random axes, "
                    f"round cap {configs.CFG3_MAX_ROUNDS}",
        "batched_fp32_recheck_rods_per_s": b_rate, "batched_ms": dt * 1e3,
        "batched_fp64_rods_per_s": b64_rate, "batched_fp64_ms": dt64b * 1e3,
        "sequential_fp64_rods_per_s": s_rate, "sequential_fp64_ms": t_seq * 1e3,
        "speedup_batched_over_fp64_single": b_rate / s_rate,
        "speedup_batching_fp64": b64_rate / s_rate,
        "speedup_precision_batched": b_rate / b64_rate,
        "identical_packings": bool(same and same_b), "accepted": acc, "target": total,
        "pair_tests_per_s": sum(s.pair_tests for s in sts) / dt,
        "max_rounds": max(s.rounds for s in sts),
        "statuses": dict(status),
        "jammed_instances": status.get("max_rounds", 0) + status.get("stalled", 0),
    }


def leg_contacts(a, out: dict) -> None:
    p = configs.cfg4(nf.NF_PREC_FP32_RECHECK)
    log("contacts: 4M rods, generate")
    with nf.Packing(p, configs.tile_options("4M")) as pk:
        pk.generate()  # warm-up
        t0 = time.perf_counter()
        st = pk.generate()[0]
        t_gen = time.perf_counter() - t0
        pk.contacts(p.contact_gap)  # warm-up
        t0 = time.perf_counter()
        c32 = pk.contacts(p.contact_gap)
        t_c = time.perf_counter() - t0
        cs = pk.contact_stats()
        eng = pk.contact_engine()
        t0 = time.perf_counter()
        h = pk.histograms(bins=64, centre_axis=0, pol_axis=(1, 0, 0))
        t_h = time.perf_counter() - t0
        rods = pk.rods()
        # the cell-pair engine on the same packing (its test count is the
        # number of pairs within centre reach)
        pk.set_option(nf.NF_OPT_CONTACT_ENGINE, nf.NF_ENGINE_CELLS)
        pk.contacts(p.contact_gap)
        t0 = time.perf_counter()
        cc = pk.contacts(p.contact_gap)
        t_cc = time.perf_counter() - t0
        csc = pk.contact_stats()
        pk.set_option(nf.NF_OPT_CONTACT_ENGINE, nf.NF_ENGINE_FOOTPRINT)
        pk.set_option(nf.NF_OPT_DIAG, 1)
        pk.contacts(p.contact_gap)
        dd_contacts = pk.contact_stats()["max_abs_d32_d64"]
        pk.set_option(nf.NF_OPT_DIAG, 0)
        pk.set_option(nf.NF_OPT_CONTACT_PREC, nf.NF_PREC_FP64)
        t0 = time.perf_counter()
        c64 = pk.contacts(p.contact_gap)
        t_c64 = time.perf_counter() - t0
        cs64 = pk.contact_stats()
    gen_prof = profiled_generation(p, configs.tile_options("4M"))
    km = cs["kernel_ms"]
    # the algorithmic bytes of the contact pass: N * 32 B * 14 tiles (27-cell
    # stencil with Newton's third law) + C * 12 B (I, J, gap)
    alg_bytes = st.n_accepted * 32 * 14 + c32.n * 12
    alg_achieved = alg_bytes / (km["pairs"] * 1e-3) / 1e9 if km["pairs"] > 0 else 0.0
    alg_cells = alg_bytes / (csc["kernel_ms"]["pairs"] * 1e-3) / 1e9 if csc["kernel_ms"]["pairs"] > 0 else 0.0
    log(f"contacts: {c32.n} pairs, footprint pass {km['pairs']:.3f} ms ({alg_achieved:.0f} GB/s), "
        f"cell-pair pass {csc['kernel_ms']['pairs']:.3f} ms; FP64 pass {cs64['kernel_ms']['pairs']:.3f} ms")
    out["extra"]["contacts"] = {
        "workload": "4,000,000 rods, 50 um cube, l 1 um, d 10 nm, uniform, vMF(+x, 2), gap <= 2 nm",
        "generate_ms": t_gen * 1e3, "generate_rods_per_s": st.n_accepted / t_gen,
        "generate_rounds": st.rounds,
        "contacts": int(c32.n), "contacts_call_ms": t_c * 1e3, "kernel_ms": km, "engine": eng,
        "pair_tests": cs["pair_tests"], "rechecks": cs["rechecks"],
        "streamed_bytes": cs["tile_bytes"],
        "pair_kernel_roofline": {"bound": "hbm", "achieved": alg_achieved, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": alg_achieved / HBM_PEAK_GBS,
                                  "algorithmic_bytes_per_pass": alg_bytes,
                                  "traffic": profiled_contact_traffic(),
                                  "traffic_source": os.path.relpath(CONTACT_TRAFFIC_SUMMARY, ROOT),
                                  "kernel": "k_seg_cross (footprint engine, FP32+recheck)",
                                  "valu": valu_roof("k_seg_cross 4M", km["pairs"])},
        "cell_engine": {"contacts_call_ms": t_cc * 1e3, "kernel_ms": csc["kernel_ms"],
                        "pairs_within_reach": csc["pair_tests"], "streamed_bytes": csc["tile_bytes"],
                        "identical_list": bool(same_contacts(c32, cc)),
                        "roofline": {"bound": "hbm", "achieved": alg_cells, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": alg_cells / HBM_PEAK_GBS,
                                     "valu": valu_roof("k_pair<2, 1> 4M", csc["kernel_ms"]["pairs"])}},
        "max_abs_d32_d64_um": dd_contacts,
        "generation_profile": gen_prof,
        "fp64_contacts_call_ms": t_c64 * 1e3, "fp64_kernel_ms": cs64["kernel_ms"],
        "identical_fp32_fp64_lists": bool(same_contacts(c32, c64)),
        "histograms_ms": t_h * 1e3,
        "histogram_totals": {k: int(v.sum()) for k, v in h.items()},
    }
    out["extra"].setdefault("chi2", {})["4M_uniform_vmf_x2"] = chi2_entry(p, rods, 0, (1, 0, 0))


def leg_network(a, out: dict) -> None:
    p = configs.cfg5(nf.NF_PREC_FP32_RECHECK)
    log("network: 8M rods, generate")
    with nf.Packing(p) as pk:
        t0 = time.perf_counter()
        st = pk.generate()[0]
        t_gen = time.perf_counter() - t0
        t0 = time.perf_counter()
        c = pk.contacts(p.contact_gap)
        t_c = time.perf_counter() - t0
        pk.network()  # warm-up
        t0 = time.perf_counter()
        info, _ = pk.network()
        t_n = time.perf_counter() - t0
        pk.node_network()  # warm-up
        t0 = time.perf_counter()
        ninfo, _ = pk.node_network()
        t_nd = time.perf_counter() - t0
    gen_prof = profiled_generation(p)
    log(f"network: {c.n} contacts, {info['n_components']} components, largest {info['largest']}, "
        f"nnz {info['nnz']}, {t_n * 1e3:.1f} ms")
    log(f"network: node level {ninfo['n_nodes']} nodes, {ninfo['n_segments']} segments, "
        f"{ninfo['n_screened']} screened contacts, built in {ninfo['device_ms']:.1f} ms on the device "
        f"({t_nd * 1e3:.1f} ms with every table copied to the host)")
    out["extra"]["network"] = {
        "workload": "8,000,000 rods, 40 um cube, l 1 um, d 10 nm, 16 agglomerates, vMF(+x, 3), "
                    "gap <= 2 nm, electrodes x = 0 / 40 um, g0 exp(-gap / 0.1 nm)",
        "generate_ms": t_gen * 1e3, "generate_rods_per_s": st.n_accepted / t_gen,
        "rounds": st.rounds, "contacts": int(c.n), "contacts_call_ms": t_c * 1e3,
        "network_call_ms": t_n * 1e3, "info": info, "generation_profile": gen_prof,
        "node_network": {"info": ninfo, "build_device_ms": ninfo["device_ms"],
                         "call_with_host_copies_ms": t_nd * 1e3},
    }


def leg_conduct(a, out: dict) -> None:
    """DC conductance of a percolating network: generation, contacts,
    network, then the Jacobi-PCG solve between the electrode planes."""
    p = nf.normalize(nf.default_params(box=(20, 20, 20), n_rods=1000000, len=1.5, radius=0.005,
                                       seed=11, contact_gap=0.005, elec_axis=0, tunnel_lambda=5e-4,
                                       precision=nf.NF_PREC_FP32_RECHECK))
    log("conduct: 1M rods, generate, contacts, network")
    with nf.Packing(p) as pk:
        pk.generate()
        c = pk.contacts(p.contact_gap)
        info, net = pk.network()
        pk.conductance(tol=1e-10, max_iter=20000)  # warm-up
        t0 = time.perf_counter()
        ci, _ = pk.conductance(tol=1e-10, max_iter=20000)
        t_s = time.perf_counter() - t0
    m = ci["unknowns"]
    # reduced-system entries: a diagonal per unknown plus two per rod-rod edge
    # inside the current-carrying components (bounded by all edges)
    nnz = m + 2 * info["n_edges"]
    # algorithmic bytes per iteration: CSR (12 B per entry) + p gathers, and
    # the vector passes (spmv: p, q; update: dinv, p, q, x rw, r rw, z; dir: z, p rw)
    per_it = 12 * nnz + 8 * nnz + 8 * m * (2 + 9 + 3)
    ms_it = ci["solve_ms"] / max(ci["iterations"], 1)
    log(f"conduct: {info['percolating']} percolating of {info['n_components']} components, {m} unknowns, "
        f"{ci['iterations']} CG iterations, {ci['solve_ms']:.1f} ms ({ms_it * 1e3:.1f} us/iteration), "
        f"G = {ci['conductance']:.6g}, residual {ci['residual']:.2e}")
    out["extra"]["conduct"] = {
        "workload": "1,000,000 rods, 20 um cube, l 1.5 um, d 10 nm, isotropic, gap <= 5 nm, "
                    "electrodes x = 0 / 20 um, g = exp(-gap / 0.5 nm); Jacobi-PCG to 1e-10",
        "contacts": int(c.n), "network": info, "solve": ci, "call_ms": t_s * 1e3,
        "ms_per_iteration": ms_it, "bytes_per_iteration_upper": per_it,
        "gb_per_s_upper": per_it / (ms_it * 1e-3) / 1e9 if ms_it > 0 else None,
        "current_mismatch": abs(ci["conductance"] - ci["current_sink"]) / max(abs(ci["conductance"]), 1e-300),
    }


def leg_prof_main(a, out: dict) -> None:
    """One generation of the headline workload and nothing else: the
    command that rocprofv3 counter summaries of its kernels are taken on."""
    p = configs.cfg2(nf.NF_PREC_FP32_RECHECK)
    with nf.Packing(p, configs.tile_options("500k")) as pk:  # the timed steps' launch shape
        st = pk.generate()[0]
    out["extra"]["prof_main"] = {"workload": "one 500k-rod headline generation", "rounds": st.rounds,
                                 "rounds_launched": st.rounds_launched, "tile_bytes": st.tile_bytes,
                                 "accepted": st.n_accepted}


def leg_prof_contacts(a, out: dict) -> None:
    """One generation of the 4M-rod contact workload, one footprint contact
    pass and one cell-pair contact pass (FP32 + recheck): the command that
    rocprofv3 counter summaries of the contact kernels are taken on."""
    p = configs.cfg4(nf.NF_PREC_FP32_RECHECK)
    with nf.Packing(p) as pk:
        st = pk.generate()[0]
        c = pk.contacts(p.contact_gap)
        cs = pk.contact_stats()
        pk.set_option(nf.NF_OPT_CONTACT_ENGINE, nf.NF_ENGINE_CELLS)
        cc = pk.contacts(p.contact_gap)
        csc = pk.contact_stats()
    out["extra"]["prof_contacts"] = {"workload": "one 4M-rod generation + footprint pass + cell-pair pass",
                                     "rounds": st.rounds, "contacts": int(c.n),
                                     "footprint_streamed_bytes": cs["tile_bytes"],
                                     "cell_engine_streamed_bytes": csc["tile_bytes"],
                                     "identical_list": bool(same_contacts(c, cc))}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads: int):
    """The oracle (FP64 C restatement of the generation: same cell list,
    round rule and exact capsule distance; OpenMP over cells) on bounded
    samples of the 20k-rod and 500k-rod workloads at the same rod density
    and laws, each the median of 5 runs, on 1 thread and on `threads`."""
    import oracle

    samples = {
        "20k_fp64": {threads: configs.cfg1(nf.NF_PREC_FP64),
                     1: configs.cfg1(nf.NF_PREC_FP64, box=(2.5, 2.5, 2.5), n_rods=2500)},
        "500k_lognormal_agglomerates": {
            threads: configs.cfg2(nf.NF_PREC_FP64, box=(5, 5, 5), n_rods=7812),
            1: configs.cfg2(nf.NF_PREC_FP64, box=(20 / 6, 20 / 6, 20 / 6), n_rods=2315)},
    }
    runs = {}
    for name, by_t in samples.items():
        for t, p in by_t.items():
            rates, prates = [], []
            for _ in range(5):
                t0 = time.perf_counter()
                _, st, _ = oracle.generate(p, threads=t)
                dt = time.perf_counter() - t0
                rates.append(st["n_accepted"] / dt)
                prates.append(st["pair_tests"] / dt)
            runs[f"{name}_{t}t"] = {"threads": t, "rods": int(p.n_rods), "box_um": p.box[0],
                                    "accepted_rods_per_s": float(np.median(rates)),
                                    "pair_tests_per_s": float(np.median(prates))}
            log(f"CPU {name} on {t} threads: {runs[f'{name}_{t}t']['accepted_rods_per_s']:.4e} rods/s")
    main = runs[f"500k_lognormal_agglomerates_{threads}t"]
    return {"value": main["accepted_rods_per_s"], "unit": "rods/s", "cores": threads, "kind": "port",
            "sample": ("oracle generation of 7,812 rods in a 5 um cube (the 500k-rod workload's laws "
                       "and rod density at 1/64 of its volume), median of 5; 1-thread and 20k-rod "
                       "samples under 'runs'; the contact list under 'contacts'"),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "runs": runs,
            "contacts": cpu_contacts(threads)}


def cpu_contacts(threads: int) -> dict:
    """The oracle's contact list (FP64 C, exact capsule distance, OpenMP over
    cells) on bounded samples of the 4M-rod contact workload (same laws and
    rod density): 500,000 rods in a 25 um cube on `threads` threads (median
    of 5) and 131,072 rods in a 16 um cube on 1 thread (median of 3).  The
    samples are packed on the GPU; the GPU contact call on the same rods is
    timed beside the oracle (median of 5) and the two lists are compared."""
    import oracle

    runs = {}
    for t, edge, n, reps in ((threads, 25.0, 500000, 5), (1, 16.0, 131072, 3)):
        p = configs.cfg4(nf.NF_PREC_FP32_RECHECK, box=(edge, edge, edge), n_rods=n)
        with nf.Packing(p, configs.tile_options("4M")) as pk:
            pk.generate()
            rods = pk.rods()
            pk.contacts(p.contact_gap)  # warm-up
            gpu_s, kern = [], []
            for _ in range(5):
                t0 = time.perf_counter()
                c = pk.contacts(p.contact_gap)
                gpu_s.append(time.perf_counter() - t0)
                kern.append(sum(pk.contact_stats()["kernel_ms"].values()))
            g64 = pk.contact_gap64(c.n)
        cpu_s = []
        for _ in range(reps):
            t0 = time.perf_counter()
            o = oracle.contacts(p, rods, p.contact_gap, threads=t)
            cpu_s.append(time.perf_counter() - t0)
        same = (c.n == len(o["I"]) and np.array_equal(c.I, o["I"]) and np.array_equal(c.col, o["J"])
                and np.array_equal(g64.view(np.uint64), o["gap64"].view(np.uint64)))
        cpu, gpu = float(np.median(cpu_s)), float(np.median(gpu_s))
        runs[f"{n}_rods_{t}t"] = {
            "threads": t, "rods": int(rods.shape[0]), "box_um": edge, "contacts": int(c.n),
            "cpu_s": cpu, "cpu_rods_per_s": rods.shape[0] / cpu,
            "gpu_call_ms": gpu * 1e3, "gpu_kernel_ms": float(np.median(kern)),
            "gpu_rods_per_s": rods.shape[0] / gpu, "identical_list": bool(same)}
        log(f"CPU contacts, {rods.shape[0]} rods on {t} threads: {cpu:.3f} s "
            f"(GPU call {gpu * 1e3:.2f} ms), {c.n} contacts, identical {same}")
    return runs


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--legs", default="main,batch,contacts,network,conduct")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 64 MiB copy kernel (reference for HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the packings of this workload are generated on a single "
              "MI355X (multi-GPU sharding is a design note only); run with --gpus 1",
              file=sys.stderr)
        return 2
    legs = set(a.legs.split(","))
    if a.calibrate_copy:
        x = np.arange(1 << 23, dtype=np.float64)
        assert np.array_equal(nf.selftest("copy", x), x)
        log("copy kernel: 67108864 bytes read, 67108864 bytes written")
    res = {"extra": {}}
    if "main" in legs:
        leg_main(a, res)
    for name, fn in (("batch", leg_batch), ("contacts", leg_contacts), ("network", leg_network),
                     ("conduct", leg_conduct), ("prof_main", leg_prof_main),
                     ("prof_contacts", leg_prof_contacts)):
        if name in legs:
            fn(a, res)
    cpu = None
    if not a.no_cpu and "main" in legs:
        threads = min(16, os.cpu_count() or 1)
        log(f"CPU oracle baseline on {threads} threads")
        cpu = cpu_baseline(threads)
        log(f"CPU: {cpu['value']:.4e} rods/s")
    out = {
        "metric": METRIC,
        "value": res.get("value"),
        "unit": "rods/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": res.get("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64-recheck",
        "data": "synthetic (Philox4x32-10 candidate stream, seed 2)",
        "config": {
            "workload": "500,000 rods, 20 um periodic cube, log-normal L (median 1 um, sigma 0.3, "
                        "[0.3, 3] um), a = 5 nm, 8 Gaussian agglomerates (sigma 1.5 um, 3x peak), "
                        "vMF(+z, kappa 5), hard core 10 nm, contacts <= 2 nm",
            "n_rods": res.get("n_rods"), "box_um": [20.0, 20.0, 20.0], "seed": 2,
            "pair_tests_per_s": res.get("pair_s"),
            "parallelism": "single GPU",
        },
        "roofline": res.get("roofline"),
        "cpu_baseline": cpu,
        "extra": res["extra"],
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
