"""Writes the golden fixtures of tests/golden from the CPU oracle.

The reference ships no golden vectors for this path, so the fixtures are
oracle outputs on small seeded cases.  They anchor regressions: the CPU
tests re-run the oracle against them, the GPU tests compare the device path
against them without running the oracle.  Each case stores its parameters,
the packing (rod bits), the contact list (including FP64 gap bits), the
histograms and the network.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import nanofiller_amd as nf  # noqa: E402
import oracle  # noqa: E402

CASES = {
    "iso_fp64": dict(box=(3, 3, 3), n_rods=400, len=1.0, radius=0.005, seed=7, contact_gap=0.05,
                     tunnel_lambda=0.01, elec_axis=0, precision=nf.NF_PREC_FP64, max_batch=413),
    "lognormal_vmf_agglom_fp32": dict(box=(4, 4, 4), n_rods=1500, len=1.0,
                                      len_law=nf.NF_LEN_LOGNORMAL, len_sigma=0.3, len_min=0.3,
                                      len_max=3.0, radius=0.005, seed=2, den_law=nf.NF_DEN_AGGLOM,
                                      n_agg=4, agg_sigma=1.0, agg_peak=3.0, ort_law=nf.NF_ORT_VMF,
                                      kappa=5.0, axis=(0, 0, 1), contact_gap=0.02,
                                      tunnel_lambda=0.005, elec_axis=2,
                                      precision=nf.NF_PREC_FP32_RECHECK, max_batch=979),
    "tiny_box_images": dict(box=(1.5, 1.5, 1.5), n_rods=40, len=2.0, radius=0.01, seed=11,
                            contact_gap=0.05, tunnel_lambda=0.01, elec_axis=1,
                            precision=nf.NF_PREC_FP64, max_batch=256),
}
HIST = dict(bins=16, centre_axis=0, pol_axis=(0.0, 0.0, 1.0), phi_bins=8, theta_bins=8)


def params(name):
    return nf.normalize(nf.default_params(**CASES[name]))


def make(name):
    p = params(name)
    rods, st, acc = oracle.generate(p, threads=1)
    c = oracle.contacts(p, rods, p.contact_gap)
    h = oracle.histograms(p, rods, c["I"], c["J"], **HIST)
    info, net = oracle.network(p, rods, c["I"], c["J"], c["gap64"], c["image"])
    arrays = {
        "rods_bits": rods.view(np.uint32),
        "round_accepts": acc,
        "I": c["I"], "J": c["J"], "image": c["image"],
        "gap_bits": c["gap"].view(np.uint32), "gap64_bits": c["gap64"].view(np.uint64),
        "h_centre": h["centre"], "h_polar": h["polar"], "h_mutual": h["mutual"],
        "h_phitheta": h["phitheta"],
        "labels": net["labels"], "row_ptr": net["row_ptr"], "col": net["col"],
        "val_bits": net["val"].view(np.uint64),
    }
    meta = {"case": name, "params": CASES[name], "stats": st, "pair_tests": c["pair_tests"],
            "network": info, "hist": HIST}
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"{name}: {len(rods)} rods, {len(c['I'])} contacts, nnz {info['nnz']}")


if __name__ == "__main__":
    for n in CASES:
        make(n)
