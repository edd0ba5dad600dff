"""Writes tests/golden/<case>_digest.json from the CPU oracle, for workloads
at full size whose packings and contact lists are too large to commit:

- headline_500k: 500,000 rods, log-normal lengths (median 1 um, sigma 0.3,
  [0.3, 3] um), 8 Gaussian agglomerates (sigma 1.5 um, peak 3), vMF(+z,
  kappa 5), a = 5 nm, 20 um periodic cube, seed 2, contacts within 2 nm;
- contacts_4M: 4,000,000 rods, l 1 um, a = 5 nm, uniform, vMF(+x, kappa 2),
  50 um cube, seed 4, contacts within 2 nm, and its three 64-bin histograms
  (centre x, polar angle and mutual angle about +x), stored whole;
- network_8M: 8,000,000 rods, l 1 um, a = 5 nm, 16 agglomerates (sigma
  2 um, peak 3), vMF(+x, kappa 3), 40 um cube, seed 5, contacts within
  2 nm, and its network between electrode planes x = 0 and x = 40 um with
  g = exp(-gap / 0.1 nm): digest of the labels, CSR and value bits.

Each fixture holds SHA-256 digests (tests/golden_cases.py) of the packing
and of the contact list, with the counters; the GPU tests compute the same
digests of the device's results.  The schedule bound max_batch is written
out, so that a fixture does not depend on the library's default for it.
Several minutes per case on 8 threads.

Usage: python tests/golden/make_full_size.py [threads] [case ...]
"""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import nanofiller_amd as nf  # noqa: E402
import oracle  # noqa: E402
from tests.golden_cases import contacts_digest, network_digest, rods_digest  # noqa: E402

CASES = {
    "headline_500k": dict(box=(20, 20, 20), n_rods=500000, len=1.0, len_law=nf.NF_LEN_LOGNORMAL,
                          len_sigma=0.3, len_min=0.3, len_max=3.0, radius=0.005, seed=2,
                          den_law=nf.NF_DEN_AGGLOM, n_agg=8, agg_sigma=1.5, agg_peak=3.0,
                          ort_law=nf.NF_ORT_VMF, kappa=5.0, axis=(0, 0, 1), contact_gap=0.002,
                          precision=nf.NF_PREC_FP32_RECHECK),
    "contacts_4M": dict(box=(50, 50, 50), n_rods=4000000, len=1.0, radius=0.005, seed=4,
                        ort_law=nf.NF_ORT_VMF, kappa=2.0, axis=(1, 0, 0), contact_gap=0.002,
                        precision=nf.NF_PREC_FP32_RECHECK),
    "network_8M": dict(box=(40, 40, 40), n_rods=8000000, len=1.0, radius=0.005, seed=5,
                       den_law=nf.NF_DEN_AGGLOM, n_agg=16, agg_sigma=2.0, agg_peak=3.0,
                       ort_law=nf.NF_ORT_VMF, kappa=3.0, axis=(1, 0, 0), contact_gap=0.002,
                       elec_axis=0, tunnel_lambda=1e-4, tunnel_g0=1.0,
                       precision=nf.NF_PREC_FP32_RECHECK),
}
# the histograms of a case, stored whole (64 bins about +x)
HIST = {"contacts_4M": dict(bins=64, centre_axis=0, pol_axis=(1.0, 0.0, 0.0))}
# the cases whose network is digested too
NETWORK = {"network_8M"}


def make(name: str, threads: int) -> None:
    kw = dict(CASES[name])
    kw["max_batch"] = int(nf.normalize(nf.default_params(**CASES[name])).max_batch)
    p = nf.normalize(nf.default_params(**kw))
    t0 = time.time()
    rods, st, acc = oracle.generate(p, threads=threads)
    t1 = time.time()
    c = oracle.contacts(p, rods, p.contact_gap, threads=threads)
    t2 = time.time()
    meta = {
        "case": name,
        "params": {k: (list(v) if isinstance(v, tuple) else v) for k, v in kw.items()},
        "stats": st,
        "round_accepts": [int(x) for x in acc],
        "rods_sha256": rods_digest(rods),
        "contacts": int(len(c["I"])),
        "contact_pair_tests": int(c["pair_tests"]),
        "contacts_sha256": contacts_digest(c["I"], c["J"], c["image"], c["gap"], c["gap64"]),
    }
    if name in HIST:
        hs = HIST[name]
        h = oracle.histograms(p, rods, c["I"], c["J"], **hs)
        meta["hist"] = {k: (list(v) if isinstance(v, tuple) else v) for k, v in hs.items()}
        for k in ("centre", "polar", "mutual"):
            meta["h_" + k] = [int(x) for x in h[k]]
    if name in NETWORK:
        info, net = oracle.network(p, rods, c["I"], c["J"], c["gap64"], c["image"])
        meta["network"] = info
        meta["network_sha256"] = network_digest(net["labels"], net["row_ptr"], net["col"], net["val"])
    with open(os.path.join(HERE, f"{name}_digest.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"{name}: {len(rods)} rods in {t1 - t0:.0f} s, {len(c['I'])} contacts in {t2 - t1:.0f} s")


if __name__ == "__main__":
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    for name in sys.argv[2:] or list(CASES):
        make(name, threads)
