"""Cell-size autotuner for the batched packings (256 uneven instances).

Instances are grouped in buckets (decade of N x aspect-ratio tercile, see
configs.bucket).  The tuner times the batched FP32 + recheck generation of
all instances: first with one cell-edge factor f for every bucket, passed as
cell_size = -f (f times the library's default edge), then varying one
bucket's factor at a time from the best of those, keeping a change when it
shortens the batch by more than 2%.  The packing does not depend on the
cell edge - every accept decision is a function of the candidate stream
alone - so tuning changes speed only; the tests check that.

Usage: python -m nanofiller_amd.autotune [--factors 0.7,0.85,1,1.2,1.4]
Writes nanofiller_amd/cell_tuning.json.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import NF_PREC_FP32_RECHECK, NFError, Packing
from . import configs


def _batch_ms(shapes, base, fmap, repeats: int = 2) -> float:
    """Best-of-`repeats` time of one batched generation of all instances with
    cell_size = -fmap[bucket] per instance."""
    ps = []
    for i, p0 in enumerate(base):
        p = p0.copy()
        p.cell_size = -float(fmap[configs.bucket(shapes[i]["n"], shapes[i]["aspect"])])
        ps.append(p)
    best = float("inf")
    with Packing(ps) as pk:
        pk.generate()  # warm-up: first-call costs of this handle
        for _ in range(repeats):
            t0 = time.perf_counter()
            pk.generate()
            best = min(best, time.perf_counter() - t0)
    return best * 1e3


def tune(factors, n_inst: int = 256, seed: int = 3, log=print) -> dict:
    """Times the batched generation itself: the best single factor for every
    bucket first, then one pass of per-bucket coordinate descent."""
    shapes = configs.cfg3_instances(n_inst, seed)
    base = configs.cfg3(NF_PREC_FP32_RECHECK, n_inst=n_inst, tuned=False, seed=seed)
    buckets = sorted({configs.bucket(s["n"], s["aspect"]) for s in shapes})
    glob = {}
    for f in factors:
        fmap = {b: f for b in buckets}
        try:
            glob[f] = _batch_ms(shapes, base, fmap)
        except NFError as e:  # a grid too fine for some instance
            log(f"global factor {f} skipped ({e})")
            continue
        log(f"global factor {f}: {glob[f]:.1f} ms")
    f0 = min(glob, key=glob.get)
    table = {b: {"cell_factor": float(f0), "ms_by_factor": {}} for b in buckets}
    current = glob[f0]
    for b in buckets:
        for f in factors:
            if f == table[b]["cell_factor"]:
                table[b]["ms_by_factor"][str(f)] = current
                continue
            fmap = {k: v["cell_factor"] for k, v in table.items()}
            fmap[b] = f
            try:
                t = _batch_ms(shapes, base, fmap)
            except NFError:
                continue
            table[b]["ms_by_factor"][str(f)] = t
            if t < current * 0.98:  # keep a change only for a clear gain
                current = t
                table[b]["cell_factor"] = float(f)
        log(f"{b}: factor {table[b]['cell_factor']} (batch {current:.1f} ms)")
    return table


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--factors", default="0.7,0.85,1.0,1.2,1.4,1.7")
    ap.add_argument("--out", default=configs.TUNING_FILE)
    a = ap.parse_args()
    factors = [float(x) for x in a.factors.split(",")]
    table = tune(factors, log=lambda m: print(m, file=sys.stderr, flush=True))
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({b: t["cell_factor"] for b, t in table.items()}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
