"""Cell-size autotuner for the batched packings (256 uneven instances).

Instances are grouped in buckets (decade of N x aspect-ratio tercile, see
configs.bucket).  For every bucket the tuner times single-instance
FP32 + recheck generations of up to `per_bucket` representative instances
(the bucket's instances closest to its median N) at each cell-edge factor
f, passed as cell_size = -f (f times the library's default edge), and keeps
the factor with the smallest total time.  The packing does not depend on the
cell edge - every accept decision is a function of the candidate stream
alone - so tuning changes speed only; the tests check that.

Usage: python -m nanofiller_amd.autotune [--factors 0.5,0.7,1,1.4] [--per-bucket 3]
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


def tune(factors, per_bucket: int = 3, n_inst: int = 256, seed: int = 3, log=print) -> dict:
    shapes = configs.cfg3_instances(n_inst, seed)
    base = configs.cfg3(NF_PREC_FP32_RECHECK, n_inst=n_inst, tuned=False, seed=seed)
    buckets: dict[str, list[int]] = {}
    for i, s in enumerate(shapes):
        buckets.setdefault(configs.bucket(s["n"], s["aspect"]), []).append(i)
    # warm the device and the library once
    with Packing(base[0]) as pk:
        pk.generate()
    table = {}
    for b, idx in sorted(buckets.items()):
        med = float(np.median([shapes[i]["n"] for i in idx]))
        reps = sorted(idx, key=lambda i: abs(shapes[i]["n"] - med))[:per_bucket]
        times = {}
        for f in factors:
            total = 0.0
            ok = True
            for i in reps:
                p = base[i].copy()
                p.cell_size = -float(f)
                try:
                    with Packing(p) as pk:
                        t0 = time.perf_counter()
                        pk.generate()
                        total += time.perf_counter() - t0
                except NFError as e:  # a grid too fine for the instance
                    log(f"{b}: factor {f} skipped ({e})")
                    ok = False
                    break
            if ok:
                times[str(f)] = total * 1e3
        best = min(times, key=times.get)
        table[b] = {"cell_factor": float(best), "ms_by_factor": times, "instances": reps}
        log(f"{b}: {len(idx)} instances, best factor {best} ({times[best]:.1f} ms for {len(reps)})")
    return table


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--factors", default="0.35,0.5,0.7,1.0,1.4")
    ap.add_argument("--per-bucket", type=int, default=3)
    ap.add_argument("--out", default=configs.TUNING_FILE)
    a = ap.parse_args()
    factors = [float(x) for x in a.factors.split(",")]
    table = tune(factors, a.per_bucket, log=lambda m: print(m, file=sys.stderr, flush=True))
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({b: t["cell_factor"] for b, t in table.items()}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
