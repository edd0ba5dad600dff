"""Autotuner: cell size of the batched packings and the launch shape of the
generation's pair passes.

Cell size (the 256 uneven packings).

Instances are grouped in buckets (decade of N x aspect-ratio tercile, see
configs.bucket).  The tuner times the batched FP32 + recheck generation of
all instances: first with one cell-edge factor f for every bucket, passed as
cell_size = -f (f times the library's default edge), then varying one
bucket's factor at a time from the best of those, keeping a change when it
shortens the batch by more than 2%.  The packing does not depend on the
cell edge - every accept decision is a function of the candidate stream
alone - so tuning changes speed only; the tests check that.

Launch shape (--tiles): per workload (the 20k-, 500k-, 4M- and 8M-rod
packings and the 256-packing batch), the generation cell edge as a factor of
the library's default (single packings without a set cell size), the waves
per CU of the reject-vs-accepted and same-round passes, and for launches
with few cells per wave the split of a dense cell (work items as a shift,
and the probes per item at least), by one pass of coordinate descent over
the generation time from the library's defaults, keeping a change for a
gain over 2%.  The batch's launches have a million cells, so only its wave
counts are tuned.  The packing does not depend on any of these
(test_tile_shape_does_not_change_packing, test_cell_size_does_not_change_packing).

Usage: python -m nanofiller_amd.autotune [--factors 0.7,0.85,1,1.2,1.4]
       python -m nanofiller_amd.autotune --tiles
Writes nanofiller_amd/cell_tuning.json (bucket cell factors; "tiles").
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import (NF_OPT_PART_SHIFT, NF_OPT_SPLIT_MIN, NF_OPT_WAVES_NEW, NF_OPT_WAVES_OLD,
               NF_PREC_FP32_RECHECK, NFError, Packing)
from . import configs

# option keys of a tile entry in cell_tuning.json
TILE_KEYS = {"waves_old": NF_OPT_WAVES_OLD, "waves_new": NF_OPT_WAVES_NEW, "part_shift": NF_OPT_PART_SHIFT,
             "split_min": NF_OPT_SPLIT_MIN}
TILE_DEFAULTS = {"cell_factor": 1.0, "waves_old": 16, "waves_new": 16, "part_shift": 1, "split_min": 32}
# cell_factor scales the library's default generation cell edge (0.85 of a
# typical pair reach; 1.18 makes it one reach), passed as cell_size = -f
TILE_CHOICES = {"cell_factor": [0.85, 1.0, 1.18, 1.35], "waves_old": [8, 12, 16], "waves_new": [8, 12, 16],
                "part_shift": [0, 1, 2], "split_min": [16, 32, 64]}


def _gen_ms(params, tiles: dict, repeats: int) -> float:
    """Best-of-`repeats` generation time with the tile options set (and the
    cell factor, when tuned, as cell_size = -factor)."""
    opts = {TILE_KEYS[k]: v for k, v in tiles.items() if k in TILE_KEYS}
    if "cell_factor" in tiles:
        params = params.copy()
        params.cell_size = -float(tiles["cell_factor"])
    best = float("inf")
    with Packing(params, opts) as pk:
        pk.generate()  # warm-up
        for _ in range(repeats):
            t0 = time.perf_counter()
            pk.generate()
            best = min(best, time.perf_counter() - t0)
    return best * 1e3


def tile_workloads() -> dict:
    """The workloads the launch shape is tuned on: parameters, keys tuned,
    repeats."""
    launch = ["waves_old", "waves_new", "part_shift", "split_min"]
    return {
        # the 20k case keeps its 1 um cells (the reference's domain grid)
        "20k": (configs.cfg1(NF_PREC_FP32_RECHECK), launch, 3),
        "500k": (configs.cfg2(NF_PREC_FP32_RECHECK, tuned=False), ["cell_factor"] + launch, 5),
        "4M": (configs.cfg4(NF_PREC_FP32_RECHECK, tuned=False), ["cell_factor"] + launch, 3),
        "8M": (configs.cfg5(NF_PREC_FP32_RECHECK, tuned=False), ["cell_factor", "waves_old", "waves_new"], 1),
        # per-bucket cell factors (tune()); its launches hold a million cells
        "batch256": (configs.cfg3(NF_PREC_FP32_RECHECK), ["waves_old", "waves_new"], 1),
    }


def tune_tiles(names=None, log=print) -> dict:
    out = {}
    for name, (params, keys, reps) in tile_workloads().items():
        if names and name not in names:
            continue
        cur = {k: v for k, v in TILE_DEFAULTS.items() if k in keys or k in TILE_KEYS}
        cur_ms = _gen_ms(params, cur, reps)
        tried = {json.dumps(cur, sort_keys=True): cur_ms}
        log(f"{name}: defaults {cur_ms:.2f} ms")
        for k in keys:
            for v in TILE_CHOICES[k]:
                if v == cur[k]:
                    continue
                trial = dict(cur, **{k: v})
                t = _gen_ms(params, trial, reps)
                tried[json.dumps(trial, sort_keys=True)] = t
                log(f"{name}: {k}={v} {t:.2f} ms")
                if t < cur_ms * 0.98:
                    cur, cur_ms = trial, t
        d0 = {k: v for k, v in TILE_DEFAULTS.items() if k in keys or k in TILE_KEYS}
        out[name] = dict(cur, ms=cur_ms, ms_defaults=tried[json.dumps(d0, sort_keys=True)],
                         tried={kk: vv for kk, vv in tried.items()})
        log(f"{name}: {cur} {cur_ms:.2f} ms")
    return out


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
    ap.add_argument("--tiles", action="store_true", help="tune the pair passes' launch shape")
    ap.add_argument("--workloads", default="", help="with --tiles: comma-separated subset")
    ap.add_argument("--out", default=configs.TUNING_FILE)
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    current = configs.load_tuning() if a.out == configs.TUNING_FILE else {}
    if a.tiles:
        tiles = tune_tiles([w for w in a.workloads.split(",") if w] or None, log=log)
        current.setdefault("tiles", {}).update(tiles)
        print(json.dumps({k: {kk: v[kk] for kk in TILE_DEFAULTS if kk in v} for k, v in tiles.items()}))
    else:
        factors = [float(x) for x in a.factors.split(",")]
        table = tune(factors, log=log)
        current.update(table)
        print(json.dumps({b: t["cell_factor"] for b, t in table.items()}))
    with open(a.out, "w") as f:
        json.dump(current, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
