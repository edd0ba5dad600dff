"""Parameter sets of the benchmark workloads (all lengths in um).

The single-packing sets take the generation cell factor the autotuner
recorded for them in cell_tuning.json ("tiles"; tuned=False: the library's
default edge); the packing does not depend on it.

cfg1  20,000 rods, l = 1, d = 10 nm, 5 um periodic cube, uniform, isotropic,
      FP64, seed 1, 1 um cells, contacts <= 2 nm.
cfg2  500,000 rods in a 20 um cube: log-normal l (median 1, sigma 0.3, in
      [0.3, 3]), d = 10 nm, background + 8 Gaussian agglomerates (sigma 1.5,
      peak 3x), vMF(+z, kappa 5), FP32 + FP64 recheck, seed 2.
cfg3  256 independent packings: N log-uniform in [2k, 200k], aspect ratio
      uniform in [20, 500] (d = 10 nm), volume fraction uniform in
      [0.1%, 2%], box edge from N, l and the fraction; even instances
      isotropic, odd ones vMF with kappa uniform in [1, 20] about a random
      axis; seed 3, instance i on Philox stream i.  Slender rods at the
      higher fractions lie past the jamming limit of random sequential
      addition (n * V_ex = 2 * vf * aspect well above 3): those instances
      end at the round cap (CFG3_MAX_ROUNDS = 100 rounds, status
      max_rounds) or stall, as the oracle's do.
cfg4  4,000,000 rods in a 50 um cube, l = 1, d = 10 nm, uniform, vMF(+x,
      kappa 2), seed 4; contacts <= 2 nm and 64-bin histograms.
cfg5  8,000,000 rods in a 40 um cube (~1% volume), l = 1, d = 10 nm,
      background + 16 agglomerates, vMF(+x, kappa 3), seed 5; contacts
      <= 2 nm, components, Kirchhoff matrix g0 exp(-gap / 1e-4 um) between
      the planes x = 0 and x = 40.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import (NF_DEN_AGGLOM, NF_LEN_LOGNORMAL, NF_ORT_VMF, NF_PREC_FP32_RECHECK, NF_PREC_FP64,
               default_params, normalize)

# Round cap of the 256-packing workload: about 46 of its instances lie past
# the jamming limit and would otherwise run the reference's 1000 rounds
# (src/cudaMain.cu:113-114) at a few accepted rods per round.
CFG3_MAX_ROUNDS = 100

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cell_tuning.json")


def cfg1(precision=NF_PREC_FP64, **kw):
    base = dict(box=(5, 5, 5), n_rods=20000, len=1.0, radius=0.005, seed=1, cell_size=1.0,
                contact_gap=0.002, precision=precision)
    base.update(kw)
    return normalize(default_params(**base))


def tuned_cell_factor(workload: str):
    """The generation cell factor the autotuner picked for a workload
    (cell_size = -factor), or None."""
    f = load_tuning().get("tiles", {}).get(workload, {}).get("cell_factor")
    return float(f) if f is not None else None


def _tuned(base: dict, workload: str, tuned: bool, kw: dict) -> dict:
    if tuned and "cell_size" not in kw:
        f = tuned_cell_factor(workload)
        if f is not None:
            base["cell_size"] = -f
    base.update(kw)
    return base


def cfg2(precision=NF_PREC_FP32_RECHECK, tuned: bool = True, **kw):
    base = dict(box=(20, 20, 20), n_rods=500000, len=1.0, len_law=NF_LEN_LOGNORMAL, len_sigma=0.3,
                len_min=0.3, len_max=3.0, radius=0.005, seed=2, den_law=NF_DEN_AGGLOM, n_agg=8,
                agg_sigma=1.5, agg_peak=3.0, ort_law=NF_ORT_VMF, kappa=5.0, axis=(0, 0, 1),
                contact_gap=0.002, precision=precision)
    return normalize(default_params(**_tuned(base, "500k", tuned, kw)))


def cfg4(precision=NF_PREC_FP32_RECHECK, tuned: bool = True, **kw):
    base = dict(box=(50, 50, 50), n_rods=4000000, len=1.0, radius=0.005, seed=4,
                ort_law=NF_ORT_VMF, kappa=2.0, axis=(1, 0, 0), contact_gap=0.002,
                precision=precision)
    return normalize(default_params(**_tuned(base, "4M", tuned, kw)))


def cfg5(precision=NF_PREC_FP32_RECHECK, tuned: bool = True, **kw):
    base = dict(box=(40, 40, 40), n_rods=8000000, len=1.0, radius=0.005, seed=5,
                den_law=NF_DEN_AGGLOM, n_agg=16, agg_sigma=2.0, agg_peak=3.0, ort_law=NF_ORT_VMF,
                kappa=3.0, axis=(1, 0, 0), contact_gap=0.002, elec_axis=0, tunnel_lambda=1e-4,
                tunnel_g0=1.0, precision=precision)
    return normalize(default_params(**_tuned(base, "8M", tuned, kw)))


def cfg3_instances(n_inst: int = 256, seed: int = 3, n_scale: float = 1.0, n_min: int = 200):
    """Shape of every cfg3 instance: dicts of N, aspect, l, vf, edge, law.
    n_scale < 1 shrinks every N (at least n_min rods) at the same volume
    fraction, aspect and law: the same mix at sizes the oracle checks."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    a = 0.005
    for i in range(n_inst):
        n = int(round(math.exp(rng.uniform(math.log(2000), math.log(200000)))))
        if n_scale != 1.0:
            n = max(n_min, int(round(n * n_scale)))
        ar = float(rng.uniform(20.0, 500.0))
        length = ar * 2 * a
        vf = float(rng.uniform(0.001, 0.02))
        edge = (n * math.pi * a * a * length / vf) ** (1.0 / 3.0)
        kappa = float(rng.uniform(1.0, 20.0))
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        out.append(dict(n=n, aspect=ar, len=length, vf=vf, edge=edge,
                        vmf=(i % 2 == 1), kappa=kappa, axis=tuple(float(x) for x in v)))
    return out


def bucket(n: int, aspect: float) -> str:
    """Autotuning bucket: decade of N x aspect-ratio tercile."""
    dec = min(max(int(math.floor(math.log10(n))), 3), 5)
    ter = 0 if aspect < 180.0 else (1 if aspect < 340.0 else 2)
    return f"n{dec}_ar{ter}"


def load_tuning() -> dict:
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            return json.load(f)
    return {}


def tile_options(workload: str) -> dict:
    """Handle options of the launch shape the autotuner picked for a
    workload ("20k", "500k", "4M", "batch256"; nanofiller_amd/autotune.py
    --tiles), {} when untuned."""
    from . import NF_OPT_PART_SHIFT, NF_OPT_SPLIT_MIN, NF_OPT_WAVES_NEW, NF_OPT_WAVES_OLD

    t = load_tuning().get("tiles", {}).get(workload, {})
    keys = {"waves_old": NF_OPT_WAVES_OLD, "waves_new": NF_OPT_WAVES_NEW, "part_shift": NF_OPT_PART_SHIFT,
            "split_min": NF_OPT_SPLIT_MIN}
    return {keys[k]: int(v) for k, v in t.items() if k in keys}


def cfg3(precision=NF_PREC_FP32_RECHECK, n_inst: int = 256, tuned: bool = True, seed: int = 3,
         cell_factor: float | None = None, n_scale: float = 1.0,
         max_rounds: int = CFG3_MAX_ROUNDS):
    """The 256 parameter sets.  The cell edge is the library's default edge
    scaled by the factor of the instance's bucket in cell_tuning.json (1.0
    when untuned), passed as a negative cell_size."""
    tune = load_tuning() if tuned else {}
    ps = []
    for i, s in enumerate(cfg3_instances(n_inst, seed, n_scale)):
        f = cell_factor if cell_factor is not None else tune.get(bucket(s["n"], s["aspect"]), {}).get(
            "cell_factor", 1.0)
        kw = dict(box=(s["edge"],) * 3, n_rods=s["n"], len=s["len"], radius=0.005, seed=seed,
                  stream=i, contact_gap=0.002, precision=precision, cell_size=-f,
                  max_rounds=max_rounds)
        if s["vmf"]:
            kw.update(ort_law=NF_ORT_VMF, kappa=s["kappa"], axis=s["axis"])
        ps.append(normalize(default_params(**kw)))
    return ps
