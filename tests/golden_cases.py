"""Loader of the golden fixtures written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

import nanofiller_amd as nf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["iso_fp64", "lognormal_vmf_agglom_fp32", "tiny_box_images"]


def load(name):
    with open(os.path.join(GOLDEN, f"{name}.json")) as f:
        meta = json.load(f)
    with np.load(os.path.join(GOLDEN, f"{name}.npz")) as z:
        arr = {k: z[k] for k in z.files}
    kw = dict(meta["params"])
    for k in ("box", "axis"):
        if k in kw:
            kw[k] = tuple(kw[k])
    p = nf.normalize(nf.default_params(**kw))
    return p, meta, arr
