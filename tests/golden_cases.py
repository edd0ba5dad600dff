"""Loader of the golden fixtures written by tests/golden/make_golden.py, and
of the full-size digest fixtures written by tests/golden/make_full_size.py."""
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


# Full-size fixtures: too large to commit as arrays, so they hold SHA-256
# digests of the oracle's packing and contact list (tests/golden/make_full_size.py).
def rods_digest(rods) -> str:
    """SHA-256 of the rods as n x 8 float32 (x y z cx cy cz l a), row-major."""
    import hashlib

    r = np.ascontiguousarray(rods, np.float32)
    assert r.ndim == 2 and r.shape[1] == 8
    return hashlib.sha256(r.tobytes()).hexdigest()


def contacts_digest(I, J, image, gap32, gap64) -> str:
    """SHA-256 of a contact list: I, J (u32), image (i32), gap32 and gap64
    bits, each array whole and in that order."""
    import hashlib

    h = hashlib.sha256()
    for a, dt in ((I, np.uint32), (J, np.uint32), (image, np.int32), (gap32, np.float32), (gap64, np.float64)):
        h.update(np.ascontiguousarray(a, dt).tobytes())
    return h.hexdigest()


def network_digest(labels, row_ptr, col, val) -> str:
    """SHA-256 of a rod network: labels, row_ptr, col (u32) and the FP64
    value bits, each array whole and in that order."""
    import hashlib

    h = hashlib.sha256()
    for a, dt in ((labels, np.uint32), (row_ptr, np.uint32), (col, np.uint32), (val, np.float64)):
        h.update(np.ascontiguousarray(a, dt).tobytes())
    return h.hexdigest()


def load_digest(name):
    with open(os.path.join(GOLDEN, f"{name}_digest.json")) as f:
        meta = json.load(f)
    kw = dict(meta["params"])
    for k in ("box", "axis"):
        if k in kw:
            kw[k] = tuple(kw[k])
    return nf.normalize(nf.default_params(**kw)), meta
