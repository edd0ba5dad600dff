"""nf_histograms_noalloc (scratch in a static device array, no allocation)
against nf_histograms, through the C ABI, on two calls in a row: the second
on a smaller packing with fewer bins, so stale counts or edges left in the
shared array by the first call would show."""
import ctypes as C

import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


def _spec(bins, pol_axis, phi_bins, theta_bins, centre_axis=0):
    s = nf.HistSpec()
    s.bins = bins
    s.centre_axis = centre_axis
    for i in range(3):
        s.pol_axis[i] = pol_axis[i]
    s.phi_bins = phi_bins
    s.theta_bins = theta_bins
    return s


def _call(fn, pk, spec):
    nb, n2 = spec.bins, spec.phi_bins * spec.theta_bins
    out = {k: np.full(nb, -7, np.int64) for k in ("centre", "polar", "mutual")}
    out["phitheta"] = np.full(max(n2, 1), -7, np.int64)
    rc = fn(pk._h, C.byref(spec), out["centre"].ctypes.data, out["polar"].ctypes.data,
            out["mutual"].ctypes.data, out["phitheta"].ctypes.data if n2 else None)
    assert rc == nf.NF_OK, nf.lib().nf_last_error(pk._h)
    return out


@pytest.mark.parametrize("order", ["noalloc_first", "alloc_first"])
def test_noalloc_histograms_match(order):
    L = nf.lib()
    runs = [
        (cases.small_cfg2(nf.NF_PREC_FP32_RECHECK), _spec(64, (0, 0, 1), 32, 32)),
        (cases.small_iso(nf.NF_PREC_FP64, n_rods=500), _spec(16, (1, 0, 0), 8, 4, centre_axis=2)),
    ]
    for p, spec in runs:
        with nf.Packing(p) as pk:
            pk.generate()
            rods = pk.rods()
            c = pk.contacts(p.contact_gap)
            if order == "noalloc_first":
                a = _call(L.nf_histograms_noalloc, pk, spec)
                b = _call(L.nf_histograms, pk, spec)
            else:
                b = _call(L.nf_histograms, pk, spec)
                a = _call(L.nf_histograms_noalloc, pk, spec)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        o = oracle.histograms(p, rods, c.I, c.col, bins=spec.bins, centre_axis=spec.centre_axis,
                              pol_axis=tuple(spec.pol_axis), phi_bins=spec.phi_bins,
                              theta_bins=spec.theta_bins)
        for k in ("centre", "polar", "mutual"):
            np.testing.assert_array_equal(a[k], o[k], err_msg=k)
        np.testing.assert_array_equal(a["phitheta"].reshape(o["phitheta"].shape), o["phitheta"])
        assert a["centre"].sum() == len(rods) and a["mutual"].sum() == c.n


def test_noalloc_rejects_oversized_requests():
    p = cases.small_iso(nf.NF_PREC_FP64, n_rods=300)
    with nf.Packing(p) as pk:
        pk.generate()
        for spec in (_spec(257, (0, 0, 1), 0, 0), _spec(8, (0, 0, 1), 128, 64)):
            out = np.zeros(8192, np.int64)
            rc = nf.lib().nf_histograms_noalloc(pk._h, C.byref(spec), out.ctypes.data,
                                                out.ctypes.data, None, out.ctypes.data)
            assert rc == nf.NF_E_INVALID
