"""VTK and COO writers (write_CNTvtk, save_coo formats) on a device packing."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle
from tests import cases

pytestmark = pytest.mark.gpu


def _packing():
    p = cases.small_iso(nf.NF_PREC_FP64, n_rods=1500, contact_gap=0.02, elec_axis=2)
    rods, _, _ = oracle.generate(p, threads=16)
    return p, rods


def test_vtk(tmp_path):
    p, rods = _packing()
    path = tmp_path / "rods.vtk"
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        pk.write_vtk(str(path))
    lines = path.read_text().splitlines()
    n = len(rods)
    assert lines[0] == "# vtk DataFile Version 2.0"
    assert lines[2] == "ASCII" and lines[3] == "DATASET UNSTRUCTURED_GRID"
    assert lines[4] == f"POINTS {n} float"
    pts = np.array([list(map(float, l.split())) for l in lines[5:5 + n]])
    assert lines[5 + n] == f"POINT_DATA {n}" and lines[6 + n] == "VECTORS CNT float"
    vec = np.array([list(map(float, l.split())) for l in lines[7 + n:7 + 2 * n]])
    assert lines[7 + 2 * n] == "SCALARS LAK float 3" and lines[8 + 2 * n] == "LOOKUP_TABLE default"
    lak = np.array([list(map(float, l.split())) for l in lines[9 + 2 * n:9 + 3 * n]])
    assert len(lines) == 9 + 3 * n
    # six significant digits
    np.testing.assert_allclose(pts, rods[:, 0:3], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(vec, rods[:, 3:6], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(lak[:, 0:2], rods[:, 6:8], rtol=1e-5)
    np.testing.assert_array_equal(lak[:, 2], np.arange(n))


def _read_coo(path):
    lines = path.read_text().splitlines()
    assert lines[0] == "#\tI\tJ\tD\tS"
    rows = [l.split("\t")[1:] for l in lines[1:]]
    if not rows:
        return np.zeros((0,), np.int64), np.zeros((0,), np.int64), np.zeros(0), np.zeros(0)
    a = np.array(rows, dtype=np.float64)
    return a[:, 0].astype(np.int64), a[:, 1].astype(np.int64), a[:, 2], a[:, 3]


def test_coo(tmp_path):
    p, rods = _packing()
    o = oracle.contacts(p, rods, p.contact_gap, threads=16)
    oi, _ = oracle.network(p, rods, o["I"], o["J"], o["gap64"], o["image"])
    with nf.Packing(p) as pk:
        pk.set_rods(rods)
        pk.contacts(p.contact_gap)
        pk.write_coo(str(tmp_path) + "/")
    I, J, D, S = _read_coo(tmp_path / "internal.srt.coo")
    zc = ((o["image"] >> 16) & 0xFF) == 0  # contacts not crossing the z electrodes
    np.testing.assert_array_equal(I, o["I"][zc])
    np.testing.assert_array_equal(J, o["J"][zc])
    np.testing.assert_allclose(D, o["gap"][zc], rtol=1e-5, atol=1e-9)
    assert np.all(S >= 0)
    Is, Js, Ds, _ = _read_coo(tmp_path / "source.elec.coo")
    Ik, Jk, Dk, _ = _read_coo(tmp_path / "sink.elec.coo")
    assert len(Is) == oi["n_source"] and len(Ik) == oi["n_sink"]
    assert np.all(Js == 4294967295) and np.all(Jk == 4294967294)
    assert np.all(np.diff(Is) > 0) and np.all(np.diff(Ik) > 0)
    assert np.all((Ds >= 0) & (Ds <= p.contact_gap * (1 + 1e-5)))
