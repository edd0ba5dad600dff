"""State files on the device: read a reference-layout file, save a packing
and parse it with struct, and round-trip through the CLI."""
import os
import subprocess

import numpy as np
import pytest

import nanofiller_amd as nf
from tests import state_format

pytestmark = pytest.mark.gpu


def _params(n=3000, box=(4.0, 4.0, 4.0), **kw):
    return nf.normalize(nf.default_params(box=box, n_rods=n, len=1.0, radius=0.005, seed=7,
                                          contact_gap=0.01, **kw))


def test_load_reference_layout_file(tmp_path):
    # 2 x 2 x 2 cells with spare capacity: empty slots (l = 0) are skipped,
    # the rest come out in slot order
    path = str(tmp_path / "s.bin")
    r0 = [0.5, 0.5, 0.5, 1.0, 0.0, 0.0, 1.0, 0.005]
    r1 = [3.5, 0.5, 0.5, 0.0, 1.0, 0.0, 0.8, 0.005]
    r2 = [3.5, 3.5, 3.5, 0.0, 0.0, 1.0, 1.2, 0.004]
    cells = {0: [r0], 1: [r1], 7: [r2]}
    state_format.write(path, (4.0, 4.0, 4.0), (2, 2, 2), 1.0, 0.005, cells, cap=[2] * 8)
    with nf.Packing(_params(n=10)) as pk:
        n = pk.load_state(path)
        rods = pk.rods()
    assert n == 3
    np.testing.assert_array_equal(rods, np.array([r0, r1, r2], np.float32))


def test_load_rejects_other_box(tmp_path):
    path = str(tmp_path / "s.bin")
    state_format.write(path, (5.0, 4.0, 4.0), (1, 1, 1), 1.0, 0.005,
                       {0: [[0.5, 0.5, 0.5, 1, 0, 0, 1, 0.005]]})
    with nf.Packing(_params(n=10)) as pk:
        with pytest.raises(nf.NFError) as e:
            pk.load_state(path)
    assert e.value.code == nf.NF_E_INVALID and "box" in str(e.value)


def test_save_matches_the_layout_and_round_trips(tmp_path):
    p = _params()
    path = str(tmp_path / "output.bin")
    with nf.Packing(p) as pk:
        pk.generate()
        rods = pk.rods()
        c0 = pk.contacts(p.contact_gap)
        pk.save_state(path)
    s = state_format.read(path)
    n = len(rods)
    assert s["n_rods"] == n and s["slots"] == n
    ext = s["ext"]
    assert s["box"] == pytest.approx(tuple(p.box))
    assert s["edge"] == pytest.approx(tuple(p.box[i] / ext[i] for i in range(3)), rel=1e-6)
    assert np.array_equal(np.cumsum(s["occ"].astype(np.int64)), s["addr"])
    d = s["data"]
    # k is the cell of the centre, and every slot of a cell holds that cell's rods
    cell = np.zeros(n, np.int64)
    mult = 1
    for a in range(3):
        ca = np.clip(np.floor(d[a].astype(np.float64) * ext[a] / p.box[a]), 0, ext[a] - 1).astype(np.int64)
        cell += ca * mult
        mult *= ext[a]
    np.testing.assert_array_equal(d[8].astype(np.int64), cell)
    slot_cell = np.repeat(np.arange(len(s["occ"])), s["occ"].astype(np.int64))
    np.testing.assert_array_equal(slot_cell, cell)
    # the same rods, permuted
    saved = d[:8].T.copy()
    key = lambda r: np.lexsort(r.T[::-1])
    np.testing.assert_array_equal(saved[key(saved)], rods[key(rods)])
    # reload: the contact graph is the same under the permutation
    with nf.Packing(p) as pk:
        assert pk.load_state(path) == n
        back = pk.rods()
        np.testing.assert_array_equal(back, saved)
        c1 = pk.contacts(p.contact_gap)
    order = key(rods)
    pos_old = np.empty(n, np.int64)
    pos_old[order] = np.arange(n)
    pos_new = np.empty(n, np.int64)
    pos_new[key(back)] = np.arange(n)
    e0 = {(min(pos_old[i], pos_old[j]), max(pos_old[i], pos_old[j]))
          for i, j in zip(c0.I.tolist(), c0.col.tolist())}
    e1 = {(min(pos_new[i], pos_new[j]), max(pos_new[i], pos_new[j]))
          for i, j in zip(c1.I.tolist(), c1.col.tolist())}
    assert c0.n == c1.n and e0 == e1


def test_cli_writes_and_reads_output_bin(tmp_path):
    cli = os.path.join(os.path.dirname(nf.__file__), "nanofiller")
    (tmp_path / "params.txt").write_text(
        "SPAT_DISTR_FILE NONE 3 3 3 10 10 10\nINC_LEN_MED 1.0\nINC_RAD_MED 0.005\n"
        "NUM_RODS 2000\nSTAT_DIST_LIM 0.01\nRNG_SEED 4\nWRT_BIN 1\n")
    r = subprocess.run([cli, "params.txt", "generate", "stat_contacts"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "output.bin").exists() and "dumped into file output.bin" in r.stdout
    line = [l for l in r.stdout.splitlines() if l.startswith("contacts (gap")][0]
    r2 = subprocess.run([cli, "params.txt", "stat_contacts"], cwd=tmp_path, capture_output=True,
                        text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr
    assert "2000 inclusions found" in r2.stdout
    line2 = [l for l in r2.stdout.splitlines() if l.startswith("contacts (gap")][0]
    assert line == line2
