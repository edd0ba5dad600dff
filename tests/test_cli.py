"""The command-line front end (reference verbs of src/NanoFiller.cpp:48-149)."""
import os
import subprocess

import numpy as np
import pytest

import nanofiller_amd as nf

CLI = os.path.join(os.path.dirname(nf.__file__), "nanofiller")
PARAMS = ("SPAT_DISTR_FILE NONE 3 3 3 10 10 10\nINC_LEN_MED 1.0\nINC_RAD_MED 0.005\n"
          "NUM_RODS 2000\nSTAT_DIST_LIM 0.01\nRNG_SEED 4\n")


def _run(tmp_path, *verbs):
    (tmp_path / "params.txt").write_text(PARAMS)
    return subprocess.run([CLI, "params.txt", *verbs], cwd=tmp_path, capture_output=True,
                          text=True, timeout=300)


def test_cli_usage_and_missing_state_file(tmp_path):
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    # without generate the rods come from output.bin, as in the reference
    r = _run(tmp_path, "write_vtk")
    assert r.returncode == 1 and "output.bin" in r.stderr


def test_cli_reports_missing_device(tmp_path):
    r = _run(tmp_path, "generate")
    if r.returncode == 0:
        pytest.skip("a GPU is visible")
    assert r.returncode == 1 and "no HIP device" in r.stderr


@pytest.mark.gpu
def test_cli_end_to_end(tmp_path):
    r = _run(tmp_path, "generate", "write_vtk", "stat_phitheta", "stat_mutangle", "stat_contacts", "stat_vdw",
             "stat_conduct", "1", "internal")
    assert r.returncode == 0, r.stderr
    assert "2000 inclusions generated!" in r.stdout
    for f in ("output.vtk", "rho.dat", "gamma.dat", "internal.srt.coo", "source.elec.coo",
              "sink.elec.coo", "Nodes.tbl.dat", "segment.LA.coo", "internal.DS.coo", "sink.DS.coo",
              "source.DS.coo"):
        assert (tmp_path / f).exists(), f
    # same packing as the library call with the same parameters
    p = nf.normalize(nf.load_params(str(tmp_path / "params.txt")))
    with nf.Packing(p) as pk:
        pk.generate()
        rods = pk.rods()
        c = pk.contacts(p.contact_gap)
        h = pk.histograms(bins=32, centre_axis=0, pol_axis=(0, 0, 1), phi_bins=32, theta_bins=32)
    lines = (tmp_path / "output.vtk").read_text().splitlines()
    pts = np.array([list(map(float, l.split())) for l in lines[5:5 + len(rods)]])
    np.testing.assert_allclose(pts, rods[:, :3], rtol=1e-5, atol=1e-6)
    gamma = (tmp_path / "gamma.dat").read_text().replace("gamma={", "").replace("}", "")
    g = np.array([int(x) for x in gamma.replace(",", " ").split()])
    np.testing.assert_array_equal(g, h["mutual"])
    rho = (tmp_path / "rho.dat").read_text().replace("rho = {", "").replace("{", "").replace("}", "")
    rv = np.array([int(x) for x in rho.replace(",", " ").split()])
    np.testing.assert_array_equal(rv, h["phitheta"].ravel())
    assert f"contacts (gap <= 0.01 um): {c.n};" in r.stdout
    assert "effective conductance" in r.stdout
    assert "node network:" in r.stdout


@pytest.mark.gpu
def test_cli_reduce(tmp_path):
    """reduce 1: dead-end pruning along x, reduced.bin with WRT_BIN; the
    pruned packing equals the library's nf_reduce on the same rods."""
    (tmp_path / "params.txt").write_text(PARAMS.replace("STAT_DIST_LIM 0.01", "STAT_DIST_LIM 0.05")
                                         + "WRT_BIN 1\n")
    r = subprocess.run([CLI, "params.txt", "generate", "reduce", "1", "stat_contacts"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    p = nf.normalize(nf.load_params(str(tmp_path / "params.txt")))
    p.elec_axis = 0
    with nf.Packing(p) as pk:
        pk.generate()
        pk.contacts(p.contact_gap)
        info, _ = pk.reduce(2)
        pruned = pk.rods()
    assert f"Inclusions remained: {info['n_after']} " in r.stdout
    assert 0 < info["n_after"] < 2000
    hd = nf.read_state_header(str(tmp_path / "reduced.bin"))
    assert hd["n_rods"] == info["n_after"]
    with nf.Packing(p) as pk:
        assert pk.load_state(str(tmp_path / "reduced.bin")) == info["n_after"]
        back = pk.rods()
    # the state file stores rods by cell: the same set
    key = lambda a: np.lexsort(a.view(np.uint32).T[::-1])
    np.testing.assert_array_equal(back[key(back)].view(np.uint32), pruned[key(pruned)].view(np.uint32))


def test_cli_milestones_are_read(tmp_path):
    """VOLUME_PERCENT_LIST is parsed (count, then percent values)."""
    (tmp_path / "params.txt").write_text(PARAMS + "VOLUME_PERCENT_LIST 3 0.01 0.02 0.03\n")
    assert nf.params_milestones(str(tmp_path / "params.txt")) == [0.01, 0.02, 0.03]
    (tmp_path / "none.txt").write_text(PARAMS)
    assert nf.params_milestones(str(tmp_path / "none.txt")) == []
    (tmp_path / "bad.txt").write_text(PARAMS + "VOLUME_PERCENT_LIST 3 0.01\n")
    with pytest.raises(nf.NFError):
        nf.params_milestones(str(tmp_path / "bad.txt"))


@pytest.mark.gpu
def test_cli_milestones_and_expand(tmp_path):
    """generate with VOLUME_PERCENT_LIST writes state_<vf>.bin dumps; expand 1
    hands the expanded packing to the later verbs (the same as nf_expand)."""
    (tmp_path / "params.txt").write_text(PARAMS + "VOLUME_PERCENT_LIST 2 0.02 0.04\n")
    r = subprocess.run([CLI, "params.txt", "generate", "expand", "1", "stat_contacts"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    dumps = [f for f in os.listdir(tmp_path) if f.startswith("state_")]
    assert dumps
    p = nf.normalize(nf.load_params(str(tmp_path / "params.txt")))
    with nf.Packing(p) as pk:
        pk.generate()
        with pk.expand(1) as ex:
            q = ex.params()
            n = len(ex.rods())
    assert f"new cnt number: {n}" in r.stdout
    assert "New box: %f - %f - %f" % (q.box[0], q.box[1], q.box[2]) in r.stdout
