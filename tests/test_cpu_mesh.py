"""Density map law (SPAT_DISTR_FILE map; read_density src/IO.cpp:11-36,
density2Tex src/den2tex.h:5-53): file format, parameter loading, and the
oracle's interpolation pinned by properties a plausible mistake breaks -
exact values at the mesh points (a half-cell shift fails), linear
interpolation along each axis separately (a transposed axis fails), the
periodic wrap, a constant map - and the acceptance statistics."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle


def _p(mesh, box=(4.0, 5.0, 6.0), **kw):
    base = dict(box=box, n_rods=10, len=0.5, radius=0.005, seed=3)
    base.update(kw)
    p = nf.default_params(**base)
    p.set_density_mesh(mesh)
    return nf.normalize(p)


def test_file_round_trip_and_order(tmp_path):
    v = (np.arange(2 * 3 * 4, dtype=np.float32) / 24.0).reshape(2, 3, 4)
    path = tmp_path / "mesh.dat"
    # the reference's layout written by hand: dims, extents, values with z fastest
    path.write_text("4 5 6\n2 3 4\n" + " ".join(f"{x:.9g}" for x in v.ravel()) + "\n")
    dims, back = nf.read_density_mesh(str(path))
    assert dims == (4.0, 5.0, 6.0) and back.shape == (2, 3, 4)
    np.testing.assert_array_equal(back, v)
    nf.write_density_mesh(str(tmp_path / "m2.dat"), dims, v)
    _, again = nf.read_density_mesh(str(tmp_path / "m2.dat"))
    np.testing.assert_array_equal(again, v)


def test_file_rejects_bad_values(tmp_path):
    path = tmp_path / "mesh.dat"
    path.write_text("1 1 1\n1 1 2\n0.5 1.5\n")
    with pytest.raises(nf.NFError) as e:
        nf.read_density_mesh(str(path))
    assert e.value.code == nf.NF_E_INVALID
    path.write_text("1 1 1\n1 1 3\n0.5 0.5\n")
    with pytest.raises(nf.NFError):
        nf.read_density_mesh(str(path))


def test_params_file_keyword(tmp_path):
    (tmp_path / "mesh.dat").write_text("3 4 5\n1 1 2\n0.25 1\n")
    (tmp_path / "p.txt").write_text(f"SPAT_DISTR_FILE {tmp_path / 'mesh.dat'}\nNUM_RODS 100\n")
    p = nf.normalize(nf.load_params(str(tmp_path / "p.txt")))
    assert p.den_law == nf.NF_DEN_MESH and list(p.den_ext) == [1, 1, 2]
    assert list(p.box) == [3.0, 4.0, 5.0]
    assert [p.den_mesh[0], p.den_mesh[1]] == [0.25, 1.0]


def test_exact_at_mesh_points():
    rng = np.random.default_rng(0)
    v = rng.random((3, 4, 5)).astype(np.float32)
    p = _p(v)
    L = p.box
    for ix in range(3):
        for iy in range(4):
            for iz in range(5):
                x, y, z = ((ix + 0.5) / 3) * L[0], ((iy + 0.5) / 4) * L[1], ((iz + 0.5) / 5) * L[2]
                assert oracle.mesh_density(p, x, y, z) == pytest.approx(float(v[ix, iy, iz]), abs=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_linear_along_each_axis_and_wrap(axis):
    n = 5
    shape = [1, 1, 1]
    shape[axis] = n
    vals = np.array([0.1, 0.9, 0.3, 0.7, 0.5], np.float32)
    v = vals.reshape(shape)
    p = _p(v)
    L = p.box[axis]
    rng = np.random.default_rng(axis)
    for t in rng.random(200):
        pos = [0.3 * p.box[0], 0.4 * p.box[1], 0.6 * p.box[2]]
        pos[axis] = t * L
        s = t * n - 0.5
        i0 = int(np.floor(s))
        f = s - i0
        expect = (1 - f) * float(vals[i0 % n]) + f * float(vals[(i0 + 1) % n])
        assert oracle.mesh_density(p, *pos) == pytest.approx(expect, abs=1e-12)


def test_constant_map_acceptance():
    p = _p(np.full((2, 2, 2), 0.25, np.float32), box=(5, 5, 5), n_rods=2000, len=0.3)
    acc = sum(oracle.candidate(p, 0, s)[0] == 1 for s in range(4000))
    # binomial(4000, 0.25): 1000 +- 27
    assert abs(acc - 1000) < 5 * np.sqrt(4000 * 0.25 * 0.75)


def test_step_map_statistics():
    """Map 1 in the lower half of x, 0.2 in the upper half (mesh points at
    cell centres, so the step is linear between them): the accepted centres'
    density follows the interpolated map."""
    v = np.ones((4, 1, 1), np.float32)
    v[2:] = 0.2
    p = _p(v, box=(8, 4, 4), n_rods=3000, len=0.2)
    rods, st, _ = oracle.generate(p, threads=8)
    x = rods[:, 0]
    edges = np.linspace(0, 8, 9)
    h, _ = np.histogram(x, edges)
    centres = 0.5 * (edges[:-1] + edges[1:])
    # expected weights: the interpolated map averaged over each bin (fine sampling)
    w = []
    for a, b in zip(edges[:-1], edges[1:]):
        xs = np.linspace(a, b, 101)[:-1] + (b - a) / 200
        w.append(np.mean([oracle.mesh_density(p, xx, 1.0, 1.0) for xx in xs]))
    w = np.array(w) / np.sum(w)
    expect = w * len(x)
    chi2 = np.sum((h - expect) ** 2 / expect)
    assert chi2 < 30.0, (h, expect)  # 7 dof; p < 1e-4 at 30
    assert centres[np.argmax(h)] < 4.0


def test_normalize_rejects_bad_maps():
    p = nf.default_params(box=(1, 1, 1), n_rods=10, len=0.2)
    p.set_density_mesh(np.full((1, 1, 2), 1.5, np.float32))
    with pytest.raises(nf.NFError):
        nf.normalize(p)
    p.set_density_mesh(np.zeros((1, 1, 2), np.float32))
    with pytest.raises(nf.NFError):
        nf.normalize(p)
