"""Density map law on the GPU: packings bit-identical to the oracle's, alone
and in a batch mixed with the other laws."""
import numpy as np
import pytest

import nanofiller_amd as nf
import oracle

pytestmark = pytest.mark.gpu

# generations whose pair-test count is compared test every pair within reach
EXACT = {nf.NF_OPT_TEST_ALL_PAIRS: 1}


def _mesh(shape, seed):
    rng = np.random.default_rng(seed)
    v = rng.random(shape).astype(np.float32)
    return (0.15 + 0.85 * v).astype(np.float32)


def _p(prec, seed=3, n=4000, shape=(8, 6, 5)):
    return nf.normalize(nf.default_params(box=(6.0, 5.0, 4.0), n_rods=n, len=0.6, radius=0.005,
                                          seed=seed, precision=prec, den_mesh=_mesh(shape, seed)))


def _same(p, rods, st):
    ro, so, _ = oracle.generate(p, threads=16)
    assert rods.shape == ro.shape
    assert np.array_equal(rods.view(np.uint32), ro.view(np.uint32))
    assert st.pair_tests == so["pair_tests"] and st.rounds == so["rounds"]
    assert st.proposals == so["proposals"] and st.candidates == so["candidates"]


@pytest.mark.parametrize("prec", [nf.NF_PREC_FP32_RECHECK, nf.NF_PREC_FP64])
def test_mesh_packing_matches_oracle(prec):
    p = _p(prec)
    with nf.Packing(p, EXACT) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
    assert st.n_accepted == p.n_rods
    _same(p, rods, st)


def test_mesh_from_params_file(tmp_path):
    v = _mesh((4, 4, 4), 9)
    nf.write_density_mesh(str(tmp_path / "mesh.dat"), (5.0, 5.0, 5.0), v)
    (tmp_path / "p.txt").write_text(f"SPAT_DISTR_FILE {tmp_path / 'mesh.dat'}\nNUM_RODS 3000\n"
                                    "INC_LEN_MED 0.5\nINC_RAD_MED 0.005\nRNG_SEED 9\n")
    p = nf.normalize(nf.load_params(str(tmp_path / "p.txt")))
    assert p.den_law == nf.NF_DEN_MESH
    with nf.Packing(p, EXACT) as pk:
        st = pk.generate()[0]
        rods = pk.rods()
    _same(p, rods, st)


def test_mesh_in_mixed_batch():
    ps = [_p(nf.NF_PREC_FP32_RECHECK, seed=4, n=2000),
          nf.normalize(nf.default_params(box=(5, 5, 5), n_rods=2500, len=0.5, radius=0.005, seed=5,
                                         den_law=nf.NF_DEN_AGGLOM, n_agg=3, agg_sigma=1.0,
                                         agg_peak=2.0, stream=1)),
          _p(nf.NF_PREC_FP32_RECHECK, seed=6, n=1500, shape=(3, 7, 2))]
    with nf.Packing(ps, EXACT) as pk:
        sts = pk.generate()
        batch = [pk.rods(i) for i in range(len(ps))]
    for i, p in enumerate(ps):
        with nf.Packing(p, EXACT) as pk:
            pk.generate()
            single = pk.rods()
        assert np.array_equal(batch[i].view(np.uint32), single.view(np.uint32)), i
    _same(ps[0], batch[0], sts[0])
