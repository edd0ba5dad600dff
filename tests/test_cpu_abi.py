"""C ABI: exported symbols, parameter defaults, keyword files, validation.

Runs without a GPU: nothing here launches a kernel.
"""
import math
import re
from pathlib import Path

import pytest

import nanofiller_amd as nf

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "nanofiller.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nf_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 20
    L = nf.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(_declared()) <= set(nf.exported_symbols())


def test_defaults_follow_simprms():
    """simPrms::simPrms defaults (src/simPrms.cpp:22-117)."""
    p = nf.default_params()
    assert p.vol_frac == pytest.approx(0.0005)
    assert p.len == 4.0 and p.radius == 0.01 and p.sep == 0.0 and p.soft_frac == 1.0
    assert p.contact_gap == 0.015 and p.max_rounds == 1000
    assert p.ort_law == nf.NF_ORT_ISO and p.den_law == nf.NF_DEN_UNIFORM


def test_normalize_rod_count_and_box():
    """N = ceil(vf * V / (pi a^2 l)); an explicit n_rods wins."""
    p = nf.normalize(nf.default_params(box=(5, 5, 5), len=1.0, radius=0.005, vol_frac=0.012566))
    assert p.n_rods == math.ceil(0.012566 * 125 / (math.pi * 0.005 ** 2 * 1.0))
    q = nf.normalize(nf.default_params(box=(5, 5, 5), len=1.0, radius=0.005, n_rods=20000))
    assert q.n_rods == 20000
    # box edges are representable in float32 after normalisation
    r = nf.normalize(nf.default_params(box=(0.1, 0.3, 0.7), len=0.05, radius=0.001, n_rods=10))
    for a in range(3):
        assert float(__import__("numpy").float32(r.box[a])) == r.box[a]
    assert q.max_batch >= 256


def test_keyword_file(tmp_path):
    """Config 1 written in the reference's keywords, plus the new ones."""
    f = tmp_path / "params.txt"
    f.write_text(
        "# config 1\n"
        "SPAT_DISTR_FILE NONE 5 5 5 10 10 10\n"
        "INC_LEN_MED 1.0\n"
        "INC_RAD_MED 0.005\n"
        "VOLUME_PERCENT 1.2566\n"
        "DEF_THETA_DEV 0\n"
        "DEF_PHI_DEV 0\n"
        "VOL_FRAC_PRECISION 0\n"
        "SELF_ALLIGN_STEP 0\n"
        "STAT_DIST_LIM 0.002\n"
        "RNG_SEED 1 3\n"
        "CELL_SIZE 1.0\n"
        "PRECISION FP64\n"
        "ELECTRODE_AXIS y\n"
        "TUNNEL_LAMBDA 0.0001\n"
        "MATH_TOLERANCE 1e-6\n")
    p = nf.load_params(str(f))
    assert tuple(p.box) == (5.0, 5.0, 5.0)
    assert p.len == 1.0 and p.radius == 0.005 and p.vol_frac == pytest.approx(0.012566)
    assert p.contact_gap == 0.002 and p.seed == 1 and p.stream == 3 and p.cell_size == 1.0
    assert p.precision == nf.NF_PREC_FP64 and p.elec_axis == 1
    assert p.ort_law == nf.NF_ORT_ISO
    n = nf.normalize(p).n_rods
    assert abs(n - 20000) <= 1


def test_keyword_laws(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text("INC_LEN_LAW LOGNORMAL 1.0 0.3 0.3 3.0\nORT_LAW VMF 5 0 0 1\n"
                 "DENSITY_LAW AGGLOM 8 1.5 3.0\nNUM_RODS 500000\nPRECISION FP32\n")
    p = nf.load_params(str(f))
    assert p.len_law == nf.NF_LEN_LOGNORMAL and (p.len, p.len_sigma, p.len_min, p.len_max) == (1.0, 0.3, 0.3, 3.0)
    assert p.ort_law == nf.NF_ORT_VMF and p.kappa == 5.0 and tuple(p.axis) == (0.0, 0.0, 1.0)
    assert p.den_law == nf.NF_DEN_AGGLOM and (p.n_agg, p.agg_sigma, p.agg_peak) == (8, 1.5, 3.0)
    assert p.n_rods == 500000 and p.precision == nf.NF_PREC_FP32_RECHECK


def test_reference_orientation_keywords_map_to_vmf(tmp_path):
    """A positive DEF_THETA_DEV selects a preferred direction: vMF, kappa = 1/dev^2."""
    f = tmp_path / "p.txt"
    f.write_text("DEF_THETA_MED 90\nDEF_PHI_MED 0\nDEF_THETA_DEV 30\n")
    p = nf.load_params(str(f))
    dev = math.radians(30)
    assert p.ort_law == nf.NF_ORT_VMF and p.kappa == pytest.approx(1 / dev ** 2)
    assert p.axis[0] == pytest.approx(1.0) and abs(p.axis[2]) < 1e-12


@pytest.mark.parametrize("text,code", [
    ("INC_LEN_LAW WEIBULL\n", nf.NF_E_INVALID),
    ("ORT_LAW VMF 5\n", nf.NF_E_INVALID),
    ("SPAT_DISTR_FILE dens.dat\n", nf.NF_E_INVALID),
    ("ELECTRODE_AXIS w\n", nf.NF_E_INVALID),
])
def test_keyword_errors(tmp_path, text, code):
    f = tmp_path / "p.txt"
    f.write_text(text)
    with pytest.raises(nf.NFError) as e:
        nf.load_params(str(f))
    assert e.value.code == code
    assert "p.txt:1" in str(e.value)


def test_missing_file():
    with pytest.raises(nf.NFError) as e:
        nf.load_params("/nonexistent/params.txt")
    assert e.value.code == nf.NF_E_IO


@pytest.mark.parametrize("kw", [
    dict(radius=-1.0), dict(len=0.0), dict(elec_axis=3), dict(box=(0, 1, 1)),
    dict(len_law=nf.NF_LEN_LOGNORMAL, len_sigma=0.3, len_min=2.0, len_max=1.0),
    dict(ort_law=nf.NF_ORT_VMF, kappa=0.0), dict(den_law=nf.NF_DEN_AGGLOM, n_agg=0),
    dict(len=40.0, box=(1, 1, 1)),
])
def test_normalize_rejects(kw):
    base = dict(box=(5, 5, 5), n_rods=100, len=1.0, radius=0.005)
    base.update(kw)
    with pytest.raises(nf.NFError) as e:
        nf.normalize(nf.default_params(**base))
    assert e.value.code == nf.NF_E_INVALID


def test_create_without_device_reports_nodevice():
    """In a container without a GPU, creation fails with NF_E_NODEVICE."""
    p = nf.normalize(nf.default_params(box=(2, 2, 2), n_rods=10, len=0.5, radius=0.005))
    import ctypes as C
    h = C.c_void_p()
    rc = nf.lib().nf_create(C.byref(p), C.byref(h))
    if rc == nf.NF_OK:
        nf.lib().nf_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc == nf.NF_E_NODEVICE
    assert b"device" in nf.lib().nf_last_error(None)


def test_version_string():
    v = nf.version()
    assert v and "gfx950" in v or "." in v
