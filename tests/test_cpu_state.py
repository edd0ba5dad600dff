"""State file header parsing (host code only; no device needed)."""
import numpy as np
import pytest

import nanofiller_amd as nf
from tests import state_format


def test_header_of_a_reference_layout_file(tmp_path):
    path = str(tmp_path / "s.bin")
    cells = {0: [[0.1, 0.2, 0.3, 1, 0, 0, 1.0, 0.005]], 5: [[1.9, 0.2, 1.7, 0, 1, 0, 0.5, 0.005]] * 2}
    state_format.write(path, (2.0, 2.0, 2.0), (2, 2, 2), 1.0, 0.005, cells, cap=[3] * 8)
    hd = nf.read_state_header(path)
    assert hd["ext"] == (2, 2, 2) and hd["slots"] == 24 and hd["n_rods"] == 3
    assert hd["box"] == pytest.approx((2.0, 2.0, 2.0)) and hd["edge"] == pytest.approx((1.0, 1.0, 1.0))
    assert hd["len"] == pytest.approx(1.0) and hd["radius"] == pytest.approx(0.005)
    assert hd["tol"] == pytest.approx(1e-6)


def test_header_rejects_truncated_and_missing_files(tmp_path):
    path = tmp_path / "s.bin"
    state_format.write(str(path), (1.0, 1.0, 1.0), (1, 1, 1), 1.0, 0.005, {0: [[0.5] * 3 + [0, 0, 1, 0.5, 0.005]]})
    b = path.read_bytes()
    path.write_bytes(b[:-4])
    with pytest.raises(nf.NFError) as e:
        nf.read_state_header(str(path))
    assert e.value.code == nf.NF_E_INVALID and "does not match" in str(e.value)
    path.write_bytes(b[:20])
    with pytest.raises(nf.NFError) as e:
        nf.read_state_header(str(path))
    assert e.value.code == nf.NF_E_INVALID
    with pytest.raises(nf.NFError) as e:
        nf.read_state_header(str(tmp_path / "missing.bin"))
    assert e.value.code == nf.NF_E_IO


def test_struct_layout_round_trip(tmp_path):
    # the fixture writer and reader agree (they pin the layout the library
    # is checked against on the device)
    path = str(tmp_path / "s.bin")
    rng = np.random.default_rng(1)
    cells = {c: [list(rng.random(8).astype(np.float32)) for _ in range(c % 3)] for c in range(27)}
    state_format.write(path, (3.0, 3.0, 3.0), (3, 3, 3), 1.0, 0.005, cells)
    s = state_format.read(path)
    assert s["n_rods"] == sum(len(v) for v in cells.values()) == s["slots"]
    assert np.array_equal(np.cumsum(s["occ"]), s["addr"])
