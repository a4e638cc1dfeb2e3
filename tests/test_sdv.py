"""SDV TT files (fmex/tt_read.c, fmex/tt_write.c) on the CPU.

Fixtures: ``tests/golden/fmex_A_d80.sdv`` and ``fmex_b_d80.sdv`` are the data
files the reference keeps beside its Fortran AMEn test (fmex/A.sdv, b.sdv,
read by fmex/test_amen.f90), copied unchanged.  They pin the format:

* the oracle reader (oracle/sdv.py) parses them; the oracle writer and the
  library's writer (ttg_sdv_write) reproduce both files byte for byte;
* the library's reader (ttg_sdv_read) returns the oracle's arrays exactly;
* the contents satisfy closed forms that fix every value: A is the 10-D
  Dirichlet Laplacian on a 256^10 grid in QTT form (80 binary modes, mode
  size 4 = 2 x 2) and b the all-ones vector, so ||b|| = 2^40,
  ||A b|| = 2^33 sqrt(1370) (A b counts, at each grid point, the coordinates
  that sit on the boundary: E[S^2] = 10 q (1 - q) + (10 q)^2, q = 1/128) and
  <b, A b> = 2^80 * 10 / 128.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import sdv as S
from oracle import tt_oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
FILES = ["fmex_A_d80.sdv", "fmex_b_d80.sdv"]


def _tt(name):
    d, n, r, core, dtype = S.read(GOLDEN / name)
    assert dtype == 0
    return O.TT(n, r, core)


def test_oracle_reader_header_and_closed_forms():
    a, b = _tt(FILES[0]), _tt(FILES[1])
    assert a.d == b.d == 80 and set(a.n.tolist()) == {4} and set(b.n.tolist()) == {2}
    A = O.TTM(np.full(80, 2), np.full(80, 2), a)
    y = O.ttm_mtimes(A, b)
    assert abs(O.tt_norm(b) - 2.0 ** 40) <= 1e-14 * 2.0 ** 40
    want = 2.0 ** 33 * np.sqrt(1370.0)
    assert abs(O.tt_norm(y) - want) <= 1e-13 * want
    assert abs(O.tt_dot(b, y) - 2.0 ** 80 * 10 / 128) <= 1e-13 * 2.0 ** 80 * 10 / 128


@pytest.mark.parametrize("name", FILES)
def test_oracle_writer_reproduces_the_files(tmp_path, name):
    d, n, r, core, _ = S.read(GOLDEN / name)
    S.write(tmp_path / "x.sdv", d, n, r, core, 0)
    assert (tmp_path / "x.sdv").read_bytes() == (GOLDEN / name).read_bytes()


def test_planted_header_mistake_is_caught(tmp_path):
    d, n, r, core, _ = S.read(GOLDEN / FILES[1])
    S.write(tmp_path / "x.sdv", d, n, r, core, 0)
    raw = bytearray((tmp_path / "x.sdv").read_bytes())
    raw[96 + 8:96 + 12] = np.asarray([d + 1], "<i4").tobytes()  # i[2] = d + 1
    assert bytes(raw) != (GOLDEN / FILES[1]).read_bytes()


@pytest.mark.parametrize("name", FILES)
def test_library_reader_and_writer_match(tmp_path, name):
    import ttgpu
    t = ttgpu.tt_read(GOLDEN / name)
    d, n, r, core, _ = S.read(GOLDEN / name)
    assert t.n.tolist() == n.tolist() and t.r.tolist() == r.tolist()
    assert t.core.tobytes() == core.tobytes()
    ttgpu.tt_save(tmp_path / "y.sdv", t)
    assert (tmp_path / "y.sdv").read_bytes() == (GOLDEN / name).read_bytes()


def test_library_float_files_and_errors(tmp_path):
    import ttgpu
    rng = np.random.default_rng(3)
    x = O.rand_tt(rng, [3, 4, 2], [1, 3, 5, 1])
    t = ttgpu.TTTensor(x.n, x.r, x.core)
    ttgpu.tt_save(tmp_path / "f.sdv", t, precision=1)
    d, n, r, core, dtype = S.read(tmp_path / "f.sdv")
    assert dtype == 2 and core.tobytes() == x.core.astype(np.float32).astype(np.float64).tobytes()
    assert ttgpu.sdv_info(tmp_path / "f.sdv") == (3, x.core.size, 2)
    u = ttgpu.tt_read(tmp_path / "f.sdv")
    assert u.core.tobytes() == core.tobytes()
    # not an SDV file, a missing file, a truncated core, a complex core
    (tmp_path / "bad.sdv").write_bytes(b"XX" + bytes(200))
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.tt_read(tmp_path / "bad.sdv")
    assert e.value.status == 1
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.tt_read(tmp_path / "missing.sdv")
    assert e.value.status == 7
    raw = (GOLDEN / FILES[1]).read_bytes()
    (tmp_path / "cut.sdv").write_bytes(raw[:-8])
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.tt_read(tmp_path / "cut.sdv")
    assert e.value.status == 7
    cplx = bytearray(raw)
    cplx[20:24] = np.asarray([1], "<i4").tobytes()  # inf[1] = 1: double complex
    (tmp_path / "c.sdv").write_bytes(bytes(cplx))
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.tt_read(tmp_path / "c.sdv")
    assert e.value.status == 1
