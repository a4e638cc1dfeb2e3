"""CPU checks of the C ABI library and the host package (no GPU compute calls).

* libttgpu.so loads and exports every function declared in include/*.h;
* the ctypes signature table covers exactly those symbols;
* host-only entry points (version, capacity) work and a context cannot be
  created without a gfx950 device (no silent fallback);
* the seeded workload generators of the host package produce the oracle's
  inputs bit for bit (small sizes).
"""
import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT

from ttgpu import _lib
from ttgpu import workloads as W
from oracle import tt_oracle as O

HEADERS = [ROOT / "include" / "ttgpu.h", ROOT / "include" / "ttgpu_testing.h"]


def _declared(path):
    text = path.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(ttg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared(HEADERS[0]) + _declared(HEADERS[1])
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_tables_match_headers():
    public = _declared(HEADERS[0])
    testing = _declared(HEADERS[1])
    assert sorted(_lib.EXPORTS) == public
    assert sorted(_lib.TEST_EXPORTS) == testing


def test_version_and_capacity_are_host_only():
    lib = _lib.load()
    assert lib.ttg_version().startswith(b"ttgpu")
    n = [8] * 32
    r = [1] + [256] * 31 + [1]
    cap = C.c_int64()
    assert lib.ttg_round_capacity(32, _lib.i64(n), _lib.i64(r), C.byref(cap)) == _lib.TTG_OK
    # QR caps of round.m:48 -> ranks [1,8,64,256 x 27,64,8,1]
    rq = list(r)
    for i in range(31):
        rq[i + 1] = min(rq[i] * n[i], rq[i + 1])
    assert cap.value == sum(rq[k] * n[k] * rq[k + 1] for k in range(32))
    bad = C.c_int64()
    assert lib.ttg_round_capacity(2, _lib.i64([4, 0]), _lib.i64([1, 3, 1]), C.byref(bad)) == _lib.TTG_BAD_ARG
    assert lib.ttg_round_capacity(2, _lib.i64([4, 4]), _lib.i64([1, 0, 1]), C.byref(bad)) == _lib.TTG_BAD_ARG


def test_no_context_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.ttg_ctx_create(0, None, C.byref(h)) == _lib.TTG_NO_DEVICE


def _same(n, r, core, tt):
    assert list(map(int, n)) == list(map(int, tt.n))
    assert list(map(int, r)) == list(map(int, tt.r))
    np.testing.assert_array_equal(core, tt.core)


def test_workloads_match_oracle_configs():
    c1 = W.round1(d=6)
    _same(*c1["z"], O.make_config(1, d=6)["z"])
    c2 = W.matvec2(d=5, rank=8)
    o2 = O.make_config(2, d=5, rank=8)
    _same(*c2["x"], o2["x"])
    na, ma, ra, ca = c2["a"]
    assert list(ra) == list(o2["a"].r) and list(na) == list(o2["a"].n)
    np.testing.assert_array_equal(ca, o2["a"].tt.core)
    c3 = W.mixed3(d=5, n=4, rank=6)
    _same(*c3["x"], O.make_config(3, d=5, n=4, rank=6)["x"])
