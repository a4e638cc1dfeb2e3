"""CPU restatement of the SDV TT file format (FP64 checker).

TEST INFRASTRUCTURE ONLY: only ``tests/`` may import this module; the
product path (``libttgpu.so`` and the ``ttgpu`` package) never calls it.

Follows ``fmex/tt_read.c:13-103`` (reader) and ``fmex/tt_write.c:21-79`` with
the header set up by its ``mexFunction`` (``tt_write.c:97-136``):

* 128-byte header ``{char txt[8]; int ver[2]; int inf[4]; char comment[64];
  int i[8]}``; the writer sets ``txt = "TT      "``, ``ver = {1, 0}``,
  ``inf = {128, dtype, 0, 0}``, ``i = {1, 0, d, 0, ...}``, comment zeros;
* ``int l, m`` with ``d = m - l + 1`` (the writer stores ``1, d``);
* ``int n[d]``, ``int r[d+1]``;
* the flat core, ``dtype`` 0 double, 1 double complex, 2 float, 3 float
  complex (``tt_read.c:71-98``); the precision argument of the writer maps
  real data to ``dtype = 0 + 2*precision`` (``tt_write.c:167``).
"""
from __future__ import annotations

import numpy as np

_HEAD = np.dtype([("txt", "S8"), ("ver", "<i4", 2), ("inf", "<i4", 4), ("comment", "S64"), ("i", "<i4", 8)])
_CORE = {0: "<f8", 1: "<c16", 2: "<f4", 3: "<c8"}


def read(path):
    """``tt_read.c:29-103``: returns ``(d, n, r, core, dtype)`` (core widened to float64/complex128)."""
    b = open(path, "rb").read()
    h = np.frombuffer(b[:128], dtype=_HEAD)[0]
    if not bytes(h["txt"]).startswith(b"TT"):
        raise ValueError("Not a SDV TT file")
    if int(h["ver"][0]) != 1:
        raise ValueError("Wrong version of SDV TT file")
    dtype = int(h["inf"][1])
    lm = np.frombuffer(b[128:136], "<i4")
    d = int(lm[1] - lm[0] + 1)
    o = 136
    n = np.frombuffer(b[o:o + 4 * d], "<i4").astype(np.int64)
    o += 4 * d
    r = np.frombuffer(b[o:o + 4 * (d + 1)], "<i4").astype(np.int64)
    o += 4 * (d + 1)
    size = int(np.sum(r[:-1] * n * r[1:]))
    dt = np.dtype(_CORE[dtype])
    core = np.frombuffer(b[o:o + size * dt.itemsize], dt)
    if core.size != size:
        raise ValueError("truncated SDV file")
    core = core.astype(np.complex128 if dtype in (1, 3) else np.float64)
    return d, n, r, core, dtype


def write(path, d, n, r, core, precision=0):
    """``tt_write.c:21-79`` with the header of ``:97-136`` (real cores)."""
    h = np.zeros((), dtype=_HEAD)
    h["txt"] = b"TT      "
    h["ver"] = [1, 0]
    h["inf"] = [128, 2 * precision, 0, 0]
    h["i"] = [1, 0, d, 0, 0, 0, 0, 0]
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(np.asarray([1, d], "<i4").tobytes())
        f.write(np.asarray(n, "<i4").tobytes())
        f.write(np.asarray(r, "<i4").tobytes())
        f.write(np.asarray(core, "<f8" if precision == 0 else "<f4").tobytes())
