"""SDV TT files straight to and from device memory (GPU).

The reference's own data files (fmex/A.sdv, fmex/b.sdv, copied unchanged to
tests/golden/) are read into HBM, multiplied and measured on the device, and
checked against closed forms: A is the 10-D Dirichlet Laplacian on a 256^10
grid (80 binary modes) and b the all-ones vector, so ||A b|| =
2^33 sqrt(1370), <b, A b> = 2^80 * 10 / 128 and round(A b) keeps the value.
Writing from device memory reproduces the files byte for byte; a float file
goes through the device narrowing and widening exactly as the host casts do.
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import ttgpu
from oracle import sdv as S

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def test_read_multiply_measure_on_device(gpu_ctx):
    nA, rA, dA = ttgpu.tt_read_dev(gpu_ctx, GOLDEN / "fmex_A_d80.sdv")
    nb, rb, db = ttgpu.tt_read_dev(gpu_ctx, GOLDEN / "fmex_b_d80.sdv")
    _, _, _, coreA, _ = S.read(GOLDEN / "fmex_A_d80.sdv")
    assert dA.cpu().numpy().tobytes() == coreA.tobytes()
    m = np.full(80, 2, np.int64)
    ry, dy = ttgpu.mtimes_dev(gpu_ctx, m, m, rA, dA, rb, db)
    want = 2.0 ** 33 * np.sqrt(1370.0)
    got = ttgpu.norm_dev(gpu_ctx, m, ry, dy)
    assert abs(got - want) <= 1e-13 * want, (got, want)
    dot = ttgpu.dot_dev(gpu_ctx, m, rb, db, ry, dy)
    assert abs(dot - 2.0 ** 80 * 10 / 128) <= 1e-13 * 2.0 ** 80 * 10 / 128
    r2, dz = ttgpu.mtimes_round_dev(gpu_ctx, m, m, rA, dA, rb, db, 1e-12)
    got2 = ttgpu.norm_dev(gpu_ctx, m, r2, dz)
    assert abs(got2 - want) <= 1e-12 * want
    print("ranks of round(A b):", sorted(set(r2.tolist())), flush=True)


@pytest.mark.parametrize("name", ["fmex_A_d80.sdv", "fmex_b_d80.sdv"])
def test_device_write_reproduces_file(gpu_ctx, tmp_path, name):
    n, r, dc = ttgpu.tt_read_dev(gpu_ctx, GOLDEN / name)
    ttgpu.tt_save_dev(gpu_ctx, tmp_path / "w.sdv", n, r, dc)
    assert (tmp_path / "w.sdv").read_bytes() == (GOLDEN / name).read_bytes()


def test_device_float_round_trip_multi_chunk(gpu_ctx, tmp_path):
    """A 36 MB core: three 16 MiB staging chunks as doubles, two as floats
    (both pinned buffers cycle)."""
    rng = np.random.default_rng(5)
    n = np.array([16, 64, 64, 16], np.int64)
    r = np.array([1, 16, 256, 256, 1], np.int64)
    ln = int(np.sum(r[:-1] * n * r[1:]))
    core = rng.standard_normal(ln)
    dc = torch.from_numpy(core).to("cuda:0")
    torch.cuda.synchronize()
    for prec in (0, 1):
        ttgpu.tt_save_dev(gpu_ctx, tmp_path / "f.sdv", n, r, dc, precision=prec)
        d, nn, rr, got, dtype = S.read(tmp_path / "f.sdv")
        want = core if prec == 0 else core.astype(np.float32).astype(np.float64)
        assert dtype == 2 * prec and got.tobytes() == want.tobytes()
        n2, r2, d2 = ttgpu.tt_read_dev(gpu_ctx, tmp_path / "f.sdv")
        assert n2.tolist() == n.tolist() and r2.tolist() == r.tolist()
        assert d2.cpu().numpy().tobytes() == want.tobytes()
    assert ln * 8 > 2 * (16 << 20) and ln * 4 > 16 << 20
