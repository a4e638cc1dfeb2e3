"""Batched rounding of uneven TT tensors (config-4 recipe at reduced sizes)
against the oracle, tensor by tensor: identical ranks and relative Frobenius
difference <= 1e-12 (plain form)."""
import numpy as np
import pytest
import torch

import ttgpu
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count,dmax,rmax", [(6, 6, 8), (24, 10, 24)])
def test_batched_matches_oracle(gpu_ctx, count, dmax, rmax):
    batch = O.make_batch(count=count, seed=40 + count, dmin=3, dmax=dmax, nset=(2, 4, 8), rmin=2, rmax=rmax)
    eps = 1e-8
    cores = [t.core for t in batch]
    cofs = np.concatenate([[0], np.cumsum([c.size for c in cores])])
    dcore = torch.tensor(np.concatenate(cores), device="cuda:0")
    res = ttgpu.round_batched_dev(gpu_ctx, [t.n for t in batch], [t.r for t in batch], dcore,
                                  cofs[:-1], eps, precision="fp64")
    assert len(res) == count
    for t, (r_t, y_t) in zip(batch, res):
        ref = O.tt_round(t, eps)
        assert list(map(int, r_t)) == ref.r.tolist()
        y = O.TT(t.n, r_t, y_t.cpu().numpy())
        assert O.tt_rel_diff(y, ref) <= 1e-12
