"""Write the small golden fixtures under tests/golden/.

1. ``ttsolve_<name>.npz``: the reference's own known-answer data,
   ``tests/tt-solve/tt-solve_<name>.mat`` (A as tt_matrix, x and rhs as
   tt_tensor; the README there states A*x ~= rhs).  Read as plain data with
   scipy.io.loadmat; nothing from the reference is executed.
2. ``oracle_cfg*.npz``: reduced-size inputs of the benchmark workloads with
   the FP64 oracle's outputs (ranks, cores, per-bond singular values).

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import tt_oracle as O  # noqa: E402

REF = Path("/root/reference/tests/tt-solve")
OUT = Path(__file__).resolve().parent


def _ints(v):
    return np.asarray(v, dtype=np.float64).ravel().astype(np.int64)


def convert_ttsolve(name):
    import scipy.io as sio
    m = sio.loadmat(str(REF / f"tt-solve_{name}.mat"), struct_as_record=False, squeeze_me=True)
    A, x, rhs = m["A"], m["x"], m["rhs"]
    tt = A.tt
    out = {
        "A_n": _ints(A.n), "A_m": _ints(A.m), "A_r": _ints(tt.r), "A_core": np.asarray(tt.core, dtype=np.float64).ravel(),
        "x_n": _ints(x.n), "x_r": _ints(x.r), "x_core": np.asarray(x.core, dtype=np.float64).ravel(),
        "rhs_n": _ints(rhs.n), "rhs_r": _ints(rhs.r), "rhs_core": np.asarray(rhs.core, dtype=np.float64).ravel(),
    }
    np.savez_compressed(OUT / f"ttsolve_{name}.npz", **out)
    print(name, {k: v.shape for k, v in out.items()})


def oracle_vectors():
    c1 = O.make_config(1, d=8)
    info = {}
    y1 = O.tt_round(c1["z"], c1["eps"], info=info)
    np.savez_compressed(OUT / "oracle_cfg1_d8.npz", n=c1["z"].n, r=c1["z"].r, core=c1["z"].core,
                        eps=c1["eps"], out_r=y1.r, out_core=y1.core)
    c2 = O.make_config(2, d=6, rank=16, seed=7)
    y2 = O.tt_round(O.ttm_mtimes(c2["a"], c2["x"]), c2["eps"])
    np.savez_compressed(OUT / "oracle_cfg2_d6.npz", a_n=c2["a"].n, a_m=c2["a"].m, a_r=c2["a"].tt.r,
                        a_core=c2["a"].tt.core, x_n=c2["x"].n, x_r=c2["x"].r, x_core=c2["x"].core,
                        eps=c2["eps"], out_r=y2.r, out_core=y2.core)
    c3 = O.make_config(3, d=8, n=4, rank=24, seed=21)
    info3 = {}
    y3 = O.tt_round(c3["x"], 1e-3, info=info3)
    sv = np.concatenate([s for s in info3["sv"] if s is not None])
    np.savez_compressed(OUT / "oracle_cfg3_d8.npz", n=c3["x"].n, r=c3["x"].r, core=c3["x"].core,
                        eps=1e-3, out_r=y3.r, out_core=y3.core, sv=sv)
    print("oracle vectors written")


if __name__ == "__main__":
    for nm in ["ex2_6-10", "ex6"]:
        convert_ttsolve(nm)
    oracle_vectors()
