#!/usr/bin/env python3
"""TT rounding benchmark on one MI355X: d=32, n=8, r=256 TT (sum of two rank-128 TTs
with graded core spectra), eps=1e-6.

One step = one ``round(X, 1e-6)`` of that tensor with its cores already resident
in HBM, in the mixed form (FP32 MFMA Gram products on the first sweep; FP64
storage, solves, contractions, r x r Cholesky and Jacobi SVD).  The FP64 plain
form is timed on the same input for the speedup, both results are compared with
the FP64 CPU oracle (the restatement of @tt_tensor/round.m), and the oracle
itself is timed on the host cores as the CPU baseline.  Bonds whose smallest
singular value provably clears the truncation threshold skip their SVD (same
tensor, same ranks); the rounds are also timed with an SVD at every bond
(``with_svd_at_every_bond``).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--plain-steps P] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "tt-toolbox_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

PEAK = {  # MI355X published dense peaks
    "fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12,
}
CLASS_NAMES = ["gemm", "jacobi", "cholesky", "trtri", "householder"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _rate(fn, reps, sync):
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    sync()
    return reps / (time.perf_counter() - t0)


def round_flops(n, r_in, r_out):
    """Algorithmic flops of round.m on a tensor with input ranks r_in and
    output ranks r_out: sweep 1 QR (4mk^2 - 4k^3/3) and R * core_{i+1}
    (2 k r n r'), sweep 2 SVD (6MN^2 + 20N^3) and core_{i-1} * U S."""
    d = len(n)
    n = [int(v) for v in n]
    r = [int(v) for v in r_in]
    rq = list(r)
    for i in range(d - 1):
        rq[i + 1] = min(rq[i] * n[i], rq[i + 1])
    ro = [int(v) for v in r_out]
    f = 0.0
    for i in range(d - 1):
        m_, k_ = rq[i] * n[i], rq[i + 1]
        if m_ >= k_:
            f += 4 * m_ * k_ * k_ - 4 * k_ ** 3 / 3
        f += 2 * k_ * r[i + 1] * n[i + 1] * r[i + 2]
    for i in range(d - 1, 0, -1):
        M_, N_ = max(rq[i], n[i] * ro[i + 1]), min(rq[i], n[i] * ro[i + 1])
        f += 6 * M_ * N_ ** 2 + 20 * N_ ** 3
        f += 2 * rq[i - 1] * n[i - 1] * rq[i] * ro[i]
    return f


def roofline(work, seconds, bound="mfma", dtype="fp64", what="whole call"):
    """Achieved rate of `work` (flops, or bytes for bound 'hbm') per call of
    `seconds`, against the MI355X peak."""
    if bound == "hbm":
        a = work / seconds / 1e9
        return {"bound": "hbm", "achieved": a, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s",
                "frac": a * 1e9 / PEAK["hbm"], "traffic": None, "scope": what}
    a = work / seconds / 1e12
    return {"bound": "mfma", "achieved": a, "peak": PEAK[dtype] / 1e12, "unit": "TFLOP/s",
            "frac": a * 1e12 / PEAK[dtype], "traffic": None, "scope": what}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def mtimes_leg(ttgpu, W, ctx, dev, fused=True):
    """Config 2: A*x (d=12, n=m=4, rank_A 5, rank_x 64), unfused, and the fused
    round(A*x, 1e-6); inputs resident in HBM, host-timed back-to-back calls."""
    import torch
    sync = torch.cuda.synchronize
    out = {}
    c2 = W.matvec2()
    na, ma, ra, acore = c2["a"]
    nx, rx, xcore = c2["x"]
    da = torch.from_numpy(acore).to(dev)
    dx = torch.from_numpy(xcore).to(dev)
    ry = ra * rx
    ysize = int(sum(ry[k] * na[k] * ry[k + 1] for k in range(len(na))))
    oy = torch.empty(ysize, dtype=torch.float64, device=dev)
    rate = _rate(lambda: ttgpu.mtimes_dev(ctx, na, ma, ra, da, rx, dx, out=oy), 50, sync)
    byt = 8.0 * (ysize + acore.size + xcore.size)
    # the product-core kernel's own duration (HIP events passed to its launch)
    ctx.profile_begin("mtimes")
    for _ in range(20):
        ttgpu.mtimes_dev(ctx, na, ma, ra, da, rx, dx, out=oy)
    prof = ctx.profile_end()
    kern_s = prof["ms"] / max(prof["launches"], 1) / 1e3
    out["config2_mtimes"] = {"workload": "A*x, d=12 n=m=4 rank_A=5 rank_x=64 -> rank 320, FP64, unfused",
                             "contractions_per_s": rate, "hbm_bytes_per_contraction": byt,
                             "roofline": roofline(byt, 1.0 / rate, "hbm",
                                                  what="whole call incl. host launch overhead"),
                             "kernel": {"name": "k_mtimes_all", "launches_per_contraction": 1,
                                        "avg_launch_us": kern_s * 1e6,
                                        "roofline": roofline(byt, kern_s, "hbm", what="k_mtimes_all launch")}}
    if not fused:
        return out
    o2 = torch.empty(ttgpu.capacity(na, ry), dtype=torch.float64, device=dev)
    d2 = len(na)
    cflops = sum(2.0 * ma[k] * ry[k] * na[k] * ry[k + 1] for k in range(d2))  # product cores
    # default route: the product (capped ranks <= 480) is formed once and
    # rounded by the Gram-chain form; TTG_GRAM=0 times the sweeps with every
    # product core built on chip inside the first sweep's products
    for key, prec, env, what in [
            ("config2_fused_mtimes_round", "fp64", None,
             "round(A*x, 1e-6), plain FP64: product formed once, Gram-chain round"),
            ("config2_fused_mtimes_round_mixed", "fp32", None,
             "round(A*x, 1e-6), mixed form: product formed once, Gram-chain round"),
            ("config2_fused_sweeps", "fp64", "0",
             "round(A*x, 1e-6), plain FP64, QR/SVD sweeps with the product cores built on chip (never in HBM)")]:
        rr, info = {}, {}
        if env is not None:
            os.environ["TTG_GRAM"] = env
        try:
            rate = _rate(lambda: rr.__setitem__("r", ttgpu.mtimes_round_dev(
                ctx, na, ma, ra, da, rx, dx, 1e-6, out=o2, precision=prec, info=info)[0]), 5, sync)
        finally:
            os.environ.pop("TTG_GRAM", None)
        fl = round_flops(na, ry, rr["r"]) + cflops
        out[key] = {"workload": what, "rounds_per_s": rate, "ranks_out": [int(v) for v in rr["r"]],
                    "path": info.get("path"), "svd_skipped": info.get("svd_skipped"),
                    "jacobi_sweeps": info.get("jacobi_sweeps"),
                    "algorithmic_gflop": fl / 1e9, "roofline": roofline(fl, 1.0 / rate)}
    return out


def completion_leg(ttgpu, W, dev):
    """Config 5: completion loss 1/2||P(X) - y||^2 and all core gradients
    (65,536 samples, d=16, n=8, model rank 32), FP32 and FP64, inputs in HBM."""
    import torch
    from ttgpu import autograd as AG
    sync = torch.cuda.synchronize
    out = {}
    c5 = W.completion5()
    n5, r5, core5 = c5["model"]
    idx = torch.from_numpy(c5["idx"]).to(dev)
    vals = torch.from_numpy(c5["vals"]).to(dev)
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        core = torch.tensor(core5, dtype=dt, device=dev, requires_grad=True)

        def step():
            loss = AG.completion_loss(core, n5, r5, idx, vals)
            torch.autograd.grad(loss, core)
        rate5 = _rate(step, 20, sync)
        S = int(c5["idx"].shape[0])
        # per sample: left chain, right chain and the gradient outer products, 2 r_k r_{k+1} flops each
        fl5 = 6.0 * S * sum(int(r5[k]) * int(r5[k + 1]) for k in range(len(n5)))
        out[f"config5_completion_grad_{name}"] = {
            "workload": "loss 1/2||P(X) - y||^2 and all core gradients, 65,536 samples, d=16 n=8 rank 32",
            "evaluations_per_s": rate5, "algorithmic_gflop": fl5 / 1e9,
            "roofline": roofline(fl5, 1.0 / rate5, dtype=name)}
    return out


def other_configs(ttgpu, W, ctx, dev):
    """Configs 1, 2 and 5 (and the <y, A x> gradient at config-1 shapes), one GPU,
    inputs resident in HBM; rates are host-timed over back-to-back calls."""
    import numpy as np
    import torch
    from ttgpu import autograd as AG
    sync = torch.cuda.synchronize
    out = {}
    # config 1: round(X + 0.5 X, 1e-10), d=20 n=4 r=40 -> 20
    c1 = W.round1()
    n1, r1, core1 = c1["z"]
    d1 = torch.from_numpy(core1).to(dev)
    o1 = torch.empty(ttgpu.capacity(n1, r1), dtype=torch.float64, device=dev)
    ro = {}
    rate = _rate(lambda: ro.__setitem__("r", ttgpu.round_dev(ctx, n1, r1, d1, c1["eps"], out=o1)[0]), 20, sync)
    fl1 = round_flops(n1, r1, ro["r"])
    out["config1_round"] = {"workload": "round(X + 0.5 X, 1e-10), d=20 n=4 r=40, FP64", "rounds_per_s": rate,
                            "ranks_out_max": int(max(ro["r"])), "algorithmic_gflop": fl1 / 1e9,
                            "roofline": roofline(fl1, 1.0 / rate)}
    # the same workload from its two summands: a X + b Y fused into the round
    nx1, rx1, cx1 = c1["x"]
    dx1 = torch.from_numpy(cx1).to(dev)
    rate = _rate(lambda: ttgpu.plus_round_dev(ctx, nx1, rx1, dx1, rx1, dx1, c1["eps"], 1.0, 0.5, out=o1), 20, sync)
    out["config1_plus_round"] = {"workload": "round(X + 0.5 X, 1e-10) from X (d=20 n=4 r=20), sum fused, FP64",
                                 "rounds_per_s": rate, "roofline": roofline(fl1, 1.0 / rate)}
    out.update(mtimes_leg(ttgpu, W, ctx, dev))
    out.update(completion_leg(ttgpu, W, dev))
    # <y, A x> and its gradients at config-1 shapes
    cb = W.bilinear1()
    nb, mb, rab, acb = cb["a"]
    _, rxb, xcb = cb["x"]
    _, ryb, ycb = cb["y"]
    ac = torch.from_numpy(acb).to(dev)
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        yc = torch.tensor(ycb, dtype=dt, device=dev, requires_grad=True)
        xc = torch.tensor(xcb, dtype=dt, device=dev, requires_grad=True)
        acc_ = ac.to(dt)

        def stepb():
            v = AG.bilinear(yc, xc, acc_, nb, mb, ryb, rab, rxb)
            torch.autograd.grad(v, (yc, xc))
        out[f"bilinear_grad_{name}"] = {"workload": "<y, A x> and its x, y gradients, d=20 n=m=4, ranks 20/5/20",
                                        "evaluations_per_s": _rate(stepb, 20, sync)}
    return out


def verify_batch(tens, out, eps, oracle_count=2):
    """Checks the batched output before its rate is reported: for every tensor
    x + x the ranks equal the generic TT ranks of x (closed form), the cores are
    balanced (C_k C_k^T = nrm0^2 I, k >= 2) and nrm0^d = ||x + x|| = 2 ||x||;
    the tensors of most QR work are also rounded by the FP64 oracle (ranks
    identical, relative difference <= 1e-12)."""
    import numpy as np
    from oracle import tt_oracle as O
    bad, worst_bal, worst_nrm = [], 0.0, 0.0
    works = []
    for i, ((n, r, core), (r_t, y_t)) in enumerate(zip(tens, out)):
        z = O.TT(n, r, core)
        rx = [1] + [int(v) // 2 for v in r[1:-1]] + [1]
        x = O.TT.from_cores([np.asfortranarray(z.core_k(k)[:rx[k], :, :rx[k + 1]]) for k in range(z.d)])
        y = O.TT(n, r_t, y_t)
        nrm0, bres = O.balance_residual(y)
        nz = 2.0 * O.tt_norm_gram(x)
        rel = abs(nrm0 ** len(n) - nz) / nz
        worst_bal, worst_nrm = max(worst_bal, bres), max(worst_nrm, rel)
        if r_t.tolist() != O.generic_ranks(n, rx) or bres > 1e-12 or rel > 1e-10:
            bad.append(i)
        c, rk = 0.0, float(r[0])
        for j in range(len(n) - 1):
            m = rk * n[j]
            k = min(m, float(r[j + 1]))
            c += m * k * k
            rk = k
        works.append(c)
    worst_ref = 0.0
    for j in np.argsort(works)[::-1][:oracle_count]:
        n, r, core = tens[j]
        ref = O.tt_round(O.TT(n, r, core), eps)
        r_t, y_t = out[j]
        err = O.tt_rel_diff(O.TT(n, r_t, y_t), ref) if r_t.tolist() == ref.r.tolist() else float("inf")
        worst_ref = max(worst_ref, err)
        if err > 1e-12:
            bad.append(int(j))
    return {"ok": not bad, "failed_tensors": bad[:16], "tensors_checked": len(out),
            "max_balance_residual": worst_bal, "max_rel_norm_error": worst_nrm,
            "oracle_tensors": oracle_count, "max_rel_diff_vs_oracle": worst_ref}


def extra_legs(ttgpu, W, ctx, dev, c, dcore, outs):
    """On the d=32, n=8, r=256 tensor: the truncating round (eps 0.1; at
    1e-3 and 1e-2 every singular value of this tensor clears the threshold),
    the round of the sum from its two rank-128 summands (sum fused), norm, dot
    and the difference norm of the two forms' results; and the TT-SVD of a
    16.7M-entry full tensor.  Host-timed back-to-back calls, inputs in HBM."""
    import numpy as np
    import torch
    sync = torch.cuda.synchronize
    n, r, _ = c["x"]
    out = {}
    o = torch.empty(ttgpu.capacity(n, r), dtype=torch.float64, device=dev)
    for prec in ("fp32", "fp64"):
        rr, info = {}, {}

        def trunc():
            rr["r"] = ttgpu.round_dev(ctx, n, r, dcore, 0.1, precision=prec, out=o, info=info)[0]
        rate = _rate(trunc, 2, sync)
        fl = round_flops(n, r, rr["r"])
        out[f"truncating_round_{prec}"] = {
            "workload": "round d=32 n=8 r=256 eps=0.1 (bonds truncate: QR/SVD sweeps, Jacobi SVD at every bond)",
            "rounds_per_s": rate, "ranks_out_max": int(max(rr["r"])), "path": info.get("path"),
            "jacobi_sweeps": info.get("jacobi_sweeps"), "algorithmic_gflop": fl / 1e9,
            "roofline": roofline(fl, 1.0 / rate)}
        log(f"truncating round {prec}: {rate:.2f}/s, ranks max {max(rr['r'])}")
    na, ra, ca = c["xa"]
    nb, rb, cb = c["xb"]
    da = torch.from_numpy(ca).to(dev)
    db = torch.from_numpy(cb).to(dev)
    sync()
    for prec in ("fp32", "fp64"):
        info = {}
        rate = _rate(lambda: ttgpu.plus_round_dev(ctx, na, ra, da, rb, db, c["eps"], precision=prec, out=o,
                                                  info=info), 5, sync)
        out[f"plus_round_{prec}"] = {"workload": "round(Xa + Xb, 1e-6) from the two rank-128 summands "
                                                 "(sum assembled on the device, then the round)",
                                     "rounds_per_s": rate, "path": info.get("path")}
    nrm = {}
    rate = _rate(lambda: nrm.__setitem__("v", ttgpu.norm_dev(ctx, n, r, dcore)), 5, sync)
    out["norm"] = {"workload": "norm(X), d=32 n=8 r=256 (left QR sweep, norm.m)", "per_s": rate}
    rate = _rate(lambda: ttgpu.dot_dev(ctx, n, r, dcore, r, dcore), 5, sync)
    qk = sum(2.0 * int(r[k]) * int(n[k]) * int(r[k + 1]) * (int(r[k]) + int(r[k + 1])) for k in range(len(n)))
    out["dot"] = {"workload": "dot(X, X), d=32 n=8 r=256 (interface chain, dot.m)", "per_s": rate,
                  "roofline": roofline(qk, 1.0 / rate)}
    (rm, ym), (rp, yp) = outs
    dn = {}
    rate = _rate(lambda: dn.__setitem__("v", ttgpu.plus_norm_dev(ctx, n, rm, ym, rp, yp)), 3, sync)
    out["difference_norm"] = {"workload": "norm(Y_mixed - Y_plain) of the two rounded results (rank 256 + 256, "
                                          "QR sweep over the fused difference)",
                              "per_s": rate, "value_relative": dn["v"] / nrm["v"]}
    # TT-SVD: a 4^12 tensor (16.7M entries) of TT rank 6 plus 1e-9 dense noise, eps 1e-6
    rng = np.random.default_rng(12)
    nn = [4] * 12
    cores = W.rand_cores(rng, nn, [1, 4] + [6] * 9 + [4, 1])
    full = cores[0][0]  # (n_1, r_1); rows in column-major order of the indices so far
    for k in range(1, 12):
        rk, nk, rk1 = cores[k].shape
        t = (full @ cores[k].reshape(rk, nk * rk1)).reshape(-1, nk, rk1)
        full = t.transpose(1, 0, 2).reshape(-1, rk1)  # index prev + N_prev * i_k
    full = np.ascontiguousarray(full[:, 0])
    full = full + 1e-9 * np.linalg.norm(full) / np.sqrt(full.size) * rng.standard_normal(full.size)
    dfull = torch.from_numpy(full).to(dev)
    sync()
    rs = {}
    rate = _rate(lambda: rs.__setitem__("r", ttgpu.full_to_tt_dev(ctx, nn, dfull, 1e-6)[0]), 3, sync)
    out["tt_svd"] = {"workload": "TT-SVD of a 4^12 full tensor (16.7M entries, TT rank 6 + 1e-9 noise), eps 1e-6",
                     "per_s": rate, "ranks_out": [int(v) for v in rs["r"]],
                     "roofline": roofline(8.0 * full.size, 1.0 / rate, "hbm",
                                          what="whole call vs one read of the full tensor")}
    out["_tt_svd_input"] = (nn, full)
    del dfull, da, db, o
    torch.cuda.empty_cache()
    return out


def batched_leg(ttgpu, W, ctx, dev, count):
    """Config 4 recipe: one call for the whole batch (the first call on a context
    also tunes the execution) against a per-tensor loop of round_dev on one stream
    (32-tensor sample, scaled by QR work)."""
    import numpy as np
    import torch
    log(f"generating {count} uneven tensors (config 4 recipe)")
    bb = W.batch4(count=count)
    tens = bb["tensors"]
    cofs = np.concatenate([[0], np.cumsum([t[2].size for t in tens])])
    dall = torch.empty(int(cofs[-1]), dtype=torch.float64, device=dev)
    for i, t in enumerate(tens):
        dall[int(cofs[i]):int(cofs[i + 1])].copy_(torch.from_numpy(t[2]))
    ns, rs = [t[0] for t in tens], [t[1] for t in tens]
    torch.cuda.synchronize()
    binfo = {}
    t0 = time.perf_counter()
    ttgpu.round_batched_dev(ctx, ns, rs, dall, cofs[:-1], bb["eps"], info=binfo)  # first call tunes
    t1 = time.perf_counter()
    res = ttgpu.round_batched_dev(ctx, ns, rs, dall, cofs[:-1], bb["eps"], info=binfo)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    check = verify_batch(tens, [(np.asarray(r_t), y_t.cpu().numpy()) for r_t, y_t in res], bb["eps"])
    batch_flops = sum(round_flops(n_, r_, r_t) for n_, r_, (r_t, _) in zip(ns, rs, res))
    del res
    log("batched output check:", check)
    m = min(32, len(tens))
    for i in range(m):
        ttgpu.round_dev(ctx, ns[i], rs[i], dall[int(cofs[i]):int(cofs[i + 1])], bb["eps"])
    torch.cuda.synchronize()
    t3 = time.perf_counter()

    def qr_work(n_, r_):
        c, rk = 0.0, float(r_[0])
        for j in range(len(n_) - 1):
            mm = rk * n_[j]
            kk = min(mm, float(r_[j + 1]))
            c += mm * kk * kk
            rk = kk
        return c
    wall = sum(qr_work(n_, r_) for n_, r_ in zip(ns, rs))
    wsamp = sum(qr_work(n_, r_) for n_, r_ in zip(ns[:m], rs[:m]))
    loop_est = (t3 - t2) * wall / wsamp
    return {
        "workload": f"{len(tens)} TT tensors, d in [8,24], n in {{2,4,8,16}}, ranks in [8,128] doubled "
                    f"by self-addition, eps=1e-8, FP64",
        "output_check": check,
        "tensors_per_s": len(tens) / (t2 - t1) if check["ok"] else None, "seconds": t2 - t1, "first_call_seconds_incl_tuning": t1 - t0,
        "per_tensor_loop_tensors_per_s": len(tens) / loop_est,
        "per_tensor_loop_sample": f"first {m} tensors timed, scaled by QR work to the batch",
        "speedup_vs_loop": loop_est / (t2 - t1), "jacobi_sweeps": binfo.get("jacobi_sweeps"),
        "algorithmic_gflop_per_batch": batch_flops / 1e9,
        "roofline": roofline(batch_flops, t2 - t1, what="whole batched call (one persistent launch)")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-rounds", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-rounding leg (512 uneven tensors, config 4 recipe)")
    ap.add_argument("--batched-count", type=int, default=512)
    ap.add_argument("--mtimes-only", action="store_true",
                    help="only run the config-2 leg (A*x and the fused round(A*x)) and exit (profiling)")
    ap.add_argument("--mtimes-unfused-only", action="store_true",
                    help="only run the unfused A*x of config 2 and exit (profiling)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched leg (profiling; prints its JSON object)")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the configs 1, 2 and 5 throughput leg")
    ap.add_argument("--no-svd-always", action="store_true",
                    help="skip the comparison rounds with an SVD at every bond (shorter profiles)")
    ap.add_argument("--skip-plain", action="store_true",
                    help="time only the mixed form (profiling)")
    ap.add_argument("--completion-only", action="store_true",
                    help="only run the completion-gradient workload (65,536 samples, d=16 n=8 rank 32) and exit")
    ap.add_argument("--plain-only", action="store_true",
                    help="time only the plain FP64 form on the metric's tensor and exit (profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy (HBM counter calibration) and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the TT rounding path has no multi-GPU form (it "
              "would shard by core ranges, passing the bond factors between GPUs); run with --gpus 1",
              flush=True)
        sys.exit(2)

    import numpy as np
    import torch
    import ttgpu
    from ttgpu import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_probe:
        nb = 256 << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": nb, "copy_probe_bytes_written": nb, "launches": 3}), flush=True)
        return
    if args.mtimes_unfused_only:
        ctx = ttgpu.Context(0)
        print(json.dumps(mtimes_leg(ttgpu, W, ctx, torch.device("cuda:0"), fused=False)), flush=True)
        return
    if args.mtimes_only:
        ctx = ttgpu.Context(0)
        print(json.dumps(mtimes_leg(ttgpu, W, ctx, torch.device("cuda:0"))), flush=True)
        return
    if args.completion_only:
        print(json.dumps(completion_leg(ttgpu, W, torch.device("cuda:0"))), flush=True)
        return
    if args.batched_only:
        ctx = ttgpu.Context(0)
        print(json.dumps({"batched": batched_leg(ttgpu, W, ctx, torch.device("cuda:0"), args.batched_count)}),
              flush=True)
        return
    log("generating input (d=32, n=8, r=256)")
    c = W.mixed3()
    n, r, core = c["x"]
    eps = c["eps"]
    dev = torch.device("cuda:0")
    dcore = torch.from_numpy(core).to(dev)
    ctx = ttgpu.Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
    cap = ttgpu.capacity(n, r)
    out_mixed = torch.empty(cap, dtype=torch.float64, device=dev)
    out_plain = torch.empty(cap, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def one(prec, out):
        return ttgpu.round_dev(ctx, n, r, dcore, eps, precision=prec, out=out)

    if args.plain_only:
        for i in range(max(args.warmup, 1)):
            one("fp64", out_plain)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        info = {}
        for i in range(args.steps):
            ttgpu.round_dev(ctx, n, r, dcore, eps, precision="fp64", out=out_plain, info=info)
        e1.record(stream)
        torch.cuda.synchronize()
        print(json.dumps({"plain_ms_per_round": e0.elapsed_time(e1) / args.steps, "info": info}), flush=True)
        return
    # warm-up (also builds the workspace)
    for i in range(max(args.warmup, 1)):
        one("fp32", out_mixed)
        log(f"warmup {i + 1} done")
    # find the dominant kernel class of the mixed form (one profiled round per class)
    per_class = {}
    for cls in CLASS_NAMES:
        ctx.profile_begin(cls)
        one("fp32", out_mixed)
        per_class[cls] = ctx.profile_end()
    dominant = max(per_class, key=lambda k: per_class[k]["ms"])
    log("per-class ms per round:", {k: round(v["ms"], 3) for k, v in per_class.items()}, "dominant:", dominant)

    # timed mixed rounds (HIP events on the context's stream; no per-kernel
    # events inside the timed region, so the captured launch sequence replays)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    info = {}
    for i in range(args.steps):
        r_mixed, y_mixed = ttgpu.round_dev(ctx, n, r, dcore, eps, precision="fp32", out=out_mixed, info=info)
    ev1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = ev0.elapsed_time(ev1) / args.steps
    y_mixed_host = y_mixed.cpu().numpy().copy()  # the later rounds reuse the output buffers
    info_mixed = dict(info)
    # the dominant kernel class's launches bracketed by HIP events on the
    # context's stream, over the same number of rounds run eagerly
    ctx.profile_begin(dominant)
    for i in range(args.steps):
        one("fp32", out_mixed)
    prof = ctx.profile_end()
    log(f"mixed: {ms_step:.3f} ms/round (host clock {1e3 * (t1 - t0) / args.steps:.3f}), info {info}")

    # plain FP64 form on the same input
    if args.skip_plain:
        print(json.dumps({"mixed_ms_per_round": ms_step, "info": info, "dominant": dominant,
                          "dominant_launches_per_round": prof["launches"] / args.steps,
                          "dominant_avg_launch_ms": prof["ms"] / max(prof["launches"], 1)}), flush=True)
        return
    one("fp64", out_plain)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(args.plain_steps):
        r_plain, y_plain = one("fp64", out_plain)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_plain = e0.elapsed_time(e1) / args.plain_steps
    y_plain_host = y_plain.cpu().numpy().copy()
    info_plain = {}
    ttgpu.round_dev(ctx, n, r, dcore, eps, precision="fp64", out=out_plain, info=info_plain)
    log(f"plain: {ms_plain:.3f} ms/round, info {info_plain}")

    # the same rounds with the no-truncation shortcut disabled (an SVD at every
    # bond), for comparison
    full_svd = {}
    os.environ["TTG_SVD_ALWAYS"] = "1"
    try:
        for prec, out in (() if args.no_svd_always else (("fp32", out_mixed), ("fp64", out_plain))):
            one(prec, out)
            torch.cuda.synchronize()
            ea = torch.cuda.Event(enable_timing=True)
            eb = torch.cuda.Event(enable_timing=True)
            ea.record(stream)
            for i in range(2):
                one(prec, out)
            eb.record(stream)
            torch.cuda.synchronize()
            full_svd["mixed_ms_per_round" if prec == "fp32" else "plain_ms_per_round"] = ea.elapsed_time(eb) / 2
    finally:
        del os.environ["TTG_SVD_ALWAYS"]
    log("SVD at every bond:", full_svd)

    # algorithmic work per round at the actual ranks (round.m: QR + apply + SVD + apply)
    ro = list(map(int, r_mixed))
    flops = round_flops(n, r, ro)

    result = {
        "metric": "TT rounds/sec d=32 n=8 r=256: mixed vs FP64 on GPU at rel err <=1e-6",
        "value": 1000.0 / ms_step,
        "unit": "rounds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64",
        "data": "synthetic (seeded: sum of two rank-128 TTs, core spectra 2^(-j/16))",
        "config": {"workload": "round d=32 n=8 r=256 (ranks [1,8,64,256x27,64,8,1] after QR caps) eps=1e-6",
                   "d": 32, "n": 8, "r": 256, "eps": eps,
                   "form": "mixed (error budget 1e-6): one-pass CholeskyQR of the TT from FP64 Gram chains, "
                           "batched factorisations; FP32 products exceed the budget at these spectra"},
        "plain_fp64": {"rounds_per_s": 1000.0 / ms_plain, "ms_per_round": ms_plain,
                       "path": "gram-preconditioned exact sweep" if info_plain.get("path") == 1 else "QR/SVD sweeps",
                       "speedup_mixed_vs_plain": ms_plain / ms_step},
        "algorithmic_gflop_per_round": flops / 1e9,
        "achieved_round_tflops": flops / (ms_step * 1e-3) / 1e12,
        "ranks_out": ro,
        "svd_skipped_bonds_per_round": info.get("svd_skipped"),
        "with_svd_at_every_bond": full_svd,
        "diag": {"qr_fallbacks": info.get("qr_fallbacks"), "jacobi_sweeps": info.get("jacobi_sweeps"),
                 "jacobi_maxsweeps": info.get("jacobi_maxsweeps"), "qr_retries": info.get("qr_retries"),
                 "per_class_ms_per_round": {k: v["ms"] for k, v in per_class.items()}},
    }
    # roofline of the dominant kernel class
    avg_ms = prof["ms"] / max(prof["launches"], 1)
    # every kernel class computes in FP64 (the mixed form's FP32 Gram
    # products are a small share of the GEMM launches)
    peak, unit, bound = PEAK["fp64"], "TFLOP/s", "mfma"
    achieved = (prof["flops"] / max(prof["launches"], 1)) / (avg_ms * 1e-3)
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(dominant)
        except Exception:
            traffic = None
    result["roofline"] = {"bound": bound, "kernel": dominant, "achieved": achieved / 1e12, "peak": peak / 1e12,
                          "unit": unit, "frac": achieved / peak, "traffic": traffic,
                          "launches_per_round": prof["launches"] / args.steps, "avg_launch_ms": avg_ms}

    # truncating round, the sum fused into the round, norm / dot / difference
    # norm and TT-SVD (one GPU, inputs in HBM)
    extra = {}
    if not args.no_other:
        ym_d = torch.from_numpy(y_mixed_host).to(dev)
        yp_d = torch.from_numpy(y_plain_host).to(dev)
        extra = extra_legs(ttgpu, W, ctx, dev, c, dcore, ((r_mixed, ym_d), (r_plain, yp_d)))
        del ym_d, yp_d
        log("extra legs:", {k: v for k, v in extra.items() if not k.startswith("_")})

    # batched rounding of uneven tensors (config-4 recipe)
    if not args.no_batched:
        del dcore, out_mixed, out_plain
        torch.cuda.empty_cache()
        result["batched"] = batched_leg(ttgpu, W, ctx, dev, args.batched_count)
        log("batched:", result["batched"])

    # the other configurations at one GPU: rounds/s, contractions/s, gradient evaluations/s
    if not args.no_other:
        result["other_configs"] = other_configs(ttgpu, W, ctx, dev)
        result["other_configs"].update({k: v for k, v in extra.items() if not k.startswith("_")})
        log("other configs:", result["other_configs"])

    # parity against the FP64 oracle (CPU) + CPU baseline timing
    if not args.no_cpu:
        from oracle import tt_oracle as O
        try:
            from threadpoolctl import threadpool_info
            cores = max([t.get("num_threads", 1) for t in threadpool_info()] or [1])
        except Exception:
            cores = os.cpu_count() or 1
        X = O.TT(n, r, core)
        times = []
        ref = None
        for i in range(args.cpu_rounds):
            t0 = time.perf_counter()
            ref = O.tt_round(X, eps)
            times.append(time.perf_counter() - t0)
            log(f"cpu oracle round {i + 1}: {times[-1]:.2f} s")
        med = sorted(times)[len(times) // 2]
        # the same round on one host thread (BLAS/LAPACK limited to one thread)
        one = None
        try:
            from threadpoolctl import threadpool_limits
            with threadpool_limits(limits=1):
                t0 = time.perf_counter()
                O.tt_round(X, eps)
                one = 1.0 / (time.perf_counter() - t0)
            log(f"cpu oracle round, 1 thread: {1.0 / one:.2f} s")
        except Exception as e:  # threadpoolctl missing: report it
            log("one-thread baseline skipped:", e)
        result["cpu_baseline"] = {"value": 1.0 / med, "unit": "rounds/s", "cores": int(cores), "kind": "port",
                                  "sample": f"{args.cpu_rounds} full rounds of the same d=32 n=8 r=256 tensor "
                                            f"(numpy/LAPACK restatement of round.m), median",
                                  "one_thread": {"value": one, "unit": "rounds/s", "cores": 1,
                                                 "sample": "1 full round, BLAS/LAPACK limited to one thread"},
                                  "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}
        ym = O.TT(n, r_mixed, y_mixed_host)
        yp = O.TT(n, r_plain, y_plain_host)
        result["rel_err_mixed_vs_oracle"] = O.tt_rel_diff(ym, ref)
        result["rel_err_plain_vs_oracle"] = O.tt_rel_diff(yp, ref)
        result["plain_ranks_match_oracle"] = list(map(int, r_plain)) == ref.r.tolist()
        result["mixed_ranks_match_oracle"] = list(map(int, r_mixed)) == ref.r.tolist()
        if not args.no_other:  # oracle rates of the small configs, same host cores
            c1 = W.round1()
            Z = O.TT(*c1["z"])
            t0 = time.perf_counter()
            for _ in range(5):
                O.tt_round(Z, c1["eps"])
            t1 = time.perf_counter()
            c2 = W.matvec2()
            na, ma, ra, acore = c2["a"]
            A2 = O.TTM(na, ma, O.TT(na * ma, ra, acore))
            X2 = O.TT(*c2["x"])
            t2 = time.perf_counter()
            Y2 = O.ttm_mtimes(A2, X2)
            t3 = time.perf_counter()
            O.tt_round(Y2, 1e-6)
            t4 = time.perf_counter()
            result["cpu_baseline_other"] = {
                "kind": "port", "cores": int(cores),
                "config1_rounds_per_s": 5 / (t1 - t0),
                "config2_mtimes_per_s": 1 / (t3 - t2),
                "config2_mtimes_then_round_per_s": 1 / (t4 - t2)}
            try:
                from threadpoolctl import threadpool_limits
                with threadpool_limits(limits=1):
                    t0 = time.perf_counter()
                    for _ in range(5):
                        O.tt_round(Z, c1["eps"])
                    result["cpu_baseline_other"]["config1_rounds_per_s_one_thread"] = 5 / (time.perf_counter() - t0)
            except Exception as e:
                log("one-thread config-1 baseline skipped:", e)
            if "_tt_svd_input" in extra:
                nn, full = extra["_tt_svd_input"]
                t0 = time.perf_counter()
                O.tt_from_full(full, 1e-6, nn)
                result["cpu_baseline_other"]["tt_svd_per_s"] = 1 / (time.perf_counter() - t0)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
