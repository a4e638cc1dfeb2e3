"""Launch interface of the Gram-chain round (repeated rounds of one shape).

prepare rounds once and keeps the captured launch sequence; launch re-rounds
whatever the input buffer holds with no host synchronisation and reports
through a device flag; the captured graph also embeds as a child node of a
caller's graph.  Every replay is checked bitwise against an eager round of the
same data (TTG_GRAPH=0, its own context and buffers) and against the oracle
(<= 1e-10, ranks identical), with eager work and synchronisations between
replays.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import ttgpu
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


def _eager(x, eps, prec, monkeypatch):
    ctx2 = ttgpu.Context(0)
    d = torch.tensor(x.core, device="cuda:0")
    torch.cuda.synchronize()
    monkeypatch.setenv("TTG_GRAPH", "0")
    r, y = ttgpu.round_dev(ctx2, x.n, x.r, d, eps, precision=prec)
    monkeypatch.delenv("TTG_GRAPH")
    out = (r, y.cpu().numpy().copy())
    ctx2.close()
    return out


@pytest.mark.parametrize("prec", ["fp32", pytest.param("fp64", marks=pytest.mark.skip(
    reason="the plain-form variant did not finish on the GPU (killed after 180 s without output); the cause "
           "is not known, so the interface refuses TTG_PREC_FP64 and this case stays off the GPU"))])
def test_prepare_then_launch_on_new_data(monkeypatch, prec):
    ctx = ttgpu.Context(0)
    xs = [O.make_config(3, d=9, n=4, rank=20, seed=s)["x"] for s in (91, 92, 93, 94)]
    n, r = xs[0].n, xs[0].r
    dcore = torch.tensor(xs[0].core, device="cuda:0")
    torch.cuda.synchronize()
    r1, y, proven = ttgpu.round_gram_prepare(ctx, n, r, dcore, 1e-6, prec)
    assert proven
    er, ey = _eager(xs[0], 1e-6, prec, monkeypatch)
    assert r1.tolist() == er.tolist() and y.cpu().numpy().tobytes() == ey.tobytes()
    dstat = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    other = O.rand_tt(np.random.default_rng(95), [3] * 5, [1, 3, 7, 7, 3, 1])
    for x in xs[1:] * 2:
        dcore.copy_(torch.from_numpy(x.core))
        dstat.zero_()
        torch.cuda.synchronize()
        ttgpu.round_gram_launch(ctx, dstat)
        torch.cuda.synchronize()
        assert int(dstat.item()) == 1
        g = y.cpu().numpy().copy()
        er, ey = _eager(x, 1e-6, prec, monkeypatch)
        assert er.tolist() == r1.tolist() and g.tobytes() == ey.tobytes()
        ref = O.tt_round(x, 1e-6)
        assert ref.r.tolist() == r1.tolist()
        assert O.tt_rel_diff(O.TT(n, r1, g), ref) <= 1e-10
        ttgpu.round(ttgpu.TTTensor(other.n, other.r, other.core), 1e-8, ctx=ttgpu.default_context())
    # a Gram-form round of another shape on this context replaces the prepared one
    ttgpu.round_dev(ctx, other.n, other.r, torch.tensor(other.core, device="cuda:0"), 1e-8, precision="fp32")
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.round_gram_launch(ctx, dstat)
    assert e.value.status == 1
    ctx.close()


def test_plain_form_is_refused(gpu_ctx):
    x = O.make_config(3, d=6, n=4, rank=8, seed=90)["x"]
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.round_gram_prepare(gpu_ctx, x.n, x.r, torch.tensor(x.core, device="cuda:0"), 1e-6, "fp64")
    assert e.value.status == 1


def test_not_proven_is_reported():
    """x + x truncates at every bond: prepare reports it, launch writes 0."""
    ctx = ttgpu.Context(0)
    rng = np.random.default_rng(96)
    x = O.rand_tt(rng, [3] * 6, [1, 3, 9, 12, 9, 3, 1])
    z = O.tt_plus(x, x)
    dcore = torch.tensor(z.core, device="cuda:0")
    torch.cuda.synchronize()
    _, _, proven = ttgpu.round_gram_prepare(ctx, z.n, z.r, dcore, 1e-10, "fp32")
    assert not proven
    dstat = torch.ones(1, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    ttgpu.round_gram_launch(ctx, dstat)
    torch.cuda.synchronize()
    assert int(dstat.item()) == 0
    # a shape outside the form (ranks above 480 after the caps) is refused up front
    wide = O.rand_tt(rng, [30] * 4, [1, 30, 600, 30, 1])
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.round_gram_prepare(ctx, wide.n, wide.r, torch.tensor(wide.core, device="cuda:0"), 1e-8, "fp32")
    assert e.value.status == 1
    ctx.close()


def test_child_graph_in_a_callers_graph(monkeypatch):
    hip = C.CDLL("libamdhip64.so")
    V = C.c_void_p
    hip.hipGraphCreate.argtypes = [C.POINTER(V), C.c_uint]
    hip.hipGraphAddChildGraphNode.argtypes = [C.POINTER(V), V, C.POINTER(V), C.c_size_t, V]
    hip.hipGraphInstantiate.argtypes = [C.POINTER(V), V, C.POINTER(V), C.c_char_p, C.c_size_t]
    hip.hipGraphLaunch.argtypes = [V, V]
    hip.hipStreamSynchronize.argtypes = [V]
    hip.hipGraphExecDestroy.argtypes = [V]
    hip.hipGraphDestroy.argtypes = [V]
    ctx = ttgpu.Context(0)
    xs = [O.make_config(3, d=8, n=4, rank=16, seed=s)["x"] for s in (97, 98, 99)]
    n, r = xs[0].n, xs[0].r
    dcore = torch.tensor(xs[0].core, device="cuda:0")
    torch.cuda.synchronize()
    r1, y, proven = ttgpu.round_gram_prepare(ctx, n, r, dcore, 1e-6, "fp32")
    assert proven
    child = ttgpu.round_gram_graph(ctx)
    parent, node, exe = V(), V(), V()
    assert hip.hipGraphCreate(C.byref(parent), 0) == 0
    assert hip.hipGraphAddChildGraphNode(C.byref(node), parent, None, 0, child) == 0
    assert hip.hipGraphInstantiate(C.byref(exe), parent, None, None, 0) == 0
    s = V(ctx.stream)
    try:
        for x in xs[1:] * 2:
            dcore.copy_(torch.from_numpy(x.core))
            torch.cuda.synchronize()
            assert hip.hipGraphLaunch(exe, s) == 0
            assert hip.hipStreamSynchronize(s) == 0
            g = y.cpu().numpy().copy()
            er, ey = _eager(x, 1e-6, "fp32", monkeypatch)
            assert er.tolist() == r1.tolist() and g.tobytes() == ey.tobytes()
            assert O.tt_rel_diff(O.TT(n, r1, g), O.tt_round(x, 1e-6)) <= 1e-10
    finally:
        hip.hipGraphExecDestroy(exe)
        hip.hipGraphDestroy(parent)
    ctx.close()
