"""GPU: the product entry points launch kernels only (no allocation, copy or
memset), so a caller can capture them in a hipGraph -- here torch.cuda.CUDAGraph
on a stream handed to the engine.  Each replay, with eager work and
synchronisations between the replays, gives the eager result bit for bit.
Contexts on separate streams run concurrently with the same bits."""
import numpy as np
import pytest

from int_spmv_amd import device as D
from int_spmv_amd import host as H

from .util import bitwise_equal

pytestmark = pytest.mark.gpu


def test_products_replay_from_a_captured_graph(engine):
    import torch
    m = H.CSR.powerlaw(n=1 << 13, target_nnz=200_000, maxlen=4096, seed=4)
    n = m.nrows
    x = H.uniform(n, seed=3)
    xd = engine.upload(x)
    dc = D.DeviceCSR(engine, m)
    dax = D.DeviceAXC(engine, H.AXC(m, x))
    dk = D.DeviceK1(engine, H.K1(m))
    d1 = D.DeviceAXT(engine, H.AXT(m, x, th=1))
    d4 = D.DeviceAXT(engine, H.AXT(m, x, th=4))
    dcom = D.DeviceAXT(engine, H.AXT(m, x, th=1, mode="COM", bs=256))
    g = D.DeviceAXTGather(engine, H.AXT(m, np.zeros(n), th=1, want_acol=True))
    adj = D.AxtAdjoint(engine, m, g)
    X3 = engine.upload(np.stack([x, -x, H.uniform(n, seed=5)], axis=1).reshape(-1))
    yb = engine.upload(H.normal(n, seed=6))
    ops = [("csr", n, lambda y: dc.spmv(xd, y)),
           ("csr_ordered", n, lambda y: dc.cf_csr(xd, y)),
           ("axc", n, lambda y: dax.int_axc(y)),
           ("k1p", n, lambda y: dk.int_k1p(xd, y)),
           ("axt_h1", n, lambda y: d1.int_axt(y)),
           ("axt_h4", n, lambda y: d4.int_axt(y)),
           ("axt_com", n, lambda y: dcom.int_axt(y)),
           ("column_map_k3", 3 * n, lambda y: g.matvec(X3, y, k=3)),
           ("adjoint_slot_map", n, lambda y: adj.adjoint(yb, y, vals=g.aval)),
           ("grad_vals", g.a.lenAX // 2, lambda v: g.grad_vals(xd, yb, v))]

    def buffers():
        return [torch.full((size,), float("nan"), dtype=torch.float64, device=engine.torch_device)
                for _, size, _ in ops]

    eager, outs = buffers(), buffers()
    for (_, _, f), y in zip(ops, eager):
        f(y)
    torch.cuda.synchronize()
    want = [e.cpu().numpy() for e in eager]
    assert not any(np.isnan(w).any() for w in want)  # every output written
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    old = engine.stream
    engine.set_stream(s)  # the launches go to the capturing stream
    try:
        with torch.cuda.graph(graph, stream=s):
            for (_, _, f), y in zip(ops, outs):
                f(y)
    finally:
        engine.set_stream(old)
    for rep in range(3):
        for y in outs:
            y.fill_(float("nan"))  # eager work between the replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for (name, _, _), y, w in zip(ops, outs, want):
            assert bitwise_equal(y.cpu().numpy(), w), (rep, name)


def test_two_contexts_on_two_streams(engine):
    """Two contexts, each on its own stream, with their own plans: launches
    interleaved on both streams without synchronisation give the bits of a
    single context (the library keeps no state outside its contexts, plans
    and batches)."""
    import torch
    m = H.CSR.powerlaw(n=1 << 13, target_nnz=200_000, maxlen=4096, seed=8)
    n = m.nrows
    x = H.uniform(n, seed=9)
    ah, a4 = H.AXT(m, x, th=1), H.AXT(m, x, th=4)
    ag = H.AXT(m, np.zeros(n), th=1, want_acol=True)
    xd = engine.upload(x)
    one = (D.DeviceAXT(engine, ah), D.DeviceAXT(engine, a4), D.DeviceAXTGather(engine, ag))
    ref = [torch.full((n,), float("nan"), dtype=torch.float64, device=engine.torch_device)
           for _ in range(3)]
    one[0].int_axt(ref[0])
    one[1].int_axt(ref[1])
    one[2].matvec(xd, ref[2])
    torch.cuda.synchronize()
    ref = [r.cpu().numpy() for r in ref]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    engs = [D.Engine(0), D.Engine(0)]
    for e, s in zip(engs, streams):
        s.wait_stream(torch.cuda.current_stream())
        e.set_stream(s)
    ops = [(D.DeviceAXT(e, ah), D.DeviceAXT(e, a4), D.DeviceAXTGather(e, ag)) for e in engs]
    outs = [[torch.full((n,), float("nan"), dtype=torch.float64, device=engine.torch_device)
             for _ in range(3)] for _ in engs]
    torch.cuda.synchronize()
    for rep in range(4):  # interleaved, no synchronisation between the streams
        for i in range(2):
            d1, d4, g = ops[i]
            d1.int_axt(outs[i][0])
            d4.int_axt(outs[i][1])
            g.matvec(xd, outs[i][2])
    torch.cuda.synchronize()
    for i in range(2):
        for j in range(3):
            assert bitwise_equal(outs[i][j].cpu().numpy(), ref[j]), (i, j)
