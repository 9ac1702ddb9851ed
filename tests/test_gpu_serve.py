"""Served form: every request's result is the same alone and in the mix.

"Alone": the request sequence applied one call at a time with
dyg_stream_apply, each call returning before the next starts.  "In the mix":
the same sequence submitted to the server from one thread while a second
thread collects results, so several requests are admitted, copied and
launched while earlier ones are still on the GPU.  Outcome codes must be
equal request by request and the final G and H equal slot for slot; the
sequence is also checked against the oracle.
"""
import threading

import numpy as np
import pytest

from conftest import f32bits

pytestmark = pytest.mark.gpu


def _requests(wl, seed, count=24):
    rng = np.random.default_rng(seed)
    G = wl.G
    out, k = [], 0
    for r in range(count):
        n = int(np.floor(np.exp(rng.uniform(0, np.log(400)))))
        if rng.uniform() < 0.6:
            u, v, w = wl.ins.u[k:k + n], wl.ins.v[k:k + n], wl.ins.w[k:k + n]
            k += n
            if r % 5 == 0:  # an existing edge and a repeat inside the request
                e = int(rng.integers(len(G)))
                u = np.append(u, [G.u[e], u[0]]); v = np.append(v, [G.v[e], v[0]])
                w = np.append(w, [0.5, 0.25]).astype(np.float32)
            out.append(("ins", u, v, w))
        else:
            idx = rng.choice(len(G), n, replace=False)
            out.append(("del", G.u[idx], G.v[idx], None))
    return out


def _same(a, b):
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(f32bits(x) if x.dtype == np.float32 else x,
                              f32bits(y) if y.dtype == np.float32 else y)


def test_serve_alone_equals_in_mix(ctx, grid_wl):
    from dygrass_amd import Server, Stream, stream_params
    from oracle import stream as S

    wl = grid_wl
    reqs = _requests(wl, 5)
    sp = stream_params(w_ins=64, w_del=64, threshold=10.0, seed=2, max_batch=512,
                       max_inserts=8192)
    g_edges, h_edges = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)

    alone = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    alone_out = [alone.apply(op, u, v, w)[0] for op, u, v, w in reqs]

    mixed = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    srv = Server(mixed, slots=4, max_request=512)
    tickets = []
    got = [None] * len(reqs)
    results = [None] * len(reqs)

    def collect():
        for i in range(len(reqs)):
            while len(tickets) <= i:
                threading.Event().wait(0.0005)
            got[i], results[i] = srv.wait(tickets[i], len(reqs[i][1]))

    th = threading.Thread(target=collect)
    th.start()
    for op, u, v, w in reqs:
        tickets.append(srv.submit(op, u, v, w))
    th.join(timeout=120)
    assert not th.is_alive()
    srv.close()
    assert tickets == sorted(tickets)
    for i, (a, b) in enumerate(zip(alone_out, got)):
        assert np.array_equal(a, b), i
        assert results[i]["n"] == len(reqs[i][1])
        assert results[i]["latency_ms"] >= results[i]["queue_ms"] >= 0.0
    _same(alone.G, mixed.G)
    _same(alone.H, mixed.H)
    assert alone.position == mixed.position

    orc = S.Stream(wl.G.V, g_edges, h_edges, thr=10.0, w_ins=64, w_del=64, seed=2)
    for (op, u, v, w), a in zip(reqs, alone_out):
        assert np.array_equal(orc.apply(op, u, v, w), a)


def test_serve_admission(ctx, grid_wl):
    from dygrass_amd import DygError, Server, Stream, stream_params

    wl = grid_wl
    sp = stream_params(w_ins=16, w_del=16, threshold=10.0, max_batch=64, max_inserts=1024)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    with pytest.raises(DygError):  # larger than a slot
        Server(st, slots=2, max_request=128)
    srv = Server(st, slots=2, max_request=64)
    with pytest.raises(DygError) as e:  # larger than max_request
        srv.submit("ins", np.arange(65), np.arange(65) + 1, np.ones(65))
    assert "max_request" in str(e.value)
    t = srv.submit("ins", wl.ins.u[:10], wl.ins.v[:10], wl.ins.w[:10])
    out, res = srv.wait(t, 10)
    assert res["n"] == 10 and len(out) == 10
    srv.close()
    # a budget the resident state already exceeds
    with pytest.raises(DygError) as e:
        Server(st, slots=1, max_request=64, budget_bytes=1 << 20)
    assert "ADMISSION" in str(e.value)
