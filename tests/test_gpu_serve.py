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

from conftest import f32bits, serve_all

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
        # walk counters and the walk kernel's own duration come with the result
        walked = np.isin(b, [S.FOUND, S.ADDED, S.RECOVERED, S.HEURISTIC]).sum()
        assert results[i]["redirects"] <= results[i]["hops"]
        assert (results[i]["hops"] > 0) == (walked > 0)
        assert 0.0 <= results[i]["walk_ms"] < results[i]["latency_ms"]
        assert results[i]["walk_ms"] > 0.0 or walked == 0
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


def test_serve_failed_copy_applies_nothing(ctx, grid_wl):
    """A request whose input copy fails (forced from the host through the
    testing hook: no copy is issued) is refused by submit, is never applied,
    and stops the server; the request before it keeps its result."""
    from dygrass_amd import DygError, Server, Stream, stream_params

    wl = grid_wl
    sp = stream_params(w_ins=64, w_del=64, threshold=10.0, seed=3, max_batch=512,
                       max_inserts=4096)
    g_edges, h_edges = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)
    a = ("ins", wl.ins.u[:50], wl.ins.v[:50], wl.ins.w[:50])  # read in place: no copy
    idx = np.arange(0, len(wl.G), len(wl.G) // 120)[:120]
    b = ("del", wl.G.u[idx], wl.G.v[idx], None)               # copied to the device

    alone = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    want, _ = alone.apply(*a)

    st = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    srv = Server(st, slots=2, max_request=512)
    ta = srv.submit(*a)
    srv.fail_next_copy()
    with pytest.raises(DygError) as e:
        srv.submit(*b)
    assert "copy failed" in str(e.value)
    got, _ = srv.wait(ta, 50)
    assert np.array_equal(got, want)
    with pytest.raises(DygError) as e:  # the server has stopped
        srv.submit(*a)
    assert "stopped after an error" in str(e.value)
    srv.close()
    _same(alone.G, st.G)
    _same(alone.H, st.H)
    assert st.position == alone.position == 50
    st.close()
    alone.close()


def test_serve_after_rejected_device_batch(ctx, grid_wl):
    """A device batch rejected for an out-of-range id leaves the stream's
    rejection word set; a served deletion request that waits for its
    recovery list sizes (more candidates than the one-workgroup post kernel
    holds) must still recover its paths, as dyg_stream_apply does."""
    from dygrass_amd import DeviceArray, DygError, Server, Stream, stream_params

    wl = grid_wl
    sp = stream_params(w_ins=64, w_del=128, threshold=10.0, seed=4, max_batch=512,
                       max_inserts=4096)  # 100 steps, 16 heuristic walkers per endpoint
    g_edges, h_edges = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)
    hu, hv = wl.H.u[::40][:150], wl.H.v[::40][:150]  # sparsifier edges: walked
    req = ("del", hu, hv, None)

    alone = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    want, wst = alone.apply(*req)
    assert wst["recovered"] + wst["heuristic"] > 0

    st = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    bad_u = DeviceArray.from_numpy(ctx, np.array([0, wl.G.V + 5], np.uint32))
    bad_v = DeviceArray.from_numpy(ctx, np.array([1, 2], np.uint32))
    with pytest.raises(DygError) as e:
        st.apply_dev("del", bad_u, bad_v)
    assert "out of range" in str(e.value)
    srv = Server(st, slots=2, max_request=512)
    got = serve_all(srv, [req])[0][0]
    srv.close()
    assert np.array_equal(got, want)
    _same(alone.G, st.G)
    _same(alone.H, st.H)
    for a in (bad_u, bad_v):
        a.close()
    st.close()
    alone.close()


def test_serve_grid3d_family_matches_oracle(ctx):
    """The served workload's family (3D grid plus random shortcuts, Poisson
    trace of mixed requests) at a reduced size: every request's outcome codes
    through the server equal the stream oracle's, and so do G and H."""
    from dygrass_amd import Server, Stream, stream_params, synth
    from oracle import stream as S

    wl = synth.config_served(0.002, 0, count=80)
    maxr = max(len(r[1]) for r in wl.requests)
    g_edges, h_edges = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)
    kw = dict(w_ins=64, w_del=512, threshold=wl.threshold, seed=0, steps_ins=32, steps_del=32)
    sp = stream_params(**{k: kw[k] for k in ("w_ins", "w_del", "threshold", "seed",
                                             "steps_ins", "steps_del")},
                       max_batch=maxr, max_inserts=wl.updates)
    st = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    srv = Server(st, slots=8, max_request=maxr)
    got, _ = serve_all(srv, wl.requests)
    srv.close()
    orc = S.Stream(wl.G.V, g_edges, h_edges, thr=wl.threshold, w_ins=64, w_del=512, seed=0,
                   steps_ins=32, steps_del=32)
    kinds = set()
    for (op, u, v, w), a in zip(wl.requests, got):
        want = orc.apply(op, u, v, w)
        assert np.array_equal(a, want)
        kinds.update(np.unique(want).tolist())
    assert {S.FOUND, S.ADDED, S.G_ONLY, S.RECOVERED} <= kinds
    for dev, (orp, ocol, ow) in ((st.G, orc.G.csr()), (st.H, orc.H.csr())):
        rp, col, w = dev.download()
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        assert np.array_equal(f32bits(w), f32bits(ow))
    st.close()


def test_serve_device_inputs_equal_host_inputs(ctx, grid_wl):
    """dyg_serve_submit_dev: the same request sequence submitted from device
    arrays (read in place, small and general-path sizes, insertions and
    deletions) gives the outcome codes and graphs of host submits; a request
    with an out-of-range id is refused at submit and the server goes on."""
    from dygrass_amd import DeviceArray, DygError, Server, Stream, stream_params

    wl = grid_wl
    reqs = _requests(wl, 9, count=18)
    rng = np.random.default_rng(9)
    big = rng.choice(len(wl.G), 2500, replace=False)  # above the one-workgroup size
    reqs.insert(5, ("del", wl.G.u[big], wl.G.v[big], None))
    reqs.insert(9, ("ins", wl.ins.u[1500:1980], wl.ins.v[1500:1980], wl.ins.w[1500:1980]))
    sp = stream_params(w_ins=64, w_del=64, threshold=10.0, seed=6, max_batch=3000,
                       max_inserts=16384)
    g_edges, h_edges = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)

    host = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    srv = Server(host, slots=4, max_request=3000)
    want, _ = serve_all(srv, reqs)
    srv.close()

    dev = Stream(ctx, wl.G.V, g_edges, h_edges, sp)
    srv = Server(dev, slots=4, max_request=3000)
    arrays = [(DeviceArray.from_numpy(ctx, np.asarray(u, np.uint32)),
               DeviceArray.from_numpy(ctx, np.asarray(v, np.uint32)),
               DeviceArray.from_numpy(ctx, np.asarray(w, np.float32)) if w is not None else None)
              for _, u, v, w in reqs]
    bu = DeviceArray.from_numpy(ctx, np.array([1, 2, wl.G.V], np.uint32))
    bv = DeviceArray.from_numpy(ctx, np.array([3, 4, 5], np.uint32))
    tickets = []
    for i, ((op, u, v, w), (du, dv, dw)) in enumerate(zip(reqs, arrays)):
        if i == 7:
            with pytest.raises(DygError, match="out of range"):
                srv.submit_dev("del", bu, bv)
        tickets.append(srv.submit_dev(op, du, dv, dw))
    got = [srv.wait(t, len(r[1]))[0] for t, r in zip(tickets, reqs)]
    srv.close()
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    _same(host.G, dev.G)
    _same(host.H, dev.H)
    assert dev.position == host.position
    for t in arrays:
        for a in t:
            if a is not None:
                a.close()
    host.close()
    dev.close()
