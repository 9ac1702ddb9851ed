"""Full-size checks on the GPU: the 1M-node random geometric graph with 200k
insertions in 10 snapshot batches (threshold 100, 256 walkers), the 2M-node
mesh with 100k deletions, the 4M-point 12-NN batched stream (against the C++
stream oracle and the plain form) and the 8M-vertex served workload.

Every one of the 200k decisions must equal the oracle's, and the final
sparsifier must equal the oracle's slot for slot.  Size-independent
properties are checked as well: H grows by exactly two slots per added edge,
a second run reproduces the first bit for bit.
"""
import os

import numpy as np
import pytest

from conftest import csr, f32bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rgg():
    from dygrass_amd import synth

    return synth.config_rgg_insert(1.0, 0)


def test_full_size_incremental_matches_oracle(ctx, rgg):
    from dygrass_amd import DeviceGraph, incremental_update, params
    from oracle import pyoracle as po

    rp, col, w = csr(rgg.H)
    p = params(rgg.walkers, 100, rgg.threshold, seed=0)
    H = DeviceGraph(ctx, rgg.H.V, rp, col, w, slack=2)
    _, n0, _ = H.info()
    found, st = incremental_update(H, rgg.ins.u, rgg.ins.v, rgg.ins.w, p)
    _, n1, _ = H.info()
    added = int((found == -1).sum())
    assert st["added"] == added and n1 - n0 == 2 * added
    assert np.all((found == -1) | ((found >= 1) & (found <= 99)))

    og = po.Graph(rgg.H.V, rp, col, w)
    threads = min(os.cpu_count() or 1, 16)
    ofound, ost = og.incremental(rgg.ins.u, rgg.ins.v, rgg.ins.w, rgg.threshold, rgg.walkers,
                                 seed=0, threads=threads)
    assert np.array_equal(found, ofound)
    grp, gcol, gw = H.download()
    orp, ocol, ow = og.export()
    assert np.array_equal(grp, orp) and np.array_equal(gcol, ocol)
    assert np.array_equal(f32bits(gw), f32bits(ow))
    assert H.check() == 0

    H2 = DeviceGraph(ctx, rgg.H.V, rp, col, w, slack=0)
    found2, _ = incremental_update(H2, rgg.ins.u, rgg.ins.v, rgg.ins.w, p)
    assert np.array_equal(found, found2)


def test_full_size_decremental_matches_oracle(ctx):
    """2M-node triangulated mesh, 100k uniform deletions in 10 snapshot batches,
    512 walkers: the recovered edge set, the sparsifier size and every counter
    equal the oracle's."""
    from dygrass_amd import DualGraph, params, synth
    from oracle import pyoracle as po

    wl = synth.config_mesh_delete(1.0, 0)
    rp, col, w = csr(wl.G)
    d = DualGraph(ctx, wl.G.V, rp, col, w, wl.H.u, wl.H.v)
    before = d.sparse_size
    log = d.run(wl.dels.u, wl.dels.v, params(512, 100, seed=0))
    added = d.added_edges()
    threads = min(os.cpu_count() or 1, 16)
    st, oadded = po.decremental_lines(wl.G.V, wl.G.u, wl.G.v, wl.G.w, wl.H.u, wl.H.v,
                                      wl.dels.u, wl.dels.v, seed=0, threads=threads)
    assert before == st.sparse_before
    assert d.sparse_size == st.sparse_after
    assert np.array_equal(added, oadded)
    assert sum(b["queued"] for b in log) == st.filtered
    assert sum(b["no_path"] for b in log) == st.no_path
    assert len(log) == 10 and st.filtered > 20000


def _keys(V, u, v):
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    return np.minimum(u, v) * V + np.maximum(u, v)


def _edges(g, V):
    """Each undirected edge of a device graph once: sorted keys, weight bits."""
    rp, col, w = g.download()
    rows = np.repeat(np.arange(V, dtype=np.int64), np.diff(rp).astype(np.int64))
    once = col.astype(np.int64) > rows
    k = rows[once] * V + col[once].astype(np.int64)
    order = np.argsort(k)
    return k[order], w[once].view(np.uint32)[order]


def test_full_size_batched_stream_properties(ctx):
    """The batched-stream workload at full size (4M-point 12-NN graph, 64
    batches, 379k updates) with the committed autotuned setting: outcome codes
    and H equal the plain form's slot for slot (the reference-shaped
    implementation: host edge maps and plans, one workgroup per edge); G ends
    as exactly G0 - deletions + insertions; H lies inside G with G's weight
    bits; every no-path insertion is in H; deleted edges are in neither."""
    from dygrass_amd import Stream, autotune, csr_from_edges, stream_params, synth
    from oracle import stream as S

    wl = synth.config_batched(1.0, 0)
    V = wl.G.V
    t = autotune.load(wl.name)
    s = autotune.setting_of(t) if t else autotune.Setting(w_ins=256, w_del=512)
    csr = (csr_from_edges(V, wl.G.u, wl.G.v, wl.G.w), csr_from_edges(V, wl.H.u, wl.H.v, wl.H.w))
    maxb = max(len(b[1]) for b in wl.batches)

    def run(plain):
        sp = stream_params(w_ins=s.w_ins, w_del=s.w_del, threshold=wl.threshold, seed=0,
                           max_batch=maxb, max_inserts=wl.updates, plain=plain,
                           steps_ins=s.steps_ins, steps_del=s.steps_del,
                           block_edges_ins=0 if plain else s.block_ins,
                           block_edges_del=0 if plain else s.block_del,
                           dense_no_desc=s.dense_no_desc and not plain)
        st = Stream(ctx, V, None, None, sp, csr=csr)
        outs = [st.apply(op, u, v, w)[0] for op, u, v, w in wl.batches]
        return st, outs

    dev, outs = run(False)
    plain, pouts = run(True)
    for a, b in zip(outs, pouts):
        assert np.array_equal(a, b)
    for x, y in zip(dev.H.download(), plain.H.download()):
        assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y)
    plain.close()
    assert dev.H.check() == 0

    ins = [(b, o) for b, o in zip(wl.batches, outs) if b[0] == "ins"]
    dels = [(b, o) for b, o in zip(wl.batches, outs) if b[0] == "del"]
    ik = np.concatenate([_keys(V, b[1], b[2]) for b, _ in ins])
    iw = np.concatenate([np.asarray(b[3], np.float32) for b, _ in ins]).view(np.uint32)
    io = np.concatenate([o for _, o in ins])
    dk = np.concatenate([_keys(V, b[1], b[2]) for b, _ in dels])
    do = np.concatenate([o for _, o in dels])
    assert np.all(np.isin(io, [S.FOUND, S.ADDED])) and np.all(np.isin(do, [
        S.G_ONLY, S.RECOVERED, S.HEURISTIC]))  # new edges only, deletions of G edges only
    g0k = _keys(V, wl.G.u, wl.G.v)
    g0w = np.asarray(wl.G.w, np.float32).view(np.uint32)
    keep = ~np.isin(g0k, dk)
    want_k = np.concatenate([g0k[keep], ik])
    want_w = np.concatenate([g0w[keep], iw])
    order = np.argsort(want_k)
    gk, gw = _edges(dev.G, V)
    assert np.array_equal(gk, want_k[order]) and np.array_equal(gw, want_w[order])
    hk, hw = _edges(dev.H, V)
    pos = np.searchsorted(gk, hk)
    assert np.all(pos < len(gk)) and np.array_equal(gk[pos], hk)  # H inside G
    assert np.array_equal(gw[pos], hw)                           # with G's weights
    assert np.all(np.isin(ik[io == S.ADDED], hk))
    assert not np.any(np.isin(dk, hk))
    dev.close()


@pytest.mark.timeout(600)
def test_full_size_batched_stream_matches_oracle(ctx):
    """The batched-stream workload at full size as the benchmark runs it
    (every batch in device memory, the list applied in one call with the
    committed autotuned setting) against the C++ stream oracle on the host's
    cores: every batch's outcome codes equal, and the final G and H equal
    slot for slot (neighbour order within every row, weight bits)."""
    from dygrass_amd import DeviceArray, Stream, autotune, csr_from_edges, stream_params, synth
    from oracle import pyoracle as po

    wl = synth.config_batched(1.0, 0)
    V = wl.G.V
    t = autotune.load(wl.name)
    s = autotune.setting_of(t) if t else autotune.Setting(w_ins=256, w_del=512)
    csr = (csr_from_edges(V, wl.G.u, wl.G.v, wl.G.w), csr_from_edges(V, wl.H.u, wl.H.v, wl.H.w))
    maxb = max(len(b[1]) for b in wl.batches)
    sp = stream_params(w_ins=s.w_ins, w_del=s.w_del, threshold=wl.threshold, seed=0,
                       max_batch=maxb, max_inserts=wl.updates, steps_ins=s.steps_ins,
                       steps_del=s.steps_del, block_edges_ins=s.block_ins,
                       block_edges_del=s.block_del, dense_no_desc=s.dense_no_desc)
    ops = [b[0] for b in wl.batches]
    off = np.concatenate([[0], np.cumsum([len(b[1]) for b in wl.batches])]).astype(np.uint64)
    du = DeviceArray.from_numpy(ctx, np.concatenate([np.asarray(b[1], np.uint32)
                                                     for b in wl.batches]))
    dv = DeviceArray.from_numpy(ctx, np.concatenate([np.asarray(b[2], np.uint32)
                                                     for b in wl.batches]))
    dw = DeviceArray.from_numpy(ctx, np.concatenate(
        [np.asarray(b[3], np.float32) if b[3] is not None else np.zeros(len(b[1]), np.float32)
         for b in wl.batches]))
    st = Stream(ctx, V, None, None, sp, csr=csr)
    got, _ = st.apply_many_dev(ops, off, du, dv, dw)

    threads = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    orc = po.StreamCpp(V, thr=wl.threshold, w_ins=s.w_ins, w_del=s.w_del, seed=0,
                       steps_ins=s.steps_ins, steps_del=s.steps_del, threads=threads, csr=csr)
    for b, (op, u, v, w) in enumerate(wl.batches):
        want = orc.apply(op, u, v, w)
        assert np.array_equal(got[int(off[b]):int(off[b + 1])], want), f"batch {b} ({op})"
    for which, g in ((0, st.G), (1, st.H)):
        rp, col, w = g.download()
        orp, ocol, ow = orc.csr(which)
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        assert np.array_equal(f32bits(w), f32bits(ow))
    assert st.H.check() == 0
    orc.close()
    st.close()


@pytest.mark.timeout(900)
def test_full_size_served_properties(ctx):
    """The served workload at full size (8M-vertex 3D grid plus 16M random
    shortcuts, ~10M-edge sparsifier, 1500 requests of log-uniform size
    1..10000, 60/40 insert/delete) with the committed autotuned setting, every
    request submitted at once: each request's outcome codes equal those of
    dyg_stream_apply applied one request at a time in admission order, G and
    H equal slot for slot; G ends as G0 - deletions + insertions (repeats as
    fp32 weight sums in order); H lies inside G with G's weight bits; every
    no-path insertion is in H; deleted edges are in neither."""
    from dygrass_amd import Server, Stream, autotune, csr_from_edges, stream_params, synth
    from oracle import stream as S

    from conftest import serve_all

    wl = synth.config_served(1.0, 0)
    V = wl.G.V
    t = autotune.load(wl.name)
    s = autotune.setting_of(t) if t else autotune.Setting(w_ins=64, w_del=512)
    maxr = max(len(r[1]) for r in wl.requests)
    csr = (csr_from_edges(V, wl.G.u, wl.G.v, wl.G.w), csr_from_edges(V, wl.H.u, wl.H.v, wl.H.w))
    sp = stream_params(w_ins=s.w_ins, w_del=s.w_del, threshold=wl.threshold, seed=0,
                       max_batch=maxr, max_inserts=wl.updates, steps_ins=s.steps_ins,
                       steps_del=s.steps_del, block_edges_ins=s.block_ins,
                       block_edges_del=s.block_del, dense_no_desc=s.dense_no_desc)

    served = Stream(ctx, V, None, None, sp, csr=csr)
    srv = Server(served, slots=8, max_request=maxr)
    got, _ = serve_all(srv, wl.requests, timeout=600)
    srv.close()
    alone = Stream(ctx, V, None, None, sp, csr=csr)
    for (op, u, v, w), g in zip(wl.requests, got):
        want, _ = alone.apply(op, u, v, w)
        assert np.array_equal(g, want)
    for a, b in ((served.G, alone.G), (served.H, alone.H)):
        for x, y in zip(a.download(), b.download()):
            assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                                  y.view(np.uint32) if y.dtype == np.float32 else y)
    alone.close()
    assert served.H.check() == 0

    ins = [(r, o) for r, o in zip(wl.requests, got) if r[0] == "ins"]
    dels = [(r, o) for r, o in zip(wl.requests, got) if r[0] == "del"]
    ik = np.concatenate([_keys(V, r[1], r[2]) for r, _ in ins])
    iw = np.concatenate([np.asarray(r[3], np.float32) for r, _ in ins])
    io = np.concatenate([o for _, o in ins])
    dk = np.concatenate([_keys(V, r[1], r[2]) for r, _ in dels])
    do = np.concatenate([o for _, o in dels])
    g0k = _keys(V, wl.G.u, wl.G.v)
    assert len(np.unique(dk)) == len(dk) and np.all(np.isin(dk, g0k))  # distinct G0 edges
    assert not np.any(np.isin(ik, dk))
    assert np.all(np.isin(do, [S.G_ONLY, S.RECOVERED, S.HEURISTIC]))
    # an insertion updates a weight exactly when its edge is in G0 or came earlier
    _, first = np.unique(ik, return_index=True)
    repeat = np.ones(len(ik), bool)
    repeat[first] = False
    in_g0 = np.isin(ik, g0k)
    assert np.array_equal(io == S.WEIGHT, repeat | in_g0)
    assert np.all(np.isin(io[~(repeat | in_g0)], [S.FOUND, S.ADDED]))
    # expected G: G0 without the deletions, the insertions' weights summed in order
    g0w = np.asarray(wl.G.w, np.float32)
    special = set(ik[repeat | in_g0].tolist())  # keys whose weight is a sum
    g0_of = dict(zip(g0k[np.isin(g0k, list(special))].tolist(),
                     g0w[np.isin(g0k, list(special))].tolist()))
    weights = {}
    for k, w in zip(ik.tolist(), iw.tolist()):
        if k not in special:
            continue
        if k not in weights:
            weights[k] = np.float32(g0_of[k]) + np.float32(w) if k in g0_of else np.float32(w)
        else:
            weights[k] = np.float32(weights[k] + np.float32(w))
    firsts = ~(repeat | in_g0)
    keep = ~np.isin(g0k, dk)
    want_k = np.concatenate([g0k[keep], ik[firsts]])
    want_w = np.concatenate([g0w[keep], iw[firsts]])
    order = np.argsort(want_k)
    want_k, want_w = want_k[order], want_w[order]
    for k, w in weights.items():
        want_w[np.searchsorted(want_k, k)] = w
    gk, gw = _edges(served.G, V)
    assert np.array_equal(gk, want_k) and np.array_equal(gw, want_w.view(np.uint32))
    hk, hw = _edges(served.H, V)
    pos = np.searchsorted(gk, hk)
    assert np.all(pos < len(gk)) and np.array_equal(gk[pos], hk)  # H inside G
    assert np.array_equal(gw[pos], hw)                           # with G's weights
    assert np.all(np.isin(ik[io == S.ADDED], hk))
    assert not np.any(np.isin(dk, hk)) and not np.any(np.isin(dk, gk))
    served.close()
