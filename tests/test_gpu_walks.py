"""GPU parity: the HIP walks against the oracle on the same inputs.

Bit-exact is required for everything integer (found_at, winner, hop counts,
paths, graph slot order) and for the float32 conductance, which both sides
compute with IEEE division (compared as bit patterns).
"""
import os
import tempfile

import numpy as np
import pytest

from conftest import csr, f32bits

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import pyoracle

    return pyoracle


def _first_batch(wl, n=None):
    I = (len(wl.ins) + 9) // 10
    n = I if n is None else n
    return wl.ins.u[:n], wl.ins.v[:n], wl.ins.w[:n]


@pytest.mark.parametrize("per_edge", [False, True])
@pytest.mark.parametrize("prune", [True, False])
def test_incremental_walks_match_oracle_grid(ctx, grid_wl, prune, per_edge):
    from dygrass_amd import DeviceGraph, params, walk_incremental

    po = _oracle()
    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w)
    s, t, we = _first_batch(grid_wl)
    p = params(512, 100, 10.0, seed=0, prune=prune, per_edge=per_edge)
    got = walk_incremental(H, s, t, we, p)
    og = po.Graph(grid_wl.H.V, rp, col, w)
    f, c, win, hops = og.walk_inc_batch(s, t, we, 10.0, 512, seed=0, threads=8)
    assert np.array_equal(got.found_at, f)
    assert np.array_equal(f32bits(got.cond), f32bits(c))
    assert np.array_equal(got.winner, win)
    if not prune:
        assert np.array_equal(got.hops.astype(np.uint64), hops)
        assert got.stats["hops"] == int(hops.sum())
    else:
        assert got.stats["hops"] <= int(hops.sum())


@pytest.mark.parametrize("W", [8, 64, 512])
def test_decision_only_walks(ctx, grid_wl, W):
    """DYG_WALK_DECISION_ONLY: found / not found equals the oracle's for every
    edge; a found edge reports a walker that really found the target within
    the bound, never with a larger conductance than the reference's winner
    (the winner maximises it), and the walks do no more hops than pruning
    alone."""
    from dygrass_amd import DeviceGraph, params, walk_incremental

    po = _oracle()
    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w)
    s, t, we = _first_batch(grid_wl, 600)
    got = walk_incremental(H, s, t, we, params(W, 100, 10.0, seed=3, decision_only=True))
    full = walk_incremental(H, s, t, we, params(W, 100, 10.0, seed=3))
    og = po.Graph(grid_wl.H.V, rp, col, w)
    f, c, _, _ = og.walk_inc_batch(s, t, we, 10.0, W, seed=3, threads=8)
    assert np.array_equal(got.found_at == -1, f == -1)
    hit = f != -1
    assert np.all((got.found_at[hit] >= 1) & (got.found_at[hit] <= 99))
    assert np.all(got.cond[hit] > 0) and np.all(got.cond[hit] <= c[hit])
    assert np.all(got.cond[~hit] == -1.0)
    assert got.stats["hops"] <= full.stats["hops"]


@pytest.mark.parametrize("per_edge", [False, True])
@pytest.mark.parametrize("W", [1, 2, 8, 32, 64, 128, 256, 1024])
def test_incremental_walker_counts(ctx, grid_wl, W, per_edge):
    from dygrass_amd import DeviceGraph, params, walk_incremental

    po = _oracle()
    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w)
    s, t, we = _first_batch(grid_wl, 64)
    ids = np.arange(1000, 1064, dtype=np.uint64)
    got = walk_incremental(H, s, t, we, params(W, 100, 10.0, seed=7, prune=False,
                                               per_edge=per_edge), edge_id=ids)
    og = po.Graph(grid_wl.H.V, rp, col, w)
    f, c, win, hops = og.walk_inc_batch(s, t, we, 10.0, W, seed=7, edge_id=ids, threads=8)
    assert np.array_equal(got.found_at, f)
    assert np.array_equal(f32bits(got.cond), f32bits(c))
    assert np.array_equal(got.winner, win)
    assert np.array_equal(got.hops.astype(np.uint64), hops)


@pytest.mark.parametrize("per_edge,host_edits", [(False, False), (True, True), (False, True)])
def test_incremental_update_matches_oracle_run(ctx, grid_wl, per_edge, host_edits):
    """Whole snapshot-batched run: every decision and the final slot order of H."""
    from dygrass_amd import DeviceGraph, incremental_update, params

    po = _oracle()
    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w, slack=1)
    found, stats = incremental_update(H, grid_wl.ins.u, grid_wl.ins.v, grid_wl.ins.w,
                                      params(512, 100, 10.0, seed=0, per_edge=per_edge,
                                             host_edits=host_edits))
    og = po.Graph(grid_wl.H.V, rp, col, w)
    starts = po.batches(len(grid_wl.ins), 10)
    expect = np.zeros(len(grid_wl.ins), np.int32)
    for b0, b1 in zip(starts[:-1], starts[1:]):
        s, t, we = grid_wl.ins.u[b0:b1], grid_wl.ins.v[b0:b1], grid_wl.ins.w[b0:b1]
        f, _, _, _ = og.walk_inc_batch(s, t, we, 10.0, 512, seed=0,
                                       edge_id=np.arange(b0, b1), threads=8)
        expect[b0:b1] = f
        for i in np.nonzero(f == -1)[0]:
            og.insert(int(s[i]), int(t[i]), float(we[i]))
    assert np.array_equal(found, expect)
    assert stats["added"] == int((expect == -1).sum())
    assert stats["batches"] == 10
    grp, gcol, gw = H.download()
    orp, ocol, ow = og.export()
    assert np.array_equal(grp, orp)
    assert np.array_equal(gcol, ocol)
    assert np.array_equal(f32bits(gw), f32bits(ow))
    assert H.check() == 0


@pytest.mark.parametrize("divisor", [1, 7, 64])
def test_incremental_run_order_any_batching(ctx, grid_wl, divisor):
    """The device-edited run sorts every batch's row entries at once, keyed
    by batch and row: one batch (no batch bits), uneven batches (a short last
    one) and many small ones all give the oracle's decisions and slot order."""
    from dygrass_amd import DeviceGraph, incremental_update, params

    po = _oracle()
    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w, slack=1)
    u, v, x = grid_wl.ins.u, grid_wl.ins.v, grid_wl.ins.w
    print(f"incremental run order: divisor {divisor}", flush=True)
    found, stats = incremental_update(H, u, v, x, params(128, 100, 10.0, seed=3),
                                      divisor=divisor)
    og = po.Graph(grid_wl.H.V, rp, col, w)
    starts = po.batches(len(u), divisor)
    expect = np.zeros(len(u), np.int32)
    for b0, b1 in zip(starts[:-1], starts[1:]):
        f, _, _, _ = og.walk_inc_batch(u[b0:b1], v[b0:b1], x[b0:b1], 10.0, 128, seed=3,
                                       edge_id=np.arange(b0, b1), threads=8)
        expect[b0:b1] = f
        for i in np.nonzero(f == -1)[0]:
            og.insert(int(u[b0 + i]), int(v[b0 + i]), float(x[b0 + i]))
    assert np.array_equal(found, expect)
    assert stats["batches"] == len(starts) - 1
    g, o = H.download(), og.export()
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    assert np.array_equal(f32bits(g[2]), f32bits(o[2]))
    assert H.check() == 0


def test_device_edits_chain_with_host_edits(ctx, grid_wl):
    """Two device-edited runs back to back, then host-planned inserts on the
    same graph (the host re-reads the device state first)."""
    from dygrass_amd import DeviceGraph, incremental_update, params

    po = _oracle()
    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w, slack=0)
    og = po.Graph(grid_wl.H.V, rp, col, w)
    u, v, x = grid_wl.ins.u, grid_wl.ins.v, grid_wl.ins.w
    halves = [(0, 1000), (1000, 2000)]
    for b0, b1 in halves:
        found, _ = incremental_update(H, u[b0:b1], v[b0:b1], x[b0:b1],
                                      params(128, 100, 10.0, seed=b0))
        ofound, _ = og.incremental(u[b0:b1], v[b0:b1], x[b0:b1], 10.0, 128, seed=b0, threads=8)
        assert np.array_equal(found, ofound)
    assert H.check() == 0
    H.insert(u[:50], v[:50], x[:50])
    for i in range(50):
        og.insert(int(u[i]), int(v[i]), float(x[i]))
    assert H.check() == 0
    g, o = H.download(), og.export()
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    assert np.array_equal(f32bits(g[2]), f32bits(o[2]))


def test_insert_relocation_keeps_slot_order(ctx):
    """Many appends to one row force relocations; order must be file order."""
    from dygrass_amd import DeviceGraph

    po = _oracle()
    V = 40
    u = np.arange(V - 1, dtype=np.uint32)
    v = u + 1
    w = np.linspace(0.5, 2.0, V - 1).astype(np.float32)
    rp, col, ww = __import__("dygrass_amd").csr_from_edges(V, u, v, w)
    G = DeviceGraph(ctx, V, rp, col, ww, slack=0)
    og = po.Graph(V, rp, col, ww)
    rng = np.random.default_rng(3)
    for _ in range(6):
        a = rng.integers(0, V, 25).astype(np.uint32)
        b = ((a + rng.integers(2, V - 1, 25)) % V).astype(np.uint32)
        x = rng.uniform(0.1, 10, 25).astype(np.float32)
        a[:10] = 5  # hammer one row
        G.insert(a, b, x)
        assert G.check() == 0
        for i in range(25):
            og.insert(int(a[i]), int(b[i]), float(x[i]))
    g = G.download()
    o = og.export()
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    assert np.array_equal(f32bits(g[2]), f32bits(o[2]))


def test_delete_swap_with_last_matches_oracle(ctx, mesh_wl):
    from dygrass_amd import DeviceGraph, synth

    po = _oracle()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "g.mtx")
        synth.write_mtx(path, mesh_wl.G.u, mesh_wl.G.v, mesh_wl.G.w, base=1)
        od = po.Dense(path)
        rp, col, w = csr(mesh_wl.G)
        G = DeviceGraph(ctx, mesh_wl.G.V, rp, col, w, slack=0, keyed=True)
        du, dv = mesh_wl.dels.u, mesh_wl.dels.v
        # include a repeated deletion and a non-edge: both must miss
        du = np.concatenate([du, du[:3], [0]]).astype(np.uint32)
        dv = np.concatenate([dv, dv[:3], [mesh_wl.G.V - 1]]).astype(np.uint32)
        hit = G.delete(du, dv)
        assert G.check() == 0
        ohit = np.array([od.remove(int(a), int(b)) for a, b in zip(du, dv)])
        assert np.array_equal(hit == 1, ohit == 1)
        g = G.download()
        o = od.graph.export()
        assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
        assert np.array_equal(f32bits(g[2]), f32bits(o[2]))


@pytest.mark.parametrize("per_edge", [False, True])
@pytest.mark.parametrize("prune", [True, False])
def test_decremental_walks_match_oracle(ctx, mesh_wl, prune, per_edge):
    from dygrass_amd import DeviceGraph, params, walk_decremental

    po = _oracle()
    rp, col, w = csr(mesh_wl.G)
    G = DeviceGraph(ctx, mesh_wl.G.V, rp, col, w, slack=0, keyed=True)
    n = min(300, len(mesh_wl.dels))
    du, dv = mesh_wl.dels.u[:n], mesh_wl.dels.v[:n]
    G.delete(du, dv)
    a, b = np.minimum(du, dv), np.maximum(du, dv)
    ids = np.arange(n, dtype=np.uint64) * 3 + 11
    got = walk_decremental(G, a, b, params(512, 100, seed=5, prune=prune, per_edge=per_edge),
                           edge_id=ids)
    grp, gcol, gw = G.download()
    og = po.Graph(mesh_wl.G.V, grp, gcol, gw)
    f, c, win, hops, path = og.walk_dec_batch(a, b, 512, seed=5, edge_id=ids, threads=8)
    assert np.array_equal(got.found_at, f)
    assert np.array_equal(f32bits(got.cond), f32bits(c))
    assert np.array_equal(got.winner, win)
    for i in range(n):
        if f[i] > 0:
            assert np.array_equal(got.path[i, : f[i] + 1], path[i, : f[i] + 1])
            assert got.path[i, 0] == a[i] and got.path[i, f[i]] == b[i]
    if not prune:
        assert np.array_equal(got.hops.astype(np.uint64), hops)


def test_heuristic_matches_oracle(ctx, mesh_wl):
    from dygrass_amd import DeviceGraph, walk_heuristic

    po = _oracle()
    rp, col, w = csr(mesh_wl.G)
    G = DeviceGraph(ctx, mesh_wl.G.V, rp, col, w, slack=0)
    ep = np.arange(0, mesh_wl.G.V, 97, dtype=np.uint32)
    keys = (np.arange(len(ep), dtype=np.uint64) << np.uint64(1)) | np.uint64(1)
    got = walk_heuristic(G, ep, keys, 16, seed=9)
    og = po.Graph(mesh_wl.G.V, rp, col, w)
    for i, (u, k) in enumerate(zip(ep, keys)):
        assert np.array_equal(got[i], og.heuristic(int(u), 16, 9, int(k)))


def test_walks_are_deterministic_and_prune_invariant(ctx, grid_wl):
    from dygrass_amd import DeviceGraph, params, walk_incremental

    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w)
    s, t, we = grid_wl.ins.u, grid_wl.ins.v, grid_wl.ins.w
    a = walk_incremental(H, s, t, we, params(256, 100, 10.0, seed=3, prune=True))
    b = walk_incremental(H, s, t, we, params(256, 100, 10.0, seed=3, prune=True))
    c = walk_incremental(H, s, t, we, params(256, 100, 10.0, seed=3, prune=False))
    d = walk_incremental(H, s, t, we, params(256, 100, 10.0, seed=3, per_edge=True))
    for x in (b, c, d):
        assert np.array_equal(a.found_at, x.found_at)
        assert np.array_equal(f32bits(a.cond), f32bits(x.cond))
        assert np.array_equal(a.winner, x.winner)


@pytest.mark.parametrize("W,block_edges", [(8, 1), (8, 7), (8, 64), (8, 256), (64, 3),
                                           (512, 256), (1, 256)])
@pytest.mark.parametrize("insert", [True, False])
def test_pool_shape_does_not_change_results(ctx, grid_wl, W, block_edges, insert):
    """Edges per workgroup is a launch-shape knob: ragged last workgroups,
    one edge per workgroup and 256 edges per workgroup all give the
    one-workgroup-per-edge kernel's results, hop counts included."""
    from dygrass_amd import DeviceGraph, params, walk_decremental, walk_incremental

    rp, col, w = csr(grid_wl.H)
    H = DeviceGraph(ctx, grid_wl.H.V, rp, col, w)
    n = 1001  # not a multiple of any block_edges above
    s, t, we = grid_wl.ins.u[:n], grid_wl.ins.v[:n], grid_wl.ins.w[:n]
    if insert:
        run = lambda p: walk_incremental(H, s, t, we, p)
    else:
        a, b = np.minimum(s, t), np.maximum(s, t)
        run = lambda p: walk_decremental(H, a, b, p)
    ref = run(params(W, 100, 10.0, seed=2, prune=False, per_edge=True))
    got = run(params(W, 100, 10.0, seed=2, prune=False, block_edges=block_edges))
    assert np.array_equal(got.found_at, ref.found_at)
    assert np.array_equal(f32bits(got.cond), f32bits(ref.cond))
    assert np.array_equal(got.winner, ref.winner)
    assert np.array_equal(got.hops, ref.hops)
    if not insert:
        for i in np.flatnonzero(ref.found_at > 0):
            k = ref.found_at[i] + 1
            assert np.array_equal(got.path[i, :k], ref.path[i, :k])


def test_incremental_update_dev_matches_host_entry(ctx, grid_wl):
    """Device-pointer entry point (inputs resident in HBM) equals the host one."""
    from dygrass_amd import (DeviceArray, DeviceGraph, DygError, incremental_update,
                             incremental_update_dev, params)

    rp, col, w = csr(grid_wl.H)
    p = params(512, 100, 10.0, seed=4)
    H1 = DeviceGraph(ctx, grid_wl.H.V, rp, col, w)
    f1, s1 = incremental_update(H1, grid_wl.ins.u, grid_wl.ins.v, grid_wl.ins.w, p)
    H2 = DeviceGraph(ctx, grid_wl.H.V, rp, col, w)
    m = len(grid_wl.ins)
    H2.reserve(m, (m + 9) // 10)
    ts = DeviceArray.from_numpy(ctx, grid_wl.ins.u)
    td = DeviceArray.from_numpy(ctx, grid_wl.ins.v)
    tw = DeviceArray.from_numpy(ctx, grid_wl.ins.w)
    tf = DeviceArray(ctx, m, np.int32)
    tc = DeviceArray(ctx, m, np.float32)
    s2 = incremental_update_dev(H2, ts.ptr, td.ptr, tw.ptr, m, tf.ptr, tc.ptr, p)
    assert np.array_equal(tf.numpy(), f1)
    # (hop counts are not compared: with pruning they depend on when a lane
    # sees another lane's find; the decisions do not)
    assert s2["added"] == s1["added"]
    g1, g2 = H1.download(), H2.download()
    assert all(np.array_equal(a, b) for a, b in zip(g1, g2))
    assert H1.check() == 0 and H2.check() == 0
    # an out-of-range id is rejected before any walk runs
    bad_src = grid_wl.ins.u.copy()
    bad_src[5] = grid_wl.H.V + 3
    tb = DeviceArray.from_numpy(ctx, bad_src)
    with pytest.raises(DygError):
        incremental_update_dev(H2, tb.ptr, td.ptr, tw.ptr, m, tf.ptr, tc.ptr, p)
