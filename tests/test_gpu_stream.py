"""GPU parity of the mixed insertion/deletion stream against the oracle.

Each batch's per-update outcome codes must be equal, and after each batch G
and H must be equal slot for slot (neighbour order within every row, weight
bits), which pins the walks, the duplicate handling, the weight updates, the
swap-with-last deletions and the recovery order together.
"""
import numpy as np
import pytest

from conftest import f32bits, mixed_batches

pytestmark = pytest.mark.gpu


def _assert_same_graph(dev_graph, rows):
    rp, col, w = dev_graph.download()
    orp, ocol, ow = rows.csr()
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(f32bits(w), f32bits(ow))
    assert dev_graph.check() == 0


@pytest.mark.parametrize("seed,plain,general,nodesc,graph",
                         [(0, False, False, False, False), (1, False, False, False, False),
                          (0, True, False, False, False), (2, False, True, False, False),
                          (3, False, True, True, False), (4, False, False, True, False),
                          (6, False, False, False, False), (6, False, False, False, True),
                          (7, False, False, True, True)])
def test_stream_matches_oracle(ctx, grid_wl, seed, plain, general, nodesc, graph):
    """Device-resident form (batches through the one-workgroup path, directly
    or as captured step graphs, or everything through the general pipeline;
    G with or without neighbour descriptors) and the plain form."""
    from dygrass_amd import Stream, stream_params
    from oracle import stream as S

    wl = grid_wl
    sp = stream_params(w_ins=64, w_del=64, threshold=10.0, seed=seed, max_batch=1100,
                       max_inserts=4096, plain=plain, general_only=general, dense_no_desc=nodesc,
                       graph=graph)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    orc = S.Stream(wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), thr=10.0,
                   w_ins=64, w_del=64, seed=seed)
    seen = set()
    for op, u, v, w in mixed_batches(wl, seed, n_ins=300 if seed != 1 else 1000):
        got, stats = st.apply(op, u, v, w)
        want = orc.apply(op, u, v, w)
        assert np.array_equal(got, want), (op, np.flatnonzero(got != want)[:10])
        assert stats["updates"] == len(u)
        seen.update(np.unique(got).tolist())
        _assert_same_graph(st.G, orc.G)
        _assert_same_graph(st.H, orc.H)
    assert st.position == orc.base
    # the batches exercise every outcome
    assert seen >= {S.WEIGHT, S.FOUND, S.ADDED, S.SKIP, S.G_ONLY, S.RECOVERED, S.INVALID}


@pytest.mark.parametrize("general", [False, True])
def test_stream_many_walkers_match_oracle(ctx, grid_wl, general):
    """512 insertion and 1024 deletion walkers per edge (two and four walkers
    per lane of a workgroup) on small batches: outcomes, G and H equal the
    oracle's after every batch."""
    from dygrass_amd import Stream, stream_params
    from oracle import stream as S

    wl = grid_wl
    seed = 9
    sp = stream_params(w_ins=512, w_del=1024, threshold=10.0, seed=seed, max_batch=1100,
                       max_inserts=4096, general_only=general, steps_ins=60, steps_del=40)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    orc = S.Stream(wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), thr=10.0,
                   w_ins=512, w_del=1024, seed=seed, steps_ins=60, steps_del=40)
    seen = set()
    for op, u, v, w in mixed_batches(wl, seed, n_ins=200, n_del=120):
        got, _ = st.apply(op, u, v, w)
        want = orc.apply(op, u, v, w)
        assert np.array_equal(got, want), (op, np.flatnonzero(got != want)[:10])
        seen.update(np.unique(got).tolist())
        _assert_same_graph(st.G, orc.G)
        _assert_same_graph(st.H, orc.H)
    assert seen >= {S.FOUND, S.ADDED, S.RECOVERED}


@pytest.mark.parametrize("nodesc", [False, True])
def test_step_graphs_replay_like_direct_launches(ctx, grid_wl, nodesc):
    """Captured step graphs (DYG_STREAM_GRAPH) against direct launches on the
    same batches: sizes in every size class and at its edges (1, 16, 17, 64,
    65, 256, 257, 1024, 1025, 2048) and beyond the one-workgroup limit, each
    class replayed several times with eager batches and waits in between;
    every batch's outcome codes and walk statistics, and G and H after every
    batch, must be equal, and the captured walk timing events must fire."""
    from dygrass_amd import Stream, stream_params

    wl = grid_wl
    rng = np.random.default_rng(11)
    G = wl.G
    sizes = [1, 16, 17, 64, 65, 256, 257, 1024, 1025, 2048, 3000, 1, 5, 16, 300, 2048, 40, 1]
    batches, k = [], 0
    for i, n in enumerate(sizes):
        if i % 3 == 2:  # deletions: G edges, sometimes H edges
            idx = rng.choice(len(G), n, replace=False)
            batches.append(("del", G.u[idx], G.v[idx], None))
        else:
            a = rng.integers(0, G.V, n)
            b = rng.integers(0, G.V, n)
            batches.append(("ins", a, b, rng.uniform(0.1, 10.0, n).astype(np.float32)))
            k += n
    streams = []
    for graph in (False, True):
        sp = stream_params(w_ins=64, w_del=32, threshold=10.0, seed=4, max_batch=3000,
                           max_inserts=k + 16, dense_no_desc=nodesc, graph=graph,
                           steps_del=40)
        streams.append(Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w),
                              sp))
    direct, graphed = streams
    for op, u, v, w in batches:
        o1, s1 = direct.apply(op, u, v, w)
        o2, s2 = graphed.apply(op, u, v, w)
        assert np.array_equal(o1, o2), (op, len(u))
        # (hop counts are work, not results: pruning depends on lane timing)
        for key in ("found", "added", "recovered", "heuristic", "candidates"):
            assert s1[key] == s2[key], (op, len(u), key)
        assert s2["walk_ms"] > 0.0 or s2["hops"] == 0
        assert s1["graph"] == 0 and s2["graph"] == (1 if len(u) <= 2048 else 0), (op, len(u))
        for g1, g2 in ((direct.G, graphed.G), (direct.H, graphed.H)):
            a1, a2 = g1.download(), g2.download()
            assert all(np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                                      y.view(np.uint32) if y.dtype == np.float32 else y)
                       for x, y in zip(a1, a2))
        assert graphed.H.check() == 0 and graphed.G.check() == 0


@pytest.mark.parametrize("general", [False, True])
def test_stream_tuned_settings_match_oracle(ctx, grid_wl, general):
    """The settings the autotuner moves, set apart per walk kind: walkers,
    step budgets and workgroup shape of the insertion walks differ from the
    deletion walks'; the step budgets are short enough that they cut walks
    and paths (found_at near the bound, longer paths unrecovered)."""
    from dygrass_amd import Stream, stream_params
    from oracle import stream as S

    wl = grid_wl
    seed = 5
    sp = stream_params(w_ins=128, w_del=32, threshold=10.0, seed=seed, max_batch=1100,
                       max_inserts=4096, general_only=general, steps_ins=12, steps_del=7,
                       block_edges_ins=3, block_edges_del=5, dense_no_desc=True)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    orc = S.Stream(wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), thr=10.0,
                   w_ins=128, w_del=32, seed=seed, steps_ins=12, steps_del=7)
    seen = set()
    for op, u, v, w in mixed_batches(wl, seed):
        got, _ = st.apply(op, u, v, w)
        want = orc.apply(op, u, v, w)
        assert np.array_equal(got, want), (op, np.flatnonzero(got != want)[:10])
        seen.update(np.unique(got).tolist())
        _assert_same_graph(st.G, orc.G)
        _assert_same_graph(st.H, orc.H)
    assert seen >= {S.FOUND, S.ADDED, S.RECOVERED}


@pytest.mark.parametrize("plain,nodesc", [(False, False), (True, False), (False, True)])
def test_stream_heuristic_recovery(ctx, plain, nodesc):
    """Deleting a bridge of G that H holds leaves no path: the heuristic runs."""
    from dygrass_amd import Stream, stream_params
    from oracle import stream as S

    # two triangles joined by the bridge (2,3); H = a spanning tree with it
    gu = np.array([0, 1, 2, 3, 4, 5, 2], np.uint32)
    gv = np.array([1, 2, 0, 4, 5, 3, 3], np.uint32)
    gw = np.array([1, 2, 3, 4, 5, 6, 7], np.float32)
    hu, hv, hw = gu[[0, 1, 3, 4, 6]], gv[[0, 1, 3, 4, 6]], gw[[0, 1, 3, 4, 6]]
    sp = stream_params(w_ins=8, w_del=8, threshold=10.0, seed=3, max_batch=16, max_inserts=64,
                       plain=plain, dense_no_desc=nodesc)
    st = Stream(ctx, 6, (gu, gv, gw), (hu, hv, hw), sp)
    orc = S.Stream(6, (gu, gv, gw), (hu, hv, hw), thr=10.0, w_ins=8, w_del=8, seed=3)
    got, _ = st.apply("del", [2], [3])
    want = orc.apply("del", [2], [3])
    assert got[0] == want[0] == S.HEURISTIC
    _assert_same_graph(st.G, orc.G)
    _assert_same_graph(st.H, orc.H)


def test_deletion_stream_equals_decremental_updater(ctx, mesh_wl):
    """A deletion-only stream fed the decremental updater's snapshot batches
    recovers what the decremental oracle (random_walk_decremental restated)
    recovers: same sparsifier size; every recovered path edge is in the final
    H unless a later batch deleted it (the reference's added set keeps it)."""
    from dygrass_amd import Stream, stream_params
    from oracle import pyoracle as po

    wl = mesh_wl
    G, H, D = wl.G, wl.H, wl.dels
    st_o, added = po.decremental_lines(G.V, G.u, G.v, G.w, H.u, H.v, D.u, D.v, seed=0,
                                       threads=8)
    sp = stream_params(w_ins=512, w_del=512, threshold=0.0, seed=0, max_batch=len(D),
                       max_inserts=4 * len(D))
    st = Stream(ctx, G.V, (G.u, G.v, G.w), (H.u, H.v, H.w), sp)
    M = len(D)
    I = (M + 9) // 10
    for b0 in range(0, M, I):
        st.apply("del", D.u[b0:b0 + I], D.v[b0:b0 + I])
    rp, col, _ = st.H.download()
    rows = np.repeat(np.arange(G.V), np.diff(rp).astype(np.int64))
    h = set(zip(rows[col > rows].tolist(), col[col > rows].tolist()))
    assert len(h) == int(st_o.sparse_after)
    gone = set(zip(np.minimum(D.u, D.v).tolist(), np.maximum(D.u, D.v).tolist()))
    for a, b in added.tolist():
        e = (min(a, b), max(a, b))
        assert (e in h) == (e not in gone), e
    assert st.G.check() == 0 and st.H.check() == 0


@pytest.mark.parametrize("n_del", [40, 800])
def test_small_deletion_batch_recovery_sizes(ctx, grid_wl, n_del):
    """A deletion batch of H edges within the one-workgroup size: a short
    candidate list is recovered inside the post kernel, a long one (2804
    candidates, more than that kernel holds) is handed to the general
    recovery; both match the oracle, and so does the insertion batch after
    them."""
    from dygrass_amd import Stream, stream_params
    from oracle import stream as S

    wl = grid_wl
    rng = np.random.default_rng(n_del)
    sp = stream_params(w_ins=64, w_del=64, threshold=10.0, seed=5, max_batch=1024,
                       max_inserts=4096)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    orc = S.Stream(wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), thr=10.0,
                   w_ins=64, w_del=64, seed=5)
    idx = rng.choice(len(wl.H), n_del, replace=False)
    batches = [("del", wl.H.u[idx], wl.H.v[idx], None),
               ("ins", wl.ins.u[:500], wl.ins.v[:500], wl.ins.w[:500])]
    for op, u, v, w in batches:
        got, stats = st.apply(op, u, v, w)
        want = orc.apply(op, u, v, w)
        assert np.array_equal(got, want), (op, np.flatnonzero(got != want)[:10])
        if op == "del":
            assert (stats["candidates"] > 2048) == (n_del == 800), stats["candidates"]
        _assert_same_graph(st.G, orc.G)
        _assert_same_graph(st.H, orc.H)


@pytest.mark.parametrize("nodesc", [False, True])
def test_batches_up_to_the_one_workgroup_limit(ctx, grid_wl, nodesc):
    """Batches near the one-workgroup size (2048 updates, two per lane) and
    tiny ones (ranked by counting) match the oracle: insertion batches of
    1900 and 50, deletion batches of 2048 (G edges, repeats, non-edges) and 30."""
    from dygrass_amd import Stream, stream_params
    from oracle import stream as S

    wl = grid_wl
    rng = np.random.default_rng(7)
    sp = stream_params(w_ins=64, w_del=64, threshold=10.0, seed=9, max_batch=4096,
                       max_inserts=8192, dense_no_desc=nodesc)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    orc = S.Stream(wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), thr=10.0,
                   w_ins=64, w_del=64, seed=9)
    gi = rng.choice(len(wl.G), 2000, replace=False)
    du = np.concatenate([wl.G.u[gi], rng.integers(0, wl.G.V, 40), wl.G.u[gi[:8]]])
    dv = np.concatenate([wl.G.v[gi], rng.integers(0, wl.G.V, 40), wl.G.v[gi[:8]]])
    hi = rng.choice(len(wl.H), 30, replace=False)
    batches = [("ins", wl.ins.u[:1900], wl.ins.v[:1900], wl.ins.w[:1900]),
               ("del", du[:2048], dv[:2048], None),
               ("ins", wl.ins.u[1900:1950], wl.ins.v[1900:1950], wl.ins.w[1900:1950]),
               ("del", wl.H.u[hi], wl.H.v[hi], None)]
    for op, u, v, w in batches:
        got, _ = st.apply(op, u, v, w)
        want = orc.apply(op, u, v, w)
        assert np.array_equal(got, want), (op, len(u), np.flatnonzero(got != want)[:10])
        _assert_same_graph(st.G, orc.G)
        _assert_same_graph(st.H, orc.H)


def _dev(ctx, a, dtype):
    from dygrass_amd import DeviceArray

    return DeviceArray.from_numpy(ctx, np.ascontiguousarray(np.asarray(a, dtype)))


@pytest.mark.parametrize("nodesc", [False, True])
def test_apply_dev_matches_host_inputs(ctx, grid_wl, nodesc):
    """dyg_stream_apply_dev (batch already in device memory) against
    dyg_stream_apply (host arrays) on the same mixed batches, one-workgroup
    and general sizes: outcome codes, G and H equal after every batch."""
    from dygrass_amd import Stream, stream_params

    wl = grid_wl
    batches = mixed_batches(wl, 8, n_ins=300)
    rng = np.random.default_rng(8)
    big = rng.choice(len(wl.G), 2600, replace=False)
    batches.append(("del", wl.G.u[big], wl.G.v[big], None))
    a = rng.integers(0, wl.G.V, 2500)
    b = rng.integers(0, wl.G.V, 2500)
    batches.append(("ins", a, b, rng.uniform(0.1, 10, 2500).astype(np.float32)))
    mk = lambda: Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w),
                        stream_params(w_ins=64, w_del=64, threshold=10.0, seed=8, max_batch=3000,
                                      max_inserts=8192, dense_no_desc=nodesc))
    host, dev = mk(), mk()
    for op, u, v, w in batches:
        o1, _ = host.apply(op, u, v, w)
        du, dv = _dev(ctx, u, np.uint32), _dev(ctx, v, np.uint32)
        dw = _dev(ctx, w, np.float32) if w is not None else None
        o2, _ = dev.apply_dev(op, du, dv, dw)
        assert np.array_equal(o1, o2), (op, len(u))
        _assert_same_graph(dev.G, _rows_of(host.G))
        _assert_same_graph(dev.H, _rows_of(host.H))
    assert dev.position == host.position


class _rows_of:
    """A device graph's download in the form _assert_same_graph expects."""

    def __init__(self, g):
        self._csr = g.download()

    def csr(self):
        return self._csr


@pytest.mark.parametrize("n", [10, 2500])
@pytest.mark.parametrize("op", ["ins", "del"])
def test_apply_dev_rejects_out_of_range_ids(ctx, grid_wl, n, op):
    """A device batch holding an id >= V fails with DYG_ERR_ARG and changes
    nothing (G, H, position); the next valid batch then gives what a stream
    that never saw the bad batch gives."""
    from dygrass_amd import DygError, Stream, stream_params

    wl = grid_wl
    rng = np.random.default_rng(n)
    mk = lambda: Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w),
                        stream_params(w_ins=32, w_del=32, threshold=10.0, seed=2, max_batch=3000,
                                      max_inserts=8192))
    st, ref = mk(), mk()
    if op == "ins":
        u = rng.integers(0, wl.G.V, n)
        v = rng.integers(0, wl.G.V, n)
        w = rng.uniform(0.1, 10, n).astype(np.float32)
    else:
        idx = rng.choice(len(wl.G), n, replace=False)
        u, v, w = wl.G.u[idx], wl.G.v[idx], None
    bad_u = np.array(u, np.uint32)
    bad_u[n // 2] = wl.G.V  # one id out of range
    before = (st.G.download(), st.H.download())
    with pytest.raises(DygError, match="out of range"):
        st.apply_dev(op, _dev(ctx, bad_u, np.uint32), _dev(ctx, v, np.uint32),
                     _dev(ctx, w, np.float32) if w is not None else None)
    for g, b in zip((st.G, st.H), before):
        a = g.download()
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert st.position == 0
    o1, _ = st.apply_dev(op, _dev(ctx, u, np.uint32), _dev(ctx, v, np.uint32),
                         _dev(ctx, w, np.float32) if w is not None else None)
    o2, _ = ref.apply(op, u, v, w)
    assert np.array_equal(o1, o2)
    _assert_same_graph(st.H, _rows_of(ref.H))
    _assert_same_graph(st.G, _rows_of(ref.G))


def test_apply_many_dev_matches_apply_dev(ctx, grid_wl):
    """dyg_stream_apply_many_dev (a batch list in one call) against per-batch
    dyg_stream_apply_dev: outcome codes, per-batch stats, G and H equal; bad
    offsets are rejected before anything changes."""
    from dygrass_amd import DygError, Stream, stream_params

    wl = grid_wl
    batches = mixed_batches(wl, 12, n_ins=300)
    rng = np.random.default_rng(12)
    big = rng.choice(len(wl.G), 2600, replace=False)
    batches.append(("del", wl.G.u[big], wl.G.v[big], None))
    batches.append(("ins", wl.ins.u[1200:1210], wl.ins.v[1200:1210], wl.ins.w[1200:1210]))
    mk = lambda: Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w),
                        stream_params(w_ins=64, w_del=64, threshold=10.0, seed=12,
                                      max_batch=3000, max_inserts=8192))
    one, many = mk(), mk()
    want, wst = [], []
    for op, u, v, w in batches:
        o, st = one.apply_dev(op, _dev(ctx, u, np.uint32), _dev(ctx, v, np.uint32),
                              _dev(ctx, w, np.float32) if w is not None else None)
        want.append(o)
        wst.append(st)
    off = np.concatenate([[0], np.cumsum([len(b[1]) for b in batches])]).astype(np.uint64)
    U = np.concatenate([np.asarray(b[1], np.uint32) for b in batches])
    Vv = np.concatenate([np.asarray(b[2], np.uint32) for b in batches])
    Wt = np.concatenate([np.asarray(b[3], np.float32) if b[3] is not None
                         else np.zeros(len(b[1]), np.float32) for b in batches])
    du, dv, dw = _dev(ctx, U, np.uint32), _dev(ctx, Vv, np.uint32), _dev(ctx, Wt, np.float32)
    ops = [b[0] for b in batches]
    bad = off.copy()
    bad[2] = bad[1] - 1  # a batch ending before it starts
    with pytest.raises(DygError, match="offsets"):
        many.apply_many_dev(ops[:2], bad[:3], du, dv, dw)
    assert many.position == 0
    got, gst = many.apply_many_dev(ops, off, du, dv, dw)
    assert np.array_equal(got, np.concatenate(want))
    for op, a, b in zip(ops, gst, wst):
        for k in ("updates", "found", "added", "skipped", "g_only", "recovered", "heuristic",
                  "weight_updates", "invalid"):
            assert a[k] == b[k], k
        # spans from the device clock; insertion walks timed, deletion walks not
        assert a["total_ms"] > 0.0
        if op == "ins":
            assert a["walk_ms"] > 0.0 or a["hops"] == 0
            assert a["walk_ms"] < a["total_ms"] or a["hops"] == 0
        else:
            assert a["walk_ms"] == 0.0
    _assert_same_graph(many.G, _rows_of(one.G))
    _assert_same_graph(many.H, _rows_of(one.H))
    assert many.position == one.position


def test_apply_many_dev_checks_ids_first_and_runs_lists_in_turn(ctx, grid_wl):
    """A list holding an out-of-range id anywhere is refused before anything
    changes; an empty batch inside a list is a no-op; two lists in a row give
    what one list of all their batches gives (the reports, events and arena
    headroom carry over between calls)."""
    from dygrass_amd import DygError, Stream, stream_params

    wl = grid_wl
    batches = mixed_batches(wl, 21, n_ins=250, n_del=150)
    batches.insert(2, ("ins", wl.ins.u[:0], wl.ins.v[:0], wl.ins.w[:0]))
    mk = lambda: Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w),
                        stream_params(w_ins=64, w_del=64, threshold=10.0, seed=21,
                                      max_batch=3000, max_inserts=8192))

    def dev_list(bs):
        off = np.concatenate([[0], np.cumsum([len(b[1]) for b in bs])]).astype(np.uint64)
        U = np.concatenate([np.asarray(b[1], np.uint32) for b in bs])
        Vv = np.concatenate([np.asarray(b[2], np.uint32) for b in bs])
        Wt = np.concatenate([np.asarray(b[3], np.float32) if b[3] is not None
                             else np.zeros(len(b[1]), np.float32) for b in bs])
        return ([b[0] for b in bs], off, _dev(ctx, U, np.uint32), _dev(ctx, Vv, np.uint32),
                _dev(ctx, Wt, np.float32))

    whole, split = mk(), mk()
    ops, off, du, dv, dw = dev_list(batches)
    want, _ = whole.apply_many_dev(ops, off, du, dv, dw)

    bad = [list(b) for b in batches[:3]]
    bad[2] = ("del", np.array([0, wl.G.V + 3], np.uint32), np.array([1, 2], np.uint32), None)
    with pytest.raises(DygError, match="out of range"):
        split.apply_many_dev(*dev_list(bad))
    assert split.position == 0
    fresh = mk()
    _assert_same_graph(split.G, _rows_of(fresh.G))
    _assert_same_graph(split.H, _rows_of(fresh.H))

    got = []
    for part in (batches[:3], batches[3:]):
        o, sts = split.apply_many_dev(*dev_list(part))
        got.append(o)
        assert [s_["updates"] for s_ in sts] == [len(b[1]) for b in part]
    assert np.array_equal(np.concatenate(got), want)
    _assert_same_graph(split.G, _rows_of(whole.G))
    _assert_same_graph(split.H, _rows_of(whole.H))
    assert split.position == whole.position == int(off[-1])


@pytest.mark.parametrize("seed,nodesc,general,small_1024",
                         [(31, False, True, False), (32, True, True, False),
                          (33, False, False, False), (34, True, False, False),
                          (35, False, False, True)])
def test_apply_many_dev_orders_ahead_match_oracle(ctx, grid_wl, seed, nodesc, general,
                                                  small_1024):
    """A list through dyg_stream_apply_many_dev, every batch in the general
    pipeline or in the one-workgroup kernels: each batch's keys and row
    entries are sorted ahead of it on the prep stream, and its kernels take
    the presorted orders instead of sorting their own.  Outcome codes of every
    batch and the final G and H (slot for slot) equal the oracle's, batch
    after batch.  small_1024: the 1200-update batch takes the general pipeline
    between one-workgroup batches."""
    from dygrass_amd import Stream, stream_params
    from oracle import stream as S

    wl = grid_wl
    batches = mixed_batches(wl, seed, n_ins=400, n_del=250)
    # a batch in the larger one-workgroup class (two updates per lane)
    batches.insert(2, ("ins", wl.ins.u[800:2000], wl.ins.v[800:2000], wl.ins.w[800:2000]))
    sp = stream_params(w_ins=64, w_del=64, threshold=10.0, seed=seed, max_batch=1300,
                       max_inserts=8192, general_only=general, dense_no_desc=nodesc,
                       small_1024=small_1024)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    orc = S.Stream(wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), thr=10.0,
                   w_ins=64, w_del=64, seed=seed)
    off = np.concatenate([[0], np.cumsum([len(b[1]) for b in batches])]).astype(np.uint64)
    U = np.concatenate([np.asarray(b[1], np.uint32) for b in batches])
    Vv = np.concatenate([np.asarray(b[2], np.uint32) for b in batches])
    Wt = np.concatenate([np.asarray(b[3], np.float32) if b[3] is not None
                         else np.zeros(len(b[1]), np.float32) for b in batches])
    got, _ = st.apply_many_dev([b[0] for b in batches], off, _dev(ctx, U, np.uint32),
                               _dev(ctx, Vv, np.uint32), _dev(ctx, Wt, np.float32))
    seen = set()
    for i, (op, u, v, w) in enumerate(batches):
        want = orc.apply(op, u, v, w)
        g = got[off[i]:off[i + 1]]
        assert np.array_equal(g, want), (i, op, np.flatnonzero(g != want)[:10])
        seen.update(np.unique(g).tolist())
    _assert_same_graph(st.G, orc.G)
    _assert_same_graph(st.H, orc.H)
    assert seen >= {S.WEIGHT, S.FOUND, S.ADDED, S.SKIP, S.G_ONLY, S.RECOVERED, S.INVALID}
