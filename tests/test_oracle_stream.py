"""The stream oracle pinned by hand-worked cases and by the pinned plain
incremental oracle (CPU only)."""
import numpy as np

from conftest import f32bits
from oracle import pyoracle as po
from oracle import stream as S


def _stream(gu, gv, gw, hidx, thr=10.0, W=8, seed=0):
    gu, gv, gw = np.array(gu), np.array(gv), np.array(gw, np.float32)
    h = np.array(hidx)
    return S.Stream(int(max(gu.max(), gv.max())) + 1, (gu, gv, gw), (gu[h], gv[h], gw[h]),
                    thr=thr, w_ins=W, w_del=W, seed=seed, threads=1)


def _edges(rows):
    return {(a, b) for a in range(rows.V) for b, _ in rows.rows[a] if a < b}


def test_weight_update_adds_in_order_in_g_and_h():
    st = _stream([0, 1], [1, 2], [1.0, 2.0], [0, 1])
    out = st.apply("ins", [1, 0], [0, 1], [0.1, 0.2])
    assert list(out) == [S.WEIGHT, S.WEIGHT]
    want = np.float32(np.float32(np.float32(1.0) + np.float32(0.1)) + np.float32(0.2))
    for g in (st.G, st.H):
        assert g.weight(0, 1) == g.weight(1, 0) == float(want)
        assert g.weight(1, 2) == 2.0


def test_insert_on_a_path_found_or_added_by_distortion():
    # path 0-1-2-3, all weights 1: every walker follows 0,1,2,3 (forced first
    # hop, redirects elsewhere); R after hop k is k
    for thr, expect in ((10.0, S.FOUND), (1.5, S.ADDED)):
        st = _stream([0, 1, 2], [1, 2, 3], [1, 1, 1], [0, 1, 2], thr=thr)
        out = st.apply("ins", [0, 3], [3, 0], [1.0, 2.0])
        # second occurrence of the new edge is a weight update of it
        assert list(out) == [expect, S.WEIGHT]
        assert st.G.weight(0, 3) == 3.0
        if expect == S.ADDED:  # R*w_e = 2 >= 1.5 at hop 2: no path, H gets w_first
            assert st.H.weight(0, 3) == 1.0
        else:  # R*w_e = 2 < 10 and t reached at hop 3
            assert not st.H.has(0, 3)


def test_insert_equals_plain_incremental_oracle(grid_wl):
    """An insertion-only stream of new edges, fed the plain updater's snapshot
    batches, makes the plain updater's decisions and H (slot order)."""
    wl = grid_wl
    n = 600
    u, v, w = wl.ins.u[:n], wl.ins.v[:n], wl.ins.w[:n]
    st = S.Stream(wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), thr=10.0,
                  w_ins=32, seed=4, threads=8)
    outs = []
    for b0 in po.batches(n, 10):
        b1 = min(n, b0 + (n + 9) // 10)
        outs.append(st.apply("ins", u[b0:b1], v[b0:b1], w[b0:b1]))
    out = np.concatenate(outs)
    rp, col, ww = S.Rows(wl.H.V, wl.H.u, wl.H.v, wl.H.w).csr()
    g = po.Graph(wl.H.V, rp, col, ww)
    found, _ = g.incremental(u, v, w, 10.0, 32, seed=4, threads=8)
    assert np.array_equal(out == S.ADDED, found == -1)
    erp, ecol, ew = g.export()
    orp, ocol, ow = st.H.csr()
    assert np.array_equal(erp, orp) and np.array_equal(ecol, ocol)
    assert np.array_equal(f32bits(ew), f32bits(ow))


def test_delete_on_a_cycle_recovers_the_other_arc():
    # 5-cycle; H = G.  Deleting (0,1) leaves the path 1-2-3-4-0: every walk
    # from 0 is forced along 0,4,3,2,1, whose edges H already holds
    st = _stream([0, 1, 2, 3, 4], [1, 2, 3, 4, 0], [1, 1, 1, 1, 1], [0, 1, 2, 3, 4])
    out = st.apply("del", [1], [0])
    assert list(out) == [S.RECOVERED]
    assert _edges(st.H) == {(1, 2), (2, 3), (3, 4), (0, 4)}
    assert _edges(st.G) == _edges(st.H)


def test_delete_recovers_a_chord_into_h():
    # 4-cycle with chord (0,2); H = the cycle.  Row 0 is [1, 3, 2]; deleting
    # (0,1) moves the last slot into the hole: [2, 3].  Every walk reaches 1
    # through 2, and the chord (0,2) or the arc 0-3-2 joins 0 and 1 in H again
    st = _stream([0, 1, 2, 3, 0], [1, 2, 3, 0, 2], [1, 2, 3, 4, 5], [0, 1, 2, 3], W=16)
    out = st.apply("del", [0], [1])
    assert list(out) == [S.RECOVERED]
    assert st.G.rows[0] == [[2, 5.0], [3, 4.0]]
    h = _edges(st.H)
    assert (0, 1) not in h and h <= _edges(st.G)
    # 0 and 1 connected in H again
    seen, todo = {0}, [0]
    while todo:
        x = todo.pop()
        for y, _ in st.H.rows[x]:
            if y not in seen:
                seen.add(y)
                todo.append(y)
    assert 1 in seen
    if (0, 2) in h:
        assert st.H.weight(0, 2) == 5.0  # G's weight


def test_delete_skip_g_only_invalid_and_repeats():
    st = _stream([0, 1, 2, 0], [1, 2, 3, 2], [1, 1, 1, 1], [0, 1, 2])
    out = st.apply("del", [0, 3, 2, 0, 1], [2, 0, 0, 2, 1], None)
    # (0,2) in G only; (3,0) absent; repeat of (0,2) skipped; self-loop
    assert list(out) == [S.G_ONLY, S.SKIP, S.SKIP, S.SKIP, S.INVALID]
    assert _edges(st.G) == {(0, 1), (1, 2), (2, 3)}
    assert _edges(st.H) == {(0, 1), (1, 2), (2, 3)}


def test_bridge_deletion_runs_the_heuristic():
    gu, gv = [0, 1, 2, 3, 4, 5, 2], [1, 2, 0, 4, 5, 3, 3]
    st = _stream(gu, gv, [1, 2, 3, 4, 5, 6, 7], [0, 1, 3, 4, 6], W=8, seed=3)
    out = st.apply("del", [2], [3])
    assert list(out) == [S.HEURISTIC]
    h = _edges(st.H)
    assert (2, 3) not in h and h <= _edges(st.G)
    # heuristic edges lie within two hops of the endpoints, in their triangles
    assert h - {(0, 1), (1, 2), (3, 4), (4, 5)} <= {(0, 2), (3, 5)}
    # each recovered edge carries G's weight
    for a, b in h:
        assert st.H.weight(a, b) == st.G.weight(a, b)


def _same_rows(a, b):
    rp, col, w = a
    orp, ocol, ow = b
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(f32bits(w), f32bits(ow))


def test_cpp_stream_equals_python_stream(grid_wl):
    """The C++ stream oracle (orc_stream_*, the full-size checker and CPU
    baseline) against the sequential Python one it restates: mixed batches
    with weight updates, in-batch repeats, self-loops, skipped and G-only
    deletions and path recovery; outcome codes and both graphs slot for slot
    after every batch, single- and multi-threaded."""
    from conftest import mixed_batches

    wl = grid_wl
    g, h = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)
    for threads, seed in ((1, 3), (8, 5)):
        py = S.Stream(wl.G.V, g, h, thr=10.0, w_ins=32, w_del=64, seed=seed, steps_ins=48,
                      steps_del=32)
        cc = po.StreamCpp(wl.G.V, g, h, thr=10.0, w_ins=32, w_del=64, seed=seed, steps_ins=48,
                          steps_del=32, threads=threads)
        kinds = set()
        for op, u, v, w in mixed_batches(wl, seed, n_ins=200, n_del=150):
            a = py.apply(op, u, v, w)
            assert np.array_equal(cc.apply(op, u, v, w), a)
            kinds.update(a.tolist())
            _same_rows(cc.csr(0), py.G.csr())
            _same_rows(cc.csr(1), py.H.csr())
        assert {S.WEIGHT, S.FOUND, S.ADDED, S.SKIP, S.G_ONLY, S.RECOVERED, S.INVALID} <= kinds
        st = cc.stats()
        assert st["batches"] == 4 and st["hops"] > 0
        cc.close()


def test_cpp_stream_heuristic_equals_python_stream():
    """Bridge deletions (no path, two-hop recovery) in both stream oracles."""
    # two triangles joined by the bridge 2-3; H = the two paths plus the bridge
    gu = [0, 1, 0, 3, 4, 3, 2]
    gv = [1, 2, 2, 4, 5, 5, 3]
    gw = np.array([1, 2, 3, 1, 2, 3, 0.5], np.float32)
    hidx = [0, 1, 3, 4, 6]
    py = _stream(gu, gv, gw, hidx, W=16, seed=2)
    gu, gv = np.array(gu), np.array(gv)
    h = np.array(hidx)
    cc = po.StreamCpp(6, (gu, gv, gw), (gu[h], gv[h], gw[h]), thr=10.0, w_ins=16, w_del=16,
                      seed=2, threads=2)
    a = py.apply("del", [3], [2])
    assert list(a) == [S.HEURISTIC]
    assert np.array_equal(cc.apply("del", [3], [2]), a)
    _same_rows(cc.csr(1), py.H.csr())
    _same_rows(cc.csr(0), py.G.csr())
