"""The oracle's decremental restatement against outcomes worked out by hand.

- swap-with-last with the reference's edge map
  (random_walk_decremental/gpu_graph.cuh:340-400) on a five-vertex graph;
- whole decremental runs (main.cu:318-468) on graphs where every walk is
  forced: a cycle (the deleted edge's only replacement is the rest of the
  cycle, so exactly the cycle edges missing from H are recovered) and a
  barbell (the deleted bridge has no replacement, so the two-hop recovery
  runs and only touches edges H already has);
- the same runs through the incremental restatement: a path sparsifier
  where inserted chords are found or not by the distortion rule alone.
"""
import os
import tempfile

import numpy as np

from oracle import pyoracle as po


def write(path, lines):
    with open(path, "w") as f:
        f.write("".join(" ".join(str(x) for x in ln) + "\n" for ln in lines))


def rows_of(g):
    rp, col, w = g.export()
    return [list(zip(col[rp[v]:rp[v + 1]].tolist(), w[rp[v]:rp[v + 1]].tolist()))
            for v in range(len(rp) - 1)]


def test_swap_with_last_by_hand():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.mtx")
        write(p, [(1, 2, 1), (1, 3, 2), (1, 4, 3), (1, 5, 4), (2, 3, 5)])
        g = po.Dense(p)
        assert g.multiplier == 10
        assert g.remove(0, 1) == 1
        # row 0 [1,2,3,4] -> last (4) into slot 0 -> [4,2,3]; row 1 [0,2] -> [2]
        assert g.remove(0, 2) == 1
        # row 0 [4,2,3] -> last (3) into slot 1 -> [4,3]; row 2 [0,1] -> [1]
        assert g.remove(1, 2) == -2  # vertex 1 would be left with degree 0
        assert g.remove(1, 0) == -1  # already gone
        assert rows_of(g.graph) == [[(4, 4.0), (3, 3.0)], [(2, 5.0)], [(1, 5.0)], [(0, 3.0)],
                                    [(0, 4.0)]]


def run_dec(lines_g, lines_h, dels):
    with tempfile.TemporaryDirectory() as d:
        write(os.path.join(d, "updated_dense.mtx"), lines_g)
        write(os.path.join(d, "updated_adj_sparse.mtx"), lines_h)
        write(os.path.join(d, "del.mtx"), dels)
        log = os.path.join(d, "log.txt")
        st = po.decremental_run(d, seed=0, threads=2, log_path=log)
        added = {tuple(int(x) for x in ln.split())
                 for ln in open(os.path.join(d, "added_edges.mtx")) if ln.strip()}
        return st, added, open(log).read().splitlines()


def test_decremental_cycle_recovers_missing_arc():
    n = 12
    G = [(i, i % n + 1, 1) for i in range(1, n + 1)]  # 1-2, ..., 12-1
    H = [e for e in G if e[:2] != (6, 7)]  # H lacks 6-7
    st, added, log = run_dec(G, H, [(1, 2)])
    # the only route from 1 to 2 without edge 1-2 is 1-12-11-...-3-2 (11 hops);
    # every walker takes it, so the recovered edges are the route's edges not in H
    assert added == {(6, 7)}
    assert st.filtered == 1 and st.no_path == 0 and st.added_edges == 1
    assert st.sparse_before == 11 and st.sparse_after == 11
    assert "No path count: 0" in log and "Added edges: 1" in log
    assert "Increased edges number: 0" in log


def test_decremental_bridge_takes_heuristic():
    G = [(1, 2, 1), (2, 3, 1), (1, 3, 1), (4, 5, 1), (5, 6, 1), (4, 6, 1), (3, 4, 1)]
    st, added, log = run_dec(G, G, [(3, 4)])
    assert added == set()
    assert st.no_path == 1 and st.filtered == 1
    assert st.sparse_before == 7 and st.sparse_after == 6  # recovery only re-adds H's own edges
    assert "No path count: 1" in log
    assert log.count(" result copied") == 2  # walk results, then recovery results
    assert "Increased edges number: %d" % (2 ** 64 - 1) in log  # size_t wrap, as the reference


def test_decremental_not_found_lines():
    G = [(1, 2, 1), (2, 3, 1), (3, 1, 1), (3, 4, 1), (4, 1, 1)]
    H = [(1, 2, 1), (2, 3, 1), (3, 4, 1)]
    st, added, log = run_dec(G, H, [(2, 4), (1, 3)])
    # (2,4) is not in G; (1,3) is in G but not in H: neither is walked
    assert "dense_map not found: a: 1 b: 3 key: 13" in log
    assert "sparse_map not found: a: 0 b: 2 key: 2" in log
    assert st.filtered == 0


def test_incremental_path_chords_by_distortion():
    # H is the path 0-1-...-9 with unit weights; a chord (0,k) of weight w is
    # found iff hop k is reached with R*w < thr at every hop 2..k: R_h = h
    with tempfile.TemporaryDirectory() as d:
        write(os.path.join(d, "adj_sparse.mtx"), [(i, i + 1, 1) for i in range(9)])
        chords = [(1, 4, 2.0), (1, 6, 2.0), (1, 9, 0.5), (1, 10, 3.0)]
        write(os.path.join(d, "ext.mtx"), chords)
        st = po.incremental_run(d, 10.0, 64, seed=0, threads=1)
        out = open(os.path.join(d, "updated_adj_sparse.mtx")).read().splitlines()
    # (0,3): R=3, 3*2=6 < 10 found.  (0,5): R at hop 5 = 5, 5*2 = 10 >= 10 -> not found.
    # (0,8): 8*0.5 = 4 found.  (0,9): hop 4 gives 4*3 = 12 >= 10 -> not found.
    assert st.n_added == 2
    assert out[-2:] == ["1 6 2", "1 10 3"]
    assert out[:9] == [f"{i + 1} {i + 2} 1" for i in range(9)]


def test_decremental_lines_equals_file_run():
    """The array entry (used at full size) equals the file run on the same data."""
    from dygrass_amd import synth

    wl = synth.config_mesh_delete(0.002, 3)
    with tempfile.TemporaryDirectory() as root:
        d = synth.write_dataset(root, wl)
        st_file = po.decremental_run(d, seed=1, threads=4)
        added_file = {tuple(int(x) - 1 for x in ln.split())
                      for ln in open(os.path.join(d, "added_edges.mtx")) if ln.strip()}
    st, added = po.decremental_lines(wl.G.V, wl.G.u, wl.G.v, wl.G.w, wl.H.u, wl.H.v,
                                     wl.dels.u, wl.dels.v, seed=1, threads=4)
    assert {tuple(x) for x in added.tolist()} == added_file
    for k in ("added_edges", "sparse_before", "sparse_after", "no_path", "filtered", "hops"):
        assert getattr(st, k) == getattr(st_file, k), k
    assert st.filtered > 0
