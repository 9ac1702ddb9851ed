"""The oracle's walks and reduction against what the mathematics fixes.

- path graphs: the non-backtracking walk is forced, so found_at, R (the
  float32 running sum of 1/w) and the distortion cut are known in closed form;
- leaves stop a walk without a draw; max_steps bounds the hop index at 99;
- cycles: the first draw picks the direction, the winner is the best arc with
  the bit-reversed tie rule;
- the block reduction equals "largest conductance, then smallest bit-reversed
  walker index" (the closed form of random_walk_incremental/main.cu:183-197);
- on small random graphs the probability that a walk reaches the target,
  computed exactly by dynamic programming over (vertex, previous vertex)
  states, matches the oracle's empirical rate.
"""
import math

import numpy as np
import pytest

from dygrass_amd.core import csr_from_edges
from oracle import pyoracle as po


def graph(V, edges):
    u = [a for a, b, _ in edges]
    v = [b for a, b, _ in edges]
    w = [x for _, _, x in edges]
    rp, col, ww = csr_from_edges(V, u, v, w)
    return po.Graph(V, rp, col, ww), (rp, col, ww)


def path_graph(n, seed=0):
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 2.0, n - 1).astype(np.float32)
    return graph(n, [(i, i + 1, float(w[i])) for i in range(n - 1)]), w


def fsum_inv(ws):
    R = np.float32(0)
    for x in ws:
        R = np.float32(R + np.float32(np.float32(1) / np.float32(x)))
    return R


@pytest.mark.parametrize("k", [1, 2, 7, 50, 99])
def test_path_graph_forced_walk(k):
    (g, _), w = path_graph(120)
    f, c, win, hops = g.walk_inc(0, k, 1.0, 1e30, 64)
    R = fsum_inv(w[:k])
    assert f == k
    assert np.float32(c) == np.float32(np.float32(1) / R)
    assert win == 0  # all walkers tie; bit-reversed index 0 wins
    assert hops == 64 * k
    fd, cd, wind, _, path = g.walk_dec(0, k, 64)
    assert fd == k and np.float32(cd) == np.float32(np.float32(1) / R) and wind == 0
    assert path[: k + 1].tolist() == list(range(k + 1))


def test_path_graph_step_bound():
    (g, _), _ = path_graph(120)
    assert g.walk_inc(0, 99, 1.0, 1e30, 4)[0] == 99
    assert g.walk_inc(0, 100, 1.0, 1e30, 4)[0] == -1
    assert g.walk_dec(0, 100, 4)[0] == -1
    assert g.walk_inc(0, 100, 1.0, 1e30, 4)[3] == 4 * 99  # 99 draws, then the bound


@pytest.mark.parametrize("L", [2, 7, 32, 48])
def test_path_graph_step_budget(L):
    """A step budget L (the tuner's 32 and 48 among them) bounds the hop index
    at L - 1, for insertion and deletion walks alike: on a path the route is
    unique, so a target d hops away is found at hop d exactly when d <= L - 1,
    after d draws per walker; otherwise every walker stops after L - 1 draws."""
    (g, _), _ = path_graph(120)
    for d in (L - 1, L):
        f, _, _, hops = g.walk_inc(0, d, 1.0, 1e30, 8, max_steps=L)
        fd, _, _, _, _ = g.walk_dec(0, d, 8, max_steps=L)
        if d <= L - 1:
            assert f == d and fd == d and hops == 8 * d
        else:
            assert f == -1 and fd == -1 and hops == 8 * (L - 1)


@pytest.mark.parametrize("thr", [3.0, 10.0, 25.0])
def test_distortion_cut_on_path(thr):
    (g, _), w = path_graph(60, seed=4)
    we = np.float32(1.7)
    for k in range(1, 50):
        f = g.walk_inc(0, k, float(we), thr, 2)[0]
        # hop 1 has no distortion test; hop h>=2 stops when R_h * w_e >= thr,
        # and that test comes before the target test
        ok = all(np.float32(fsum_inv(w[:h]) * we) < np.float32(thr) for h in range(2, k + 1))
        assert f == (k if ok else -1), k


def test_leaf_stops_without_draw():
    g, _ = graph(6, [(0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.0), (4, 5, 1.0)])
    f, c, win, hops = g.walk_inc(0, 5, 1.0, 1e30, 16)
    assert f == -1 and c == -1.0 and win == 0
    assert hops == 16 * 3  # draws at 0, 1, 2; vertex 3 is a leaf


def bitrev(j, bits):
    return int(format(j, f"0{bits}b")[::-1], 2) if bits else 0


def test_cycle_direction_and_tie_rule():
    n = 12
    rng = np.random.default_rng(1)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    g, _ = graph(n, [(i, (i + 1) % n, float(w[i])) for i in range(n)])
    s, t, W, seed, eid = 0, 5, 64, 3, 17
    # row 0 holds [1, n-1]: draw % 2 == 0 goes clockwise
    cw = fsum_inv(w[0:5])
    ccw = fsum_inv([w[(n - 1 - i) % n] for i in range(n - t)])
    dirs = [po.draw(seed, po.TAG_DELETE, eid, j, 0) % 2 for j in range(W)]
    c_cw, c_ccw = np.float32(1) / cw, np.float32(1) / ccw
    best = 0 if c_cw >= c_ccw else 1
    cands = [j for j in range(W) if dirs[j] == best]
    f, c, win, _, path = g.walk_dec(s, t, W, seed=seed, eid=eid)
    assert win == min(cands, key=lambda j: bitrev(j, 6))
    assert f == (5 if best == 0 else n - t)
    assert np.float32(c) == (c_cw if best == 0 else c_ccw)


@pytest.mark.parametrize("W", [1, 2, 4, 8, 64, 256, 1024])
def test_tree_reduction_closed_form(W):
    rng = np.random.default_rng(W)
    bits = int(math.log2(W))
    for _ in range(50):
        c = rng.choice(np.array([-1.0, 0.25, 0.5, 0.75], np.float32), W)
        got = po.tree_reduce(c)
        mx = c.max()
        expect = min((j for j in range(W) if c[j] == mx), key=lambda j: bitrev(j, bits))
        assert got == expect


def exact_hit_probability(rows, s, t, max_steps):
    """P(walk from s reaches t) by dynamic programming over (cur, prev)."""
    p_found = 0.0
    state = {}
    d = len(rows[s])
    for x in rows[s]:
        if x == t:
            p_found += 1.0 / d
        else:
            state[(x, s)] = state.get((x, s), 0.0) + 1.0 / d
    step = 2
    while step < max_steps and state:
        nxt = {}
        for (cur, prev), p in state.items():
            row = rows[cur]
            dg = len(row)
            if dg == 1:
                continue
            for i in range(dg - 1):
                x = row[i] if row[i] != prev else row[dg - 1]
                q = p / (dg - 1)
                if x == t:
                    p_found += q
                else:
                    nxt[(x, cur)] = nxt.get((x, cur), 0.0) + q
        state = nxt
        step += 1
    return p_found


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hit_probability_brute_force(seed):
    rng = np.random.default_rng(seed)
    V = 9
    edges = set()
    for i in range(1, V):  # random tree + a few chords
        edges.add((int(rng.integers(0, i)), i))
    while len(edges) < V + 3:
        a, b = sorted(int(x) for x in rng.integers(0, V, 2))
        if a != b:
            edges.add((a, b))
    edges = sorted(edges)
    g, (rp, col, _) = graph(V, [(a, b, 1.0) for a, b in edges])
    rows = [col[rp[v]:rp[v + 1]].tolist() for v in range(V)]
    s, t, L = 0, V - 1, 12
    p = exact_hit_probability(rows, s, t, L)
    N = 20000
    hits = 0
    for j in range(N):
        _, f, _ = g.trace(2, s, t, 1.0, 0.0, j, seed=11, eid=5, max_steps=L)
        hits += f != -1
    sigma = math.sqrt(p * (1 - p) / N) + 1e-9
    assert abs(hits / N - p) < 5 * sigma + 1e-6, (hits / N, p)


def test_winner_path_is_its_trajectory():
    rng = np.random.default_rng(5)
    V = 30
    edges = {(i, (i + 1) % V) for i in range(V)}
    while len(edges) < 60:
        a, b = sorted(int(x) for x in rng.integers(0, V, 2))
        if a != b:
            edges.add((a, b))
    g, _ = graph(V, [(a, b, float(rng.uniform(0.5, 2))) for a, b in sorted(edges)])
    for t in range(1, 10):
        f, c, win, _, path = g.walk_dec(0, t, 128, seed=2, eid=t)
        if f > 0:
            traj, ff, R = g.trace(2, 0, t, 0.0, 0.0, win, seed=2, eid=t)
            assert ff == f and traj[: f + 1].tolist() == path[: f + 1].tolist()
            assert np.float32(c) == np.float32(np.float32(1) / np.float32(R))


def test_heuristic_paths_follow_edges():
    rng = np.random.default_rng(8)
    V = 40
    edges = {(i, i + 1) for i in range(V - 1)}
    while len(edges) < 90:
        a, b = sorted(int(x) for x in rng.integers(0, V, 2))
        if a != b:
            edges.add((a, b))
    g, (rp, col, _) = graph(V, [(a, b, 1.0) for a, b in sorted(edges)])
    rows = [col[rp[v]:rp[v + 1]].tolist() for v in range(V)]
    for u in range(V):
        p3 = g.heuristic(u, 16, 1, u * 2 + 1)
        for a, b, c in p3:
            assert a == u and b in rows[u] and c in rows[b]
            # drawing u again redirects to n1's last slot, which may itself be u
            assert c != u or rows[b][-1] == u


@pytest.mark.parametrize("m", [0, 1, 2, 5, 9, 10, 11, 12, 19, 2000, 2005])
def test_snapshot_batches(m):
    I = (m + 9) // 10
    expect = [0] + [min(m, k * I) for k in range(1, (m + I - 1) // I + 1)] if m else [0]
    assert po.batches(m, 10).tolist() == expect
