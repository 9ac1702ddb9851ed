"""Sequential restatement of the mixed insertion/deletion update stream.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU-baseline leg,
as the checker.  The product never imports it.

Semantics (dynamic_update/gpu_graph.cuh:361-439 preprocessStreamEdges,
:451-465 dualGraphEdgeweightsUpdate, :620-700 updateSparsiferFromResult),
with the walks, the block reduction and the heuristic of the two plain
updaters (random_walk_incremental/main.cu:70-211,
random_walk_decremental/main.cu:77-290, gpu_graph.cuh:509-615) and the
reference defects resolved as in those updaters (heuristic hop two from
n1's own row; the weight update adds to the weight, not to the neighbour
id, dynamic_update/gpu_graph.cuh:461-462).

State: G (dense) and H (sparsifier, H a subgraph of G), both simple graphs
kept as per-vertex neighbour lists in slot order.  A batch is one operation
kind over a list of edges, processed with snapshot semantics:

insertion batch, in batch order:
  * (a,b) already in G: its weight grows by w in G, and in H if there;
  * otherwise (a,b) is appended to G (row a gets b, row b gets a) and kept;
  the kept edges walk in H as it stood at batch start (insertion walk,
  source = the edge's first vertex as given); each kept edge that finds no
  path is appended to H, in batch order, with its own weight.

deletion batch, in batch order:
  * (a,b) not in G: skipped;
  * otherwise removed from G (swap-with-last in both rows) and, when in H,
    from H too; those are kept;
  the kept edges walk in G after all of the batch's deletions (deletion walk
  from min(a,b) to max(a,b)); then, in batch order, every hop (x,y) of a
  found walk's path that H lacks is appended to H with G's weight; then, for
  every kept edge without a path, both endpoints (min first) run the 16-walker
  heuristic in G and each (u,n1), (n1,n2) that H lacks is appended to H with
  G's weight.

A walker standing on a vertex of degree 0 stops (isolated source or leaf).
Edge ids for the draw streams: the update's position in the whole stream.
Self-loops are skipped with outcome INVALID.
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as po

# per-update outcome codes (the library reports the same values)
WEIGHT = 1      # insertion of an existing edge: weight update
FOUND = 2       # insertion: a path was found, H unchanged
ADDED = 3       # insertion: no path, appended to H
SKIP = 4        # deletion of an edge not in G
G_ONLY = 5      # deletion from G only
RECOVERED = 6   # deletion from G and H, a path was found and recovered into H
HEURISTIC = 7   # deletion from G and H, no path: heuristic edges into H
INVALID = 8     # self-loop


def f32(x) -> float:
    return float(np.float32(x))


class Rows:
    """A simple graph as neighbour lists in slot order: rows[v] = [[nbr, w], ...]."""

    def __init__(self, V: int, u=(), v=(), w=()):
        self.V = V
        self.rows = [[] for _ in range(V)]
        for a, b, x in zip(u, v, w):
            self.append(int(a), int(b), f32(x))

    def pos(self, a: int, b: int) -> int:
        for i, (x, _) in enumerate(self.rows[a]):
            if x == b:
                return i
        return -1

    def has(self, a: int, b: int) -> bool:
        return self.pos(a, b) >= 0

    def weight(self, a: int, b: int) -> float:
        return self.rows[a][self.pos(a, b)][1]

    def append(self, a: int, b: int, w: float) -> None:
        self.rows[a].append([b, w])
        self.rows[b].append([a, w])

    def add_weight(self, a: int, b: int, dw: float) -> None:
        for x, y in ((a, b), (b, a)):
            s = self.rows[x][self.pos(x, y)]
            s[1] = f32(np.float32(s[1]) + np.float32(dw))

    def _remove(self, a: int, b: int) -> None:
        r = self.rows[a]
        p = self.pos(a, b)
        r[p] = r[-1]
        r.pop()

    def delete(self, a: int, b: int) -> None:
        self._remove(a, b)
        self._remove(b, a)

    def csr(self):
        deg = np.array([len(r) for r in self.rows], np.uint64)
        rp = np.zeros(self.V + 1, np.uint64)
        rp[1:] = np.cumsum(deg)
        col = np.array([x for r in self.rows for x, _ in r], np.uint32)
        w = np.array([x for r in self.rows for _, x in r], np.float32)
        return rp, col, w

    def graph(self) -> "po.Graph":
        rp, col, w = self.csr()
        return po.Graph(self.V, rp, col, w)


class Stream:
    """G, H and the walk parameters of one update stream."""

    def __init__(self, V, g_edges, h_edges, thr=100.0, w_ins=256, w_del=512, seed=0,
                 max_steps=100, heur_walkers=16, threads=8, steps_ins=None, steps_del=None):
        self.G = Rows(V, *g_edges)
        self.H = Rows(V, *h_edges)
        self.thr, self.w_ins, self.w_del = float(thr), w_ins, w_del
        self.seed, self.heur = seed, heur_walkers
        # walk bounds of the insertion and deletion walks (the reference: 100 each)
        self.steps_ins = max_steps if steps_ins is None else steps_ins
        self.steps_del = max_steps if steps_del is None else steps_del
        self.threads = threads
        self.base = 0  # stream position of the next update

    def insert_batch(self, u, v, w) -> np.ndarray:
        n = len(u)
        out = np.zeros(n, np.int32)
        keep = []
        for i in range(n):
            s, t, x = int(u[i]), int(v[i]), f32(w[i])
            a, b = min(s, t), max(s, t)
            if a == b:
                out[i] = INVALID
            elif self.G.has(a, b):
                self.G.add_weight(a, b, x)
                if self.H.has(a, b):
                    self.H.add_weight(a, b, x)
                out[i] = WEIGHT
            else:
                self.G.append(a, b, x)
                keep.append(i)
        if keep:
            k = np.array(keep)
            hg = self.H.graph()
            f, _, _, _ = hg.walk_inc_batch(np.asarray(u)[k], np.asarray(v)[k],
                                           np.asarray(w, np.float32)[k], self.thr, self.w_ins,
                                           seed=self.seed, edge_id=(self.base + k).astype(np.uint64),
                                           threads=self.threads, max_steps=self.steps_ins)
            for j, i in enumerate(keep):
                if f[j] == -1:
                    a, b = min(int(u[i]), int(v[i])), max(int(u[i]), int(v[i]))
                    self.H.append(a, b, f32(w[i]))
                    out[i] = ADDED
                else:
                    out[i] = FOUND
        self.base += n
        return out

    def _recover(self, x: int, y: int) -> None:
        if x == y:
            return
        a, b = min(x, y), max(x, y)
        if not self.H.has(a, b):
            self.H.append(a, b, self.G.weight(a, b))

    def delete_batch(self, u, v) -> np.ndarray:
        n = len(u)
        out = np.zeros(n, np.int32)
        keep = []
        for i in range(n):
            a, b = min(int(u[i]), int(v[i])), max(int(u[i]), int(v[i]))
            if a == b:
                out[i] = INVALID
            elif not self.G.has(a, b):
                out[i] = SKIP
            else:
                self.G.delete(a, b)
                if self.H.has(a, b):
                    self.H.delete(a, b)
                    keep.append(i)
                else:
                    out[i] = G_ONLY
        if keep:
            k = np.array(keep)
            src = np.minimum(np.asarray(u)[k], np.asarray(v)[k]).astype(np.uint32)
            dst = np.maximum(np.asarray(u)[k], np.asarray(v)[k]).astype(np.uint32)
            gg = self.G.graph()
            f, _, _, _, path = gg.walk_dec_batch(src, dst, self.w_del, seed=self.seed,
                                                 edge_id=(self.base + k).astype(np.uint64),
                                                 threads=self.threads,
                                                 max_steps=self.steps_del)
            for j, i in enumerate(keep):
                if f[j] > 0:
                    for q in range(1, f[j] + 1):
                        self._recover(int(path[j, q - 1]), int(path[j, q]))
                    out[i] = RECOVERED
            for j, i in enumerate(keep):
                if f[j] <= 0:
                    eid = self.base + int(i)
                    for side, e in ((0, int(src[j])), (1, int(dst[j]))):
                        p3 = gg.heuristic(e, self.heur, self.seed, (eid << 1) | side)
                        for uu, n1, n2 in p3:
                            self._recover(int(uu), int(n1))
                            self._recover(int(n1), int(n2))
                    out[i] = HEURISTIC
        self.base += n
        return out

    def apply(self, op: str, u, v, w=None) -> np.ndarray:
        if op == "ins":
            return self.insert_batch(u, v, w)
        return self.delete_batch(u, v)
