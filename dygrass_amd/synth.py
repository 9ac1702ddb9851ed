"""Seeded synthetic workloads (libdygrass_synth.so) and the named configurations.

A workload is a dense graph G, a sparsifier H (maximum-weight spanning tree
of G plus a sample of off-tree edges) and an update stream.  Everything is
deterministic in (configuration, seed).  `scale` shrinks a configuration for
tests; 1.0 is the full size.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import ptr


@dataclass
class Edges:
    V: int
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    xy: np.ndarray | None = None

    def __len__(self) -> int:
        return len(self.u)


def _take(h) -> Edges:
    s = L.synth_lib()
    try:
        n = s.dyg_synth_ne(h)
        V = s.dyg_synth_nv(h)
        u = np.ctypeslib.as_array(s.dyg_synth_u(h), (n,)).copy() if n else np.zeros(0, np.uint32)
        v = np.ctypeslib.as_array(s.dyg_synth_v(h), (n,)).copy() if n else np.zeros(0, np.uint32)
        w = np.ctypeslib.as_array(s.dyg_synth_w(h), (n,)).copy() if n else np.zeros(0, np.float32)
        xyp = s.dyg_synth_xy(h)
        xy = np.ctypeslib.as_array(xyp, (2 * V,)).copy() if xyp else None
    finally:
        s.dyg_synth_free(h)
    return Edges(V, u, v, w, xy)


def grid2d(nx: int, ny: int, wlo: float, whi: float, seed: int) -> Edges:
    return _take(L.synth_lib().dyg_synth_grid2d(nx, ny, wlo, whi, seed))


def rgg_knn(n: int, k: int, seed: int) -> Edges:
    return _take(L.synth_lib().dyg_synth_rgg_knn(n, k, seed))


def trimesh(nx: int, ny: int, wlo: float, whi: float, seed: int) -> Edges:
    return _take(L.synth_lib().dyg_synth_trimesh(nx, ny, wlo, whi, seed))


def grid3d_shortcuts(n: int, shortcuts: int, wlo: float, whi: float, seed: int) -> Edges:
    return _take(L.synth_lib().dyg_synth_grid3d_sc(n, shortcuts, wlo, whi, seed))


def sparsifier(G: Edges, extra: int, seed: int) -> Edges:
    h = L.synth_lib().dyg_synth_sparsifier(G.V, ptr(G.u, L.u32p), ptr(G.v, L.u32p),
                                            ptr(G.w, L.f32p), len(G), extra, seed)
    H = _take(h)
    H.xy = G.xy
    return H


def insertions(G: Edges, count: int, dlo: int, dhi: int, wlo: float, whi: float, seed: int,
               geometric: bool = False) -> Edges:
    xy = G.xy if geometric else None
    h = L.synth_lib().dyg_synth_insertions(G.V, ptr(G.u, L.u32p), ptr(G.v, L.u32p), len(G),
                                            ptr(xy, L.f32p), count, dlo, dhi, wlo, whi, seed)
    return _take(h)


def insertions_walk(G: Edges, count: int, dlo: int, dhi: int, wlo: float, whi: float,
                    seed: int) -> Edges:
    """New edges to where a non-backtracking walk of dlo..dhi hops stops."""
    h = L.synth_lib().dyg_synth_insertions_walk(G.V, ptr(G.u, L.u32p), ptr(G.v, L.u32p), len(G),
                                                 count, dlo, dhi, wlo, whi, seed)
    return _take(h)


def deletions(G: Edges, count: int, seed: int) -> Edges:
    h = L.synth_lib().dyg_synth_deletions(G.V, ptr(G.u, L.u32p), ptr(G.v, L.u32p),
                                           ptr(G.w, L.f32p), len(G), count, seed)
    return _take(h)


def write_mtx(path: str, u, v, w=None, base: int = 1) -> None:
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    wp = None if w is None else np.ascontiguousarray(w, np.float32)
    rc = L.synth_lib().dyg_synth_write_mtx(path.encode(), ptr(u, L.u32p), ptr(v, L.u32p),
                                           ptr(wp, L.f32p), len(u), base)
    if rc != 0:
        raise OSError(f"cannot write {path}")


def _seed(config: int, purpose: int, seed: int) -> int:
    """SplitMix64 of (seed, config, purpose)."""
    z = (seed * 0x9E3779B97F4A7C15 + config * 0xBF58476D1CE4E5B9 + purpose * 0x94D049BB133111EB
         + 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


@dataclass
class Workload:
    name: str
    G: Edges
    H: Edges
    ins: Edges | None = None
    dels: Edges | None = None
    threshold: float = 0.0
    walkers: int = 512
    batches: int = 10
    meta: dict = field(default_factory=dict)


def config_grid_insert(scale: float = 1.0, seed: int = 0) -> Workload:
    """128x128 grid, weights log-uniform [0.1,10], H = tree + 5% V, 2000
    insertions at hop distance 2..6, threshold 10, 512 walkers."""
    side = max(8, int(round(128 * math.sqrt(scale))))
    G = grid2d(side, side, 0.1, 10.0, _seed(1, 0, seed))
    H = sparsifier(G, int(0.05 * G.V), _seed(1, 1, seed))
    ins = insertions(G, max(1, int(2000 * scale)), 2, 6, 0.1, 10.0, _seed(1, 2, seed))
    return Workload(f"grid{side}x{side}-ins", G, H, ins=ins, threshold=10.0, walkers=512)


def config_rgg_insert(scale: float = 1.0, seed: int = 0) -> Workload:
    """1M-point 8-NN random geometric graph (w = 1/dist), H = tree + 2% V,
    10 batches of 20k insertions within 5 hops, threshold 100, 256 walkers."""
    n = max(1000, int(1_000_000 * scale))
    G = rgg_knn(n, 8, _seed(2, 0, seed))
    H = sparsifier(G, int(0.02 * n), _seed(2, 1, seed))
    ins = insertions(G, max(10, int(200_000 * scale)), 2, 5, 0.0, 0.0, _seed(2, 2, seed),
                     geometric=True)
    return Workload(f"rgg{n}-knn8-ins", G, H, ins=ins, threshold=100.0, walkers=256)


def config_mesh_delete(scale: float = 1.0, seed: int = 0) -> Workload:
    """2M-node triangulated mesh, weights log-uniform [0.5,2], H = tree + 3% V,
    10 batches of 10k uniform deletions, 512 walkers, 100 steps."""
    side = max(16, int(round(1414 * math.sqrt(scale))))
    G = trimesh(side, side, 0.5, 2.0, _seed(3, 0, seed))
    H = sparsifier(G, int(0.03 * G.V), _seed(3, 1, seed))
    dels = deletions(G, max(10, int(100_000 * scale)), _seed(3, 2, seed))
    return Workload(f"mesh{side}x{side}-del", G, H, dels=dels, walkers=512)


def config_rgg12(scale: float = 1.0, seed: int = 0) -> Workload:
    """4M-point 12-NN random geometric graph, H = tree + 2% V."""
    n = max(1000, int(4_000_000 * scale))
    G = rgg_knn(n, 12, _seed(4, 0, seed))
    H = sparsifier(G, int(0.02 * n), _seed(4, 1, seed))
    return Workload(f"rgg{n}-knn12", G, H, walkers=256, threshold=100.0)


def config_grid3d(scale: float = 1.0, seed: int = 0) -> Workload:
    """200^3 6-neighbour grid plus 16M shortcuts, weights log-uniform [0.1,10],
    H = tree + about 2M edges."""
    n = max(8, int(round(200 * scale ** (1 / 3))))
    V = n ** 3
    G = grid3d_shortcuts(n, int(16_000_000 * V / 8_000_000), 0.1, 10.0, _seed(5, 0, seed))
    H = sparsifier(G, int(0.25 * V), _seed(5, 1, seed))
    return Workload(f"grid3d{n}-sc", G, H, walkers=256, threshold=100.0)


@dataclass
class StreamWorkload:
    """G, H and a list of batches (op, u, v, w) with op "ins" or "del"."""
    name: str
    G: Edges
    H: Edges
    batches: list
    threshold: float = 100.0
    w_ins: int = 256
    w_del: int = 512

    @property
    def updates(self) -> int:
        return sum(len(b[1]) for b in self.batches)


def stream_batches(G: Edges, nbatch: int, max_size: int, frac_ins: float, seed: int,
                   geometric: bool, wlo: float = 0.0, whi: float = 0.0) -> list:
    """`nbatch` batches with sizes floor(exp(U[0, ln max_size])), each an
    insertion batch with probability frac_ins, else a deletion batch.
    Insertions: new edges within 2..5 hops of G (weights as the graph's);
    deletions: distinct edges of G drawn uniformly, consumed in order."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(0.0, math.log(max_size), nbatch))).astype(np.int64)
    sizes = np.maximum(sizes, 1)
    ops = np.where(rng.uniform(size=nbatch) < frac_ins, "ins", "del")
    n_ins = int(sizes[ops == "ins"].sum())
    n_del = int(sizes[ops == "del"].sum())
    ins = insertions(G, max(n_ins, 1), 2, 5, wlo, whi, seed ^ 0x51, geometric=geometric)
    dels = deletions(G, max(n_del, 1), seed ^ 0xD7)
    out, ki, kd = [], 0, 0
    for op, k in zip(ops, sizes):
        k = int(k)
        if op == "ins":
            k = min(k, len(ins) - ki)
            out.append(("ins", ins.u[ki:ki + k], ins.v[ki:ki + k], ins.w[ki:ki + k]))
            ki += k
        else:
            k = min(k, len(dels) - kd)
            out.append(("del", dels.u[kd:kd + k], dels.v[kd:kd + k], None))
            kd += k
    return out


def config_batched(scale: float = 1.0, seed: int = 0) -> StreamWorkload:
    """4M-point 12-NN random geometric graph (w = 1/dist), H = tree + 2% V;
    64 batches of log-uniform size in [1, 50000], 70% insertions (within 5
    hops, w = 1/dist) and 30% deletions (uniform over G); threshold 100,
    256 insertion walkers, 512 deletion walkers."""
    wl = config_rgg12(scale, seed)
    mx = max(2, int(50_000 * min(1.0, scale * 4)))
    b = stream_batches(wl.G, 64, mx, 0.7, _seed(4, 2, seed), geometric=True)
    return StreamWorkload(f"rgg{wl.G.V}-knn12-stream", wl.G, wl.H, b, 100.0, 256, 512)


@dataclass
class ServedWorkload:
    """G, H and a request trace: arrival times (s) and (op, u, v, w)."""
    name: str
    G: Edges
    H: Edges
    arrivals: np.ndarray
    requests: list
    threshold: float = 100.0
    w_ins: int = 256
    w_del: int = 512

    @property
    def updates(self) -> int:
        return sum(len(r[1]) for r in self.requests)


def config_served(scale: float = 1.0, seed: int = 0, count: int = 1500,
                  rate: float = 1000.0) -> ServedWorkload:
    """8M-vertex 3D grid plus shortcuts (w log-uniform [0.1,10]), H = tree +
    ~2M edges; `count` requests with Poisson arrivals at `rate` per second,
    sizes log-uniform in [1, 10000], 60% insertions (endpoints 2..5 walk hops
    apart, weights log-uniform [0.1,10]) and 40% deletions (uniform over G)."""
    wl = config_grid3d(scale, seed)
    rng = np.random.default_rng(_seed(5, 2, seed))
    mx = max(2, int(10_000 * min(1.0, scale * 8)))
    sizes = np.maximum(np.floor(np.exp(rng.uniform(0.0, math.log(mx), count))), 1).astype(int)
    ops = np.where(rng.uniform(size=count) < 0.6, "ins", "del")
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    ins = insertions_walk(wl.G, max(1, int(sizes[ops == "ins"].sum())), 2, 5, 0.1, 10.0,
                          _seed(5, 3, seed))
    dels = deletions(wl.G, max(1, int(sizes[ops == "del"].sum())), _seed(5, 4, seed))
    reqs, ki, kd = [], 0, 0
    for op, k in zip(ops, sizes):
        k = int(k)
        if op == "ins":
            k = min(k, len(ins) - ki)
            reqs.append(("ins", ins.u[ki:ki + k], ins.v[ki:ki + k], ins.w[ki:ki + k]))
            ki += k
        else:
            k = min(k, len(dels) - kd)
            reqs.append(("del", dels.u[kd:kd + k], dels.v[kd:kd + k], None))
            kd += k
    keep = [i for i, r in enumerate(reqs) if len(r[1]) > 0]
    return ServedWorkload(f"{wl.name}-served", wl.G, wl.H, arrivals[keep],
                          [reqs[i] for i in keep], 100.0, 256, 512)


CONFIGS = {
    "grid-insert": config_grid_insert,
    "rgg-insert": config_rgg_insert,
    "mesh-delete": config_mesh_delete,
    "rgg12": config_rgg12,
    "grid3d": config_grid3d,
}


def write_dataset(root: str, wl: Workload) -> str:
    """Write the reference's file set for `wl` under root/dataset/<name>/ and
    return that directory: adj_sparse.mtx (0-based H), ext.mtx (1-based
    insertions), dense.mtx (1-based G); for deletions updated_adj_sparse.mtx,
    updated_dense.mtx and del.mtx (1-based)."""
    import os

    d = os.path.join(root, "dataset", wl.name)
    os.makedirs(d, exist_ok=True)
    write_mtx(os.path.join(d, "dense.mtx"), wl.G.u, wl.G.v, wl.G.w, base=1)
    if wl.ins is not None:
        write_mtx(os.path.join(d, "adj_sparse.mtx"), wl.H.u, wl.H.v, wl.H.w, base=0)
        write_mtx(os.path.join(d, "ext.mtx"), wl.ins.u, wl.ins.v, wl.ins.w, base=1)
    if wl.dels is not None:
        write_mtx(os.path.join(d, "updated_adj_sparse.mtx"), wl.H.u, wl.H.v, wl.H.w, base=1)
        write_mtx(os.path.join(d, "updated_dense.mtx"), wl.G.u, wl.G.v, wl.G.w, base=1)
        write_mtx(os.path.join(d, "del.mtx"), wl.dels.u, wl.dels.v, None, base=1)
    return d
