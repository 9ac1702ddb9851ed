"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p

TAG_INSERT, TAG_DELETE, TAG_HEUR = 1, 2, 3


class RunStats(C.Structure):
    _fields_ = [("n_updates", C.c_uint64), ("n_added", C.c_uint64), ("n_batches", C.c_uint64),
                ("hops", C.c_uint64), ("sparse_before", C.c_uint64),
                ("sparse_after", C.c_uint64), ("added_edges", C.c_uint64),
                ("no_path", C.c_uint64), ("filtered", C.c_uint64),
                ("walk_seconds", C.c_double), ("total_seconds", C.c_double)]


class StreamStats(C.Structure):
    _fields_ = [("updates", C.c_uint64), ("batches", C.c_uint64), ("hops", C.c_uint64),
                ("walk_seconds", C.c_double), ("total_seconds", C.c_double)]


_SIGS = {
    "orc_stream_new": (vp, [C.c_uint32, u64p, u32p, f32p, u64p, u32p, f32p, C.c_float,
                            C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                            C.c_uint64, C.c_int]),
    "orc_stream_apply": (C.c_int, [vp, C.c_int, u32p, u32p, f32p, C.c_uint32, i32p]),
    "orc_stream_graph": (vp, [vp, C.c_int]),
    "orc_stream_get_stats": (None, [vp, C.POINTER(StreamStats)]),
    "orc_stream_free": (None, [vp]),
    "orc_philox4x32_10": (None, [u32p, u32p, u32p]),
    "orc_draw": (C.c_uint32, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
    "orc_tree_reduce": (C.c_uint32, [f32p, C.c_uint32]),
    "orc_graph_new": (vp, [C.c_uint32, u64p, u32p, f32p]),
    "orc_graph_free": (None, [vp]),
    "orc_graph_nv": (C.c_uint32, [vp]),
    "orc_graph_nslots": (C.c_uint64, [vp]),
    "orc_graph_export": (None, [vp, u64p, u32p, f32p]),
    "orc_graph_insert": (None, [vp, C.c_uint32, C.c_uint32, C.c_float]),
    "orc_walk_inc_edge": (None, [vp, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint32,
                                 C.c_uint32, C.c_uint64, C.c_uint64, i32p, f32p, u32p, u64p]),
    "orc_walk_dec_edge": (None, [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                 C.c_uint64, i32p, f32p, u32p, u64p, u32p]),
    "orc_walk_trace": (C.c_uint32, [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                    C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, u32p, i32p,
                                    f32p]),
    "orc_walk_heuristic": (None, [vp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, u32p]),
    "orc_walk_inc_batch": (None, [vp, C.c_uint32, u32p, u32p, f32p, u64p, C.c_float, C.c_uint32,
                                  C.c_uint32, C.c_uint64, C.c_int, i32p, f32p, u32p, u64p]),
    "orc_walk_dec_batch": (None, [vp, C.c_uint32, u32p, u32p, u64p, C.c_uint32, C.c_uint32,
                                  C.c_uint64, C.c_int, i32p, f32p, u32p, u64p, u32p]),
    "orc_batches": (C.c_uint32, [C.c_uint32, C.c_uint32, u32p]),
    "orc_format_float": (C.c_int, [C.c_float, C.c_char_p, C.c_int]),
    "orc_decremental_lines": (C.c_int, [C.c_uint32, u32p, u32p, f32p, C.c_uint64, u32p, u32p,
                                        C.c_uint64, u32p, u32p, C.c_uint32, C.c_uint64, C.c_int,
                                        C.POINTER(RunStats), u32p, u32p, C.c_uint64]),
    "orc_incremental_arrays": (C.c_int, [vp, C.c_uint32, u32p, u32p, f32p, C.c_uint32, C.c_float,
                                         C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, i32p,
                                         C.POINTER(RunStats)]),
    "orc_dense_open": (vp, [C.c_char_p]),
    "orc_dense_remove": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
    "orc_dense_graph": (vp, [vp]),
    "orc_dense_multiplier": (C.c_int64, [vp]),
    "orc_dense_free": (None, [vp]),
    "orc_incremental_run": (C.c_int, [C.c_char_p, C.c_float, C.c_uint32, C.c_uint64, C.c_int,
                                      C.c_char_p, C.POINTER(RunStats)]),
    "orc_decremental_run": (C.c_int, [C.c_char_p, C.c_uint64, C.c_int, C.c_char_p,
                                      C.POINTER(RunStats)]),
    "orc_read_graph0": (C.c_int64, [C.c_char_p, u64p, u64p, u32p, f32p]),
    "orc_read_targets": (C.c_int64, [C.c_char_p, C.c_int, C.c_uint32, u32p, f32p]),
}

_lib = None


def build() -> None:
    """Compile liboracle.so (and the reference-parser dumps when the
    reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a, ct):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ct)


def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c, u32p), _p(k, u32p), _p(out, u32p))
    return out


def draw(seed: int, tag: int, edge_id: int, walker: int, d: int) -> int:
    return lib().orc_draw(seed, tag, edge_id, walker, d)


def format_float(f: float) -> str:
    buf = C.create_string_buffer(64)
    n = lib().orc_format_float(f, buf, 64)
    return buf.value[:n].decode()


def tree_reduce(cond) -> int:
    c = np.ascontiguousarray(cond, np.float32)
    return lib().orc_tree_reduce(_p(c, f32p), len(c))


def batches(m: int, divisor: int = 10) -> np.ndarray:
    out = np.zeros(m + 2, np.uint32)
    n = lib().orc_batches(m, divisor, _p(out, u32p))
    return out[: n + 1]


class Graph:
    """Per-row slot lists (the oracle's graph)."""

    def __init__(self, V: int, row_ptr, col, w):
        rp = np.ascontiguousarray(row_ptr, np.uint64)
        cl = np.ascontiguousarray(col, np.uint32)
        ww = np.ascontiguousarray(w, np.float32)
        self.V = V
        self.h = lib().orc_graph_new(V, _p(rp, u64p), _p(cl, u32p), _p(ww, f32p))

    def insert(self, u: int, v: int, w: float) -> None:
        lib().orc_graph_insert(self.h, u, v, w)

    def export(self):
        n = lib().orc_graph_nslots(self.h)
        rp = np.zeros(self.V + 1, np.uint64)
        col = np.zeros(max(n, 1), np.uint32)
        w = np.zeros(max(n, 1), np.float32)
        lib().orc_graph_export(self.h, _p(rp, u64p), _p(col, u32p), _p(w, f32p))
        return rp, col[:n], w[:n]

    def walk_inc(self, s, t, w_e, thr, W, seed=0, eid=0, max_steps=100):
        f, c, win, h = C.c_int32(), C.c_float(), C.c_uint32(), C.c_uint64()
        lib().orc_walk_inc_edge(self.h, s, t, w_e, thr, W, max_steps, seed, eid, C.byref(f),
                                C.byref(c), C.byref(win), C.byref(h))
        return f.value, c.value, win.value, h.value

    def walk_dec(self, s, t, W, seed=0, eid=0, max_steps=100):
        f, c, win, h = C.c_int32(), C.c_float(), C.c_uint32(), C.c_uint64()
        path = np.zeros(max_steps, np.uint32)
        lib().orc_walk_dec_edge(self.h, s, t, W, max_steps, seed, eid, C.byref(f), C.byref(c),
                                C.byref(win), C.byref(h), _p(path, u32p))
        return f.value, c.value, win.value, h.value, path

    def trace(self, kind, s, t, w_e, thr, walker, seed=0, eid=0, max_steps=100):
        traj = np.zeros(max_steps + 1, np.uint32)
        f, R = C.c_int32(), C.c_float()
        n = lib().orc_walk_trace(self.h, kind, s, t, w_e, thr, max_steps, seed, eid, walker,
                                 _p(traj, u32p), C.byref(f), C.byref(R))
        return traj[:n], f.value, R.value

    def heuristic(self, u, nwalk, seed, key) -> np.ndarray:
        out = np.zeros(nwalk * 3, np.uint32)
        lib().orc_walk_heuristic(self.h, u, nwalk, seed, key, _p(out, u32p))
        return out.reshape(nwalk, 3)

    def walk_inc_batch(self, src, dst, w, thr, W, seed=0, edge_id=None, threads=1,
                       max_steps=100):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = np.ascontiguousarray(w, np.float32)
        n = len(src)
        eid = np.ascontiguousarray(np.arange(n) if edge_id is None else edge_id, np.uint64)
        f = np.zeros(max(n, 1), np.int32)
        c = np.zeros(max(n, 1), np.float32)
        win = np.zeros(max(n, 1), np.uint32)
        h = np.zeros(max(n, 1), np.uint64)
        lib().orc_walk_inc_batch(self.h, n, _p(src, u32p), _p(dst, u32p), _p(w, f32p),
                                 _p(eid, u64p), thr, W, max_steps, seed, threads, _p(f, i32p),
                                 _p(c, f32p), _p(win, u32p), _p(h, u64p))
        return f[:n], c[:n], win[:n], h[:n]

    def incremental(self, src, dst, w, thr, W, seed=0, threads=1, divisor=10, max_steps=100):
        """Whole snapshot-batched insertion run; modifies this graph."""
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = np.ascontiguousarray(w, np.float32)
        m = len(src)
        f = np.zeros(max(m, 1), np.int32)
        st = RunStats()
        lib().orc_incremental_arrays(self.h, m, _p(src, u32p), _p(dst, u32p), _p(w, f32p),
                                     divisor, thr, W, max_steps, seed, threads, _p(f, i32p),
                                     C.byref(st))
        return f[:m], st

    def walk_dec_batch(self, src, dst, W, seed=0, edge_id=None, threads=1, max_steps=100):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        n = len(src)
        eid = np.ascontiguousarray(np.arange(n) if edge_id is None else edge_id, np.uint64)
        f = np.zeros(max(n, 1), np.int32)
        c = np.zeros(max(n, 1), np.float32)
        win = np.zeros(max(n, 1), np.uint32)
        h = np.zeros(max(n, 1), np.uint64)
        path = np.zeros(max(n, 1) * max_steps, np.uint32)
        lib().orc_walk_dec_batch(self.h, n, _p(src, u32p), _p(dst, u32p), _p(eid, u64p), W,
                                 max_steps, seed, threads, _p(f, i32p), _p(c, f32p),
                                 _p(win, u32p), _p(h, u64p), _p(path, u32p))
        return f[:n], c[:n], win[:n], h[:n], path.reshape(-1, max_steps)[:n]

    def __del__(self):
        try:
            if self.h:
                lib().orc_graph_free(self.h)
                self.h = None
        except Exception:
            pass


class Dense:
    """Dense graph parsed from a 1-based weighted .mtx with the reference's
    edge map; remove() restates the dense half of removeAndUpdateDelEdges."""

    def __init__(self, path: str):
        self.h = lib().orc_dense_open(path.encode())
        if not self.h:
            raise OSError(path)
        self.graph = Graph.__new__(Graph)
        self.graph.h = lib().orc_dense_graph(self.h)
        self.graph.V = lib().orc_graph_nv(self.graph.h)

    @property
    def multiplier(self) -> int:
        return lib().orc_dense_multiplier(self.h)

    def remove(self, u: int, v: int) -> int:
        return lib().orc_dense_remove(self.h, u, v)

    def __del__(self):
        try:
            if self.h:
                self.graph.h = None  # owned by the dense object
                lib().orc_dense_free(self.h)
                self.h = None
        except Exception:
            pass


def decremental_lines(V, gu, gv, gw, hu, hv, du, dv, seed=0, threads=1):
    """Whole decremental loop on arrays; returns (stats, added pairs [k,2])."""
    a = [np.ascontiguousarray(x, np.uint32) for x in (gu, gv, hu, hv, du, dv)]
    gw = np.ascontiguousarray(gw, np.float32)
    cap = 4 * len(a[4]) * 100 + 16
    au = np.zeros(cap, np.uint32)
    av = np.zeros(cap, np.uint32)
    st = RunStats()
    rc = lib().orc_decremental_lines(V, _p(a[0], u32p), _p(a[1], u32p), _p(gw, f32p), len(a[0]),
                                     _p(a[2], u32p), _p(a[3], u32p), len(a[2]), _p(a[4], u32p),
                                     _p(a[5], u32p), len(a[4]), seed, threads, C.byref(st),
                                     _p(au, u32p), _p(av, u32p), cap)
    if rc != 0:
        raise RuntimeError(f"oracle decremental run failed ({rc})")
    n = int(st.added_edges)
    return st, np.stack([au[:n], av[:n]], axis=1)


def incremental_run(dirpath: str, thr: float, W: int, seed: int = 0, threads: int = 1,
                    log_path: str | None = None) -> RunStats:
    st = RunStats()
    rc = lib().orc_incremental_run(dirpath.encode(), thr, W, seed, threads,
                                   log_path.encode() if log_path else None, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle incremental run failed ({rc})")
    return st


def decremental_run(dirpath: str, seed: int = 0, threads: int = 1,
                    log_path: str | None = None) -> RunStats:
    st = RunStats()
    rc = lib().orc_decremental_run(dirpath.encode(), seed, threads,
                                   log_path.encode() if log_path else None, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle decremental run failed ({rc})")
    return st


def read_graph0(path: str):
    lines = C.c_uint64()
    V = lib().orc_read_graph0(path.encode(), C.byref(lines), None, None, None)
    if V < 0:
        raise OSError(path)
    begin = np.zeros(V + 1, np.uint64)
    adj = np.zeros(max(2 * lines.value, 1), np.uint32)
    w = np.zeros(max(2 * lines.value, 1), np.float32)
    lib().orc_read_graph0(path.encode(), None, _p(begin, u64p), _p(adj, u32p), _p(w, f32p))
    n = 2 * lines.value
    return V, begin, adj[:n], w[:n]


def read_targets(path: str, weighted: bool = True, divisor: int = 10):
    m = lib().orc_read_targets(path.encode(), int(weighted), divisor, None, None)
    if m < 0:
        raise OSError(path)
    nodes = np.zeros(max(2 * m, 1), np.uint32)
    w = np.zeros(max(m, 1), np.float32)
    lib().orc_read_targets(path.encode(), int(weighted), divisor, _p(nodes, u32p), _p(w, f32p))
    return nodes[: 2 * m], w[:m]


def _csr_of(V, u, v, w):
    """Row lists in slot order from an edge list: edge (a,b,w) appends b to
    a's row, then a to b's row (the order every graph here is built in)."""
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    w = np.ascontiguousarray(w, np.float32)
    m = len(u)
    src = np.empty(2 * m, np.uint32)
    dst = np.empty(2 * m, np.uint32)
    ww = np.empty(2 * m, np.float32)
    src[0::2], src[1::2] = u, v
    dst[0::2], dst[1::2] = v, u
    ww[0::2], ww[1::2] = w, w
    order = np.argsort(src, kind="stable")
    rp = np.zeros(V + 1, np.uint64)
    np.cumsum(np.bincount(src, minlength=V), out=rp[1:])
    return rp, np.ascontiguousarray(dst[order]), np.ascontiguousarray(ww[order])


class StreamCpp:
    """The stream oracle in C++ (orc_stream_*): same semantics and outcome
    codes as oracle/stream.py, fast enough for full-size workloads and used
    as the CPU baseline of the stream forms (walks OpenMP-parallel over each
    batch's edges, graph edits sequential).  Insertion walks decide only.
    g_edges / h_edges: (u, v, w) edge lists, or csr=((rp, col, w), (rp, col, w))."""

    def __init__(self, V, g_edges=None, h_edges=None, thr=100.0, w_ins=256, w_del=512, seed=0,
                 steps_ins=100, steps_del=100, heur_walkers=16, threads=1, csr=None):
        self.V = int(V)
        if csr is None:
            csr = (_csr_of(V, *g_edges), _csr_of(V, *h_edges))
        (grp, gcol, gw), (hrp, hcol, hw) = csr
        a = [np.ascontiguousarray(x, t) for x, t in ((grp, np.uint64), (gcol, np.uint32),
                                                      (gw, np.float32), (hrp, np.uint64),
                                                      (hcol, np.uint32), (hw, np.float32))]
        self.h = lib().orc_stream_new(self.V, _p(a[0], u64p), _p(a[1], u32p), _p(a[2], f32p),
                                      _p(a[3], u64p), _p(a[4], u32p), _p(a[5], f32p), thr, w_ins,
                                      w_del, steps_ins, steps_del, heur_walkers, seed, threads)

    def apply(self, op: str, u, v, w=None) -> np.ndarray:
        u = np.ascontiguousarray(u, np.uint32)
        v = np.ascontiguousarray(v, np.uint32)
        n = len(u)
        ww = np.ascontiguousarray(w, np.float32) if op == "ins" else None
        out = np.zeros(max(n, 1), np.int32)
        rc = lib().orc_stream_apply(self.h, 0 if op == "ins" else 1, _p(u, u32p), _p(v, u32p),
                                    _p(ww, f32p), n, _p(out, i32p))
        if rc != 0:
            raise ValueError("oracle stream: vertex id out of range")
        return out[:n]

    def csr(self, which: int):
        """(row_ptr, col, w) of G (0) or H (1) in slot order."""
        g = lib().orc_stream_graph(self.h, which)
        V = lib().orc_graph_nv(g)
        n = lib().orc_graph_nslots(g)
        rp = np.zeros(V + 1, np.uint64)
        col = np.zeros(max(n, 1), np.uint32)
        w = np.zeros(max(n, 1), np.float32)
        lib().orc_graph_export(g, _p(rp, u64p), _p(col, u32p), _p(w, f32p))
        return rp, col[:n], w[:n]

    def stats(self) -> dict:
        st = StreamStats()
        lib().orc_stream_get_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in StreamStats._fields_}

    def close(self) -> None:
        if self.h:
            lib().orc_stream_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
