"""Python host interface over the C ABI.

Mirrors the reference's units one to one (citations on each method):
``DeviceGraph`` is the device graph store (gpu_graph / GPU_Dual_Graph's dense
half), ``walk_incremental`` / ``walk_decremental`` / ``walk_heuristic`` are the
three kernels with their result copies, ``incremental_update`` is the
incremental CLI's batch loop and ``DualGraph`` the decremental dual graph.
All compute runs in libdygrass.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check, ptr


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


@dataclass
class WalkResult:
    found_at: np.ndarray  # int32, winner's hop index or -1
    cond: np.ndarray      # float32, winner's 1/R or -1
    winner: np.ndarray    # uint32, winning walker
    hops: np.ndarray      # uint32, neighbour draws of each edge
    path: np.ndarray | None = None  # uint32 [n, max_steps] (deletions)
    stats: dict | None = None


def params(walkers: int, max_steps: int = 100, threshold: float = 0.0, seed: int = 0,
           prune: bool = True, per_edge: bool = False, host_edits: bool = False,
           block_edges: int = 0, decision_only: bool = False) -> L.WalkParams:
    """Walk parameters.  per_edge: one workgroup per edge (the reference's
    launch shape) instead of the work-queue kernel; host_edits: plan each
    batch's appends on the host (dyg_incremental_update only); block_edges:
    edges per workgroup of the work-queue kernel (0 = automatic);
    decision_only: insertion walks stop once any walker found the target
    (found / not found exact, found_at and cond from some finder)."""
    flags = ((0 if prune else L.DYG_WALK_NO_PRUNE) | (L.DYG_WALK_PER_EDGE if per_edge else 0)
             | (L.DYG_RUN_HOST_EDITS if host_edits else 0)
             | (L.DYG_WALK_DECISION_ONLY if decision_only else 0))
    return L.WalkParams(walkers, max_steps, threshold, flags, seed, block_edges, 0)


class Context:
    """One HIP device + stream (dyg_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = C.c_void_p()
        check(self._lib.dyg_ctx_create(device, C.byref(h)))
        self.handle = h

    def sync(self) -> None:
        check(self._lib.dyg_ctx_sync(self.handle), self.handle)

    @property
    def stream(self) -> int:
        return self._lib.dyg_ctx_stream(self.handle)

    def close(self) -> None:
        if self.handle:
            self._lib.dyg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device buffer holding one numpy array's worth of data (dyg_dev_alloc)."""

    def __init__(self, ctx: Context, n: int, dtype):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        h = C.c_void_p()
        check(ctx._lib.dyg_dev_alloc(ctx.handle, self.n * self.dtype.itemsize, C.byref(h)),
              ctx.handle)
        self.handle = h

    @classmethod
    def from_numpy(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        check(ctx._lib.dyg_copy_to_device(ctx.handle, d.handle, a.ctypes.data_as(C.c_void_p),
                                          a.nbytes), ctx.handle)
        return d

    @property
    def ptr(self) -> int:
        return self.handle.value

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        check(self.ctx._lib.dyg_copy_to_host(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                             self.handle, out.nbytes), self.ctx.handle)
        return out

    def close(self) -> None:
        if self.handle:
            self.ctx._lib.dyg_dev_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def csr_from_edges(V: int, u, v, w) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """CSR in the reference's slot order: line (a,b,w) appends b to a's row and
    then a to b's row (random_walk_incremental/functions.cpp:168-232)."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    w = np.asarray(w, dtype=np.float32)
    m = len(u)
    src = np.empty(2 * m, dtype=np.int64)
    dst = np.empty(2 * m, dtype=np.int64)
    ww = np.empty(2 * m, dtype=np.float32)
    src[0::2], src[1::2] = u, v
    dst[0::2], dst[1::2] = v, u
    ww[0::2], ww[1::2] = w, w
    order = np.argsort(src, kind="stable")
    row_ptr = np.zeros(V + 1, dtype=np.uint64)
    np.add.at(row_ptr, src + 1, 1)
    row_ptr = np.cumsum(row_ptr, dtype=np.uint64)
    return row_ptr, dst[order].astype(np.uint32), ww[order]


class DeviceGraph:
    """Device graph store with per-row slack (dyg_graph)."""

    def __init__(self, ctx: Context, V: int, row_ptr, col, w, slack: int = 2, keyed: bool = False):
        self.ctx = ctx
        self._lib = ctx._lib
        self.V = int(V)
        rp, cl, ww = _u64(row_ptr), _u32(col), _f32(w)
        h = C.c_void_p()
        check(self._lib.dyg_graph_upload(ctx.handle, self.V, ptr(rp, L.u64p), ptr(cl, L.u32p),
                                         ptr(ww, L.f32p), slack, int(keyed), C.byref(h)),
              ctx.handle)
        self.handle = h

    @classmethod
    def from_edges(cls, ctx: Context, V: int, u, v, w, slack: int = 2, keyed: bool = False):
        rp, cl, ww = csr_from_edges(V, u, v, w)
        return cls(ctx, V, rp, cl, ww, slack=slack, keyed=keyed)

    def info(self) -> tuple[int, int, int]:
        V, n, cap = C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(self._lib.dyg_graph_info(self.handle, C.byref(V), C.byref(n), C.byref(cap)))
        return V.value, n.value, cap.value

    def download(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        _, n, _ = self.info()
        rp = np.zeros(self.V + 1, dtype=np.uint64)
        col = np.zeros(max(n, 1), dtype=np.uint32)
        w = np.zeros(max(n, 1), dtype=np.float32)
        check(self._lib.dyg_graph_download(self.handle, ptr(rp, L.u64p), ptr(col, L.u32p),
                                           ptr(w, L.f32p)), self.ctx.handle)
        return rp, col[:n], w[:n]

    def check(self) -> int:
        """Live slots whose copy of the neighbour's row descriptor is stale (0 = consistent)."""
        bad = C.c_uint64()
        check(self._lib.dyg_graph_check(self.handle, C.byref(bad)), self.ctx.handle)
        return bad.value

    def reserve(self, max_inserts: int, max_batch: int) -> None:
        """Pre-allocate arena and device-edit buffers (dyg_graph_reserve)."""
        check(self._lib.dyg_graph_reserve(self.handle, max_inserts, max_batch), self.ctx.handle)

    def insert(self, u, v, w) -> None:
        """random_walk_incremental/gpu_graph.cuh:319-325,357-384 (append in order)."""
        u, v, w = _u32(u), _u32(v), _f32(w)
        check(self._lib.dyg_graph_insert(self.handle, ptr(u, L.u32p), ptr(v, L.u32p),
                                         ptr(w, L.f32p), len(u)), self.ctx.handle)

    def delete(self, u, v) -> np.ndarray:
        """random_walk_decremental/gpu_graph.cuh:338-400 (swap-with-last)."""
        u, v = _u32(u), _u32(v)
        hit = np.zeros(max(len(u), 1), dtype=np.int8)
        check(self._lib.dyg_graph_delete(self.handle, ptr(u, L.u32p), ptr(v, L.u32p), len(u),
                                         ptr(hit, L.i8p)), self.ctx.handle)
        return hit[: len(u)]

    @classmethod
    def borrowed(cls, ctx: Context, handle, V: int, owner) -> "DeviceGraph":
        """A view of a graph owned by another object (kept alive by `owner`)."""
        g = cls.__new__(cls)
        g.ctx, g._lib, g.V, g.handle, g._owner = ctx, ctx._lib, int(V), C.c_void_p(handle), owner
        return g

    def close(self) -> None:
        if self.handle:
            if getattr(self, "_owner", None) is None:
                self._lib.dyg_graph_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def walk_incremental(H: DeviceGraph, src, dst, w, p: L.WalkParams, edge_id=None) -> WalkResult:
    """random_walk_incremental/main.cu:70-211 + get_result (gpu_graph.cuh:551-554)."""
    src, dst, w = _u32(src), _u32(dst), _f32(w)
    n = len(src)
    eid = None if edge_id is None else _u64(edge_id)
    f = np.zeros(max(n, 1), np.int32)
    c = np.zeros(max(n, 1), np.float32)
    win = np.zeros(max(n, 1), np.uint32)
    hops = np.zeros(max(n, 1), np.uint32)
    st = L.WalkStats()
    check(H._lib.dyg_walk_incremental(H.handle, ptr(src, L.u32p), ptr(dst, L.u32p),
                                      ptr(w, L.f32p), ptr(eid, L.u64p), n, C.byref(p),
                                      ptr(f, L.i32p), ptr(c, L.f32p), ptr(win, L.u32p),
                                      ptr(hops, L.u32p), C.byref(st)), H.ctx.handle)
    return WalkResult(f[:n], c[:n], win[:n], hops[:n], None,
                      dict(hops=st.hops, redirects=st.redirects, leaf_stops=st.leaf_stops,
                           kernel_ms=st.kernel_ms))


def walk_decremental(G: DeviceGraph, src, dst, p: L.WalkParams, edge_id=None) -> WalkResult:
    """random_walk_decremental/main.cu:77-220 + get_result (gpu_graph.cuh:831-841)."""
    src, dst = _u32(src), _u32(dst)
    n = len(src)
    eid = None if edge_id is None else _u64(edge_id)
    f = np.zeros(max(n, 1), np.int32)
    c = np.zeros(max(n, 1), np.float32)
    win = np.zeros(max(n, 1), np.uint32)
    hops = np.zeros(max(n, 1), np.uint32)
    path = np.zeros(max(n, 1) * p.max_steps, np.uint32)
    st = L.WalkStats()
    check(G._lib.dyg_walk_decremental(G.handle, ptr(src, L.u32p), ptr(dst, L.u32p),
                                      ptr(eid, L.u64p), n, C.byref(p), ptr(f, L.i32p),
                                      ptr(c, L.f32p), ptr(win, L.u32p), ptr(hops, L.u32p),
                                      ptr(path, L.u32p), C.byref(st)), G.ctx.handle)
    return WalkResult(f[:n], c[:n], win[:n], hops[:n], path.reshape(-1, p.max_steps)[:n],
                      dict(hops=st.hops, redirects=st.redirects, leaf_stops=st.leaf_stops,
                           kernel_ms=st.kernel_ms))


def walk_heuristic(G: DeviceGraph, endpoints, keys, walkers: int = 16, seed: int = 0) -> np.ndarray:
    """random_walk_decremental/main.cu:243-290 (hop two from n1's row)."""
    ep, k = _u32(endpoints), _u64(keys)
    n = len(ep)
    out = np.zeros(max(n, 1) * walkers * 3, np.uint32)
    check(G._lib.dyg_walk_heuristic(G.handle, ptr(ep, L.u32p), ptr(k, L.u64p), n, walkers, seed,
                                    ptr(out, L.u32p)), G.ctx.handle)
    return out[: n * walkers * 3].reshape(n, walkers, 3)


def incremental_update(H: DeviceGraph, src, dst, w, p: L.WalkParams, divisor: int = 10):
    """The incremental CLI's batch loop (random_walk_incremental/main.cu:297-340)."""
    src, dst, w = _u32(src), _u32(dst), _f32(w)
    m = len(src)
    f = np.zeros(max(m, 1), np.int32)
    st = L.RunStats()
    check(H._lib.dyg_incremental_update(H.handle, ptr(src, L.u32p), ptr(dst, L.u32p),
                                        ptr(w, L.f32p), m, divisor, C.byref(p), ptr(f, L.i32p),
                                        C.byref(st)), H.ctx.handle)
    stats = {k: getattr(st, k) for k, _ in L.RunStats._fields_}
    return f[:m], stats


def incremental_update_dev(H: DeviceGraph, src_ptr: int, dst_ptr: int, w_ptr: int, m: int,
                           found_ptr: int, cond_ptr: int, p: L.WalkParams, divisor: int = 10):
    """The batch loop on inputs already in device memory (dyg_incremental_update_dev).
    Pointers are device addresses (e.g. torch tensor data_ptr()); the caller
    makes sure the data is ready before the call."""
    st = L.RunStats()
    check(H._lib.dyg_incremental_update_dev(H.handle, src_ptr, dst_ptr, w_ptr, m, divisor,
                                            C.byref(p), found_ptr, cond_ptr, C.byref(st)),
          H.ctx.handle)
    return {k: getattr(st, k) for k, _ in L.RunStats._fields_}


class DualGraph:
    """GPU_Dual_Graph (random_walk_decremental/gpu_graph.cuh:79-687): G on the
    device with its edge map, H as a host key set."""

    def __init__(self, ctx: Context, V: int, g_row_ptr, g_col, g_w, h_u, h_v):
        self.ctx = ctx
        self._lib = ctx._lib
        rp, cl, ww = _u64(g_row_ptr), _u32(g_col), _f32(g_w)
        hu, hv = _u32(h_u), _u32(h_v)
        h = C.c_void_p()
        check(self._lib.dyg_dual_create(ctx.handle, V, ptr(rp, L.u64p), ptr(cl, L.u32p),
                                        ptr(ww, L.f32p), ptr(hu, L.u32p), ptr(hv, L.u32p),
                                        len(hu), C.byref(h)), ctx.handle)
        self.handle = h

    @property
    def multiplier(self) -> int:
        return self._lib.dyg_dual_key_multiplier(self.handle)

    def remove(self, u, v, ids=None) -> tuple[np.ndarray, int]:
        """removeAndUpdateDelEdges (gpu_graph.cuh:328-477)."""
        u, v = _u32(u), _u32(v)
        n = len(u)
        idv = None if ids is None else _u64(ids)
        status = np.zeros(max(n, 1), np.int8)
        q = C.c_uint32()
        check(self._lib.dyg_dual_remove(self.handle, ptr(u, L.u32p), ptr(v, L.u32p),
                                        ptr(idv, L.u64p), n, ptr(status, L.i8p), C.byref(q)),
              self.ctx.handle)
        return status[:n], q.value

    def walk(self, p: L.WalkParams) -> tuple[int, int]:
        """random_walk_decremental + incrementalUpdateSparse (gpu_graph.cuh:509-559)."""
        nop, add = C.c_uint32(), C.c_uint64()
        check(self._lib.dyg_dual_walk(self.handle, C.byref(p), C.byref(nop), C.byref(add), None),
              self.ctx.handle)
        return nop.value, add.value

    def heuristic(self, walkers: int = 16, seed: int = 0) -> None:
        """random_walk_small + heuristicRecovery (gpu_graph.cuh:577-615)."""
        check(self._lib.dyg_dual_heuristic(self.handle, walkers, seed), self.ctx.handle)

    @property
    def sparse_size(self) -> int:
        return self._lib.dyg_dual_sparse_size(self.handle)

    def added_edges(self) -> np.ndarray:
        n = self._lib.dyg_dual_added_count(self.handle)
        u = np.zeros(max(n, 1), np.uint32)
        v = np.zeros(max(n, 1), np.uint32)
        check(self._lib.dyg_dual_added_edges(self.handle, ptr(u, L.u32p), ptr(v, L.u32p)))
        return np.stack([u[:n], v[:n]], axis=1)

    def run(self, du, dv, p: L.WalkParams, divisor: int = 10, heuristic_walkers: int = 16):
        """Whole decremental loop (random_walk_decremental/main.cu:363-444)."""
        du, dv = _u32(du), _u32(dv)
        M = len(du)
        I = (M + divisor - 1) // divisor if M else 0
        log = []
        b0 = 0
        while True:
            n = min(I, M - b0)
            status, q = self.remove(du[b0:b0 + n], dv[b0:b0 + n], np.arange(b0, b0 + n))
            nop, add = self.walk(p)
            log.append(dict(batch_start=b0, queued=q, no_path=nop, added=add))
            if nop:
                self.heuristic(heuristic_walkers, p.seed)
            b0 += n
            if b0 >= M:
                break
        return log

    def close(self) -> None:
        if self.handle:
            self._lib.dyg_dual_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_params(w_ins: int = 256, w_del: int = 512, threshold: float = 100.0, seed: int = 0,
                  max_batch: int = 50000, max_inserts: int = 0, heur_walkers: int = 16,
                  slack: int = 2, max_steps: int = 100, plain: bool = False,
                  block_edges: int = 0, general_only: bool = False,
                  dense_no_desc: bool = False, steps_ins: int | None = None,
                  steps_del: int | None = None, block_edges_ins: int | None = None,
                  block_edges_del: int | None = None, graph: bool = False,
                  small_1024: bool = False) -> L.StreamParams:
    """Stream parameters; plain=True selects the reference-shaped form (host
    edge maps and edit plans, one workgroup per edge, a wait per step);
    general_only=True keeps small batches on the general pipeline;
    dense_no_desc=True keeps no neighbour descriptors in G (no refresh after
    its edits, deletion walks read the row table per hop).  steps_* and
    block_edges_* set the insertion and deletion walks apart (default:
    max_steps and block_edges for both); graph=True runs batches of at most
    2048 updates as captured hipGraphs of their one-workgroup step (apply()
    only; the served loop launches directly); small_1024=True keeps the
    one-workgroup path to batches of at most 1024 updates."""
    sp = L.StreamParams()
    sp.ins = params(w_ins, max_steps if steps_ins is None else steps_ins, threshold, seed,
                    block_edges=block_edges if block_edges_ins is None else block_edges_ins)
    sp.del_ = params(w_del, max_steps if steps_del is None else steps_del, 0.0, seed,
                     block_edges=block_edges if block_edges_del is None else block_edges_del)
    sp.heur_walkers, sp.max_batch, sp.max_inserts = heur_walkers, max_batch, max_inserts
    sp.slack = slack
    sp.flags = ((L.DYG_STREAM_PLAIN if plain else 0)
                | (L.DYG_STREAM_GENERAL_ONLY if general_only else 0)
                | (L.DYG_STREAM_DENSE_NO_DESC if dense_no_desc else 0)
                | (L.DYG_STREAM_GRAPH if graph else 0)
                | (L.DYG_STREAM_SMALL_1024 if small_1024 else 0))
    return sp


_STATS_NAMES = [k for k, _ in L.StreamStats._fields_]
_STATS_DTYPE = np.ctypeslib.as_array((L.StreamStats * 1)()).dtype


def _stats_dicts(sts, nb: int) -> list:
    """Per-batch dyg_stream_stats as dicts, read through one structured view."""
    rows = np.frombuffer(sts, dtype=_STATS_DTYPE, count=nb).tolist()
    return [dict(zip(_STATS_NAMES, r)) for r in rows]


class Stream:
    """Mixed insertion/deletion update stream on device-resident G and H
    (dyg_stream; dynamic_update/main.cu:92-124 batch loop semantics)."""

    def __init__(self, ctx: Context, V: int, g_edges, h_edges, sp: L.StreamParams,
                 csr: tuple | None = None):
        """g_edges / h_edges: (u, v, w) edge lists; or csr = (G's, H's)
        (row_ptr, col, w) from csr_from_edges, built once for many streams."""
        self.ctx = ctx
        self._lib = ctx._lib
        self.V = int(V)
        if csr is None:
            csr = (csr_from_edges(V, *g_edges), csr_from_edges(V, *h_edges))
        (grp, gcol, gw), (hrp, hcol, hw) = csr
        h = C.c_void_p()
        check(self._lib.dyg_stream_create(ctx.handle, self.V, ptr(grp, L.u64p), ptr(gcol, L.u32p),
                                          ptr(gw, L.f32p), ptr(hrp, L.u64p), ptr(hcol, L.u32p),
                                          ptr(hw, L.f32p), C.byref(sp), C.byref(h)), ctx.handle)
        self.handle = h
        self.G = DeviceGraph.borrowed(ctx, self._lib.dyg_stream_graph(h, 0), V, self)
        self.H = DeviceGraph.borrowed(ctx, self._lib.dyg_stream_graph(h, 1), V, self)

    @property
    def position(self) -> int:
        return self._lib.dyg_stream_position(self.handle)

    def apply(self, op: str, u, v, w=None) -> tuple[np.ndarray, dict]:
        """One batch: op "ins" or "del".  Returns per-update outcome codes and stats."""
        u, v = _u32(u), _u32(v)
        n = len(u)
        ww = _f32(w) if w is not None else None
        out = np.empty(max(n, 1), np.int32)
        st = L.StreamStats()
        code = L.DYG_STREAM_INSERT if op == "ins" else L.DYG_STREAM_DELETE
        check(self._lib.dyg_stream_apply(self.handle, code, L.addr(u), L.addr(v), L.addr(ww), n,
                                         L.addr(out), C.byref(st)),
              self.ctx.handle)
        return out[:n], {k: getattr(st, k) for k, _ in L.StreamStats._fields_}

    def apply_dev(self, op: str, u, v, w=None, n: int | None = None) -> tuple[np.ndarray, dict]:
        """One batch whose inputs are already in device memory: u, v (and w for
        insertions) are DeviceArray objects or raw device addresses; n defaults
        to len(u).  Same results as apply()."""
        du = u.ptr if isinstance(u, DeviceArray) else int(u)
        dv = v.ptr if isinstance(v, DeviceArray) else int(v)
        dw = (w.ptr if isinstance(w, DeviceArray) else int(w)) if w is not None else None
        if n is None:
            n = u.n if isinstance(u, DeviceArray) else 0
        out = np.empty(max(n, 1), np.int32)
        st = L.StreamStats()
        code = L.DYG_STREAM_INSERT if op == "ins" else L.DYG_STREAM_DELETE
        check(self._lib.dyg_stream_apply_dev(self.handle, code, du, dv, dw, n, L.addr(out),
                                             C.byref(st)), self.ctx.handle)
        return out[:n], {k: getattr(st, k) for k, _ in L.StreamStats._fields_}

    def apply_many_dev(self, ops, offsets, u, v, w=None) -> tuple[np.ndarray, list]:
        """Batches b = 0.. in one call: ops[b] ("ins" / "del") over updates
        offsets[b]:offsets[b+1] of the device arrays u, v, w (DeviceArray or
        raw addresses; w entries of deletion batches are ignored).  Returns every
        update's outcome code and the per-batch stats."""
        nb = len(ops)
        code = np.array([L.DYG_STREAM_INSERT if o == "ins" else L.DYG_STREAM_DELETE
                         for o in ops], np.int32)
        off = np.ascontiguousarray(offsets, np.uint64)
        assert len(off) == nb + 1
        addr = lambda a: (a.ptr if isinstance(a, DeviceArray) else int(a)) if a is not None else None
        out = np.empty(max(int(off[-1]), 1), np.int32)
        sts = (L.StreamStats * max(nb, 1))()
        check(self._lib.dyg_stream_apply_many_dev(self.handle, nb, L.addr(code), L.addr(off),
                                                  addr(u), addr(v), addr(w), L.addr(out), sts),
              self.ctx.handle)
        stats = _stats_dicts(sts, nb)
        return out[:int(off[-1])], stats

    def close(self) -> None:
        if self.handle:
            self._lib.dyg_stream_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop over a Stream (dyg_serve): requests are admitted into
    slots as slots and the memory budget allow, their inputs are copied while
    earlier requests run, and they are applied in admission order."""

    def __init__(self, stream: Stream, slots: int = 4, max_request: int = 10000,
                 budget_bytes: int = 0):
        self.stream = stream
        self._lib = stream._lib
        h = C.c_void_p()
        check(self._lib.dyg_serve_open(stream.handle, slots, max_request, budget_bytes,
                                       C.byref(h)), stream.ctx.handle)
        self.handle = h

    def submit(self, op: str, u, v, w=None) -> int:
        """Blocks until admitted; returns the ticket."""
        u, v = _u32(u), _u32(v)
        ww = _f32(w) if w is not None else None
        t = C.c_uint64()
        code = L.DYG_STREAM_INSERT if op == "ins" else L.DYG_STREAM_DELETE
        check(self._lib.dyg_serve_submit(self.handle, code, L.addr(u), L.addr(v), L.addr(ww),
                                         len(u), C.byref(t)),
              self.stream.ctx.handle)
        return t.value

    def submit_dev(self, op: str, u, v, w=None, n: int | None = None) -> int:
        """A request already in device memory (DeviceArray objects or raw device
        addresses; kept unchanged until its wait returns): ids are checked on
        the device before admission.  Returns the ticket."""
        du = u.ptr if isinstance(u, DeviceArray) else int(u)
        dv = v.ptr if isinstance(v, DeviceArray) else int(v)
        dw = (w.ptr if isinstance(w, DeviceArray) else int(w)) if w is not None else None
        if n is None:
            n = u.n if isinstance(u, DeviceArray) else 0
        t = C.c_uint64()
        code = L.DYG_STREAM_INSERT if op == "ins" else L.DYG_STREAM_DELETE
        check(self._lib.dyg_serve_submit_dev(self.handle, code, du, dv, dw, n, C.byref(t)),
              self.stream.ctx.handle)
        return t.value

    def wait(self, ticket: int, n: int) -> tuple[np.ndarray, dict]:
        out = np.zeros(max(n, 1), np.int32)
        res = L.ServeResult()
        check(self._lib.dyg_serve_wait(self.handle, ticket, L.addr(out), C.byref(res)),
              self.stream.ctx.handle)
        return out[:n], {k: getattr(res, k) for k, _ in L.ServeResult._fields_}

    def fail_next_copy(self) -> None:
        """Testing hook (dyg_serve_fail_next_copy): the next admitted request
        that copies its input reports the copy as failed."""
        check(self._lib.dyg_serve_fail_next_copy(self.handle), self.stream.ctx.handle)

    def close(self) -> None:
        if self.handle:
            self._lib.dyg_serve_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_mtx_auto(path: str):
    """Tolerant .mtx reader (dyg_mtx_read_auto): returns (info dict, V, u, v, w)
    with 0-based ids, self-loops and mirrored entries dropped."""
    lib = L.lib()
    info = L.MtxInfo()
    h = C.c_void_p()
    check(lib.dyg_mtx_read_auto(path.encode(), C.byref(info), C.byref(h)))
    try:
        n = lib.dyg_mtx_nlines(h)
        V = lib.dyg_mtx_nv(h)
        if n:
            u = np.ctypeslib.as_array(lib.dyg_mtx_line_u(h), shape=(n,)).copy()
            v = np.ctypeslib.as_array(lib.dyg_mtx_line_v(h), shape=(n,)).copy()
            w = np.ctypeslib.as_array(lib.dyg_mtx_line_w(h), shape=(n,)).copy()
        else:
            u = v = np.zeros(0, np.uint32)
            w = np.zeros(0, np.float32)
    finally:
        lib.dyg_mtx_free(h)
    return {k: getattr(info, k) for k, _ in L.MtxInfo._fields_}, V, u, v, w
