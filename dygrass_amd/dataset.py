"""The harness's file-set steps around the two updaters, without Julia.

dyGRASS.jl prepares and post-processes a graph directory around the CLIs:

- prepareDecremental (dyGRASS.jl:408-427): updated_dense.mtx = the lines of
  dense.mtx followed by those of ext.mtx (the dense graph after the
  insertions), the decremental updater's dense input;
- decremental_test (dyGRASS.jl:467-549): after both updaters ran, the final
  dense graph is (dense ∪ ext) minus del.mtx and the final sparsifier is
  updated_adj_sparse.mtx minus del.mtx plus added_edges.mtx, each added edge
  taking its weight from the final dense graph; kappa of the pair is the
  quality check.

Edges are keyed as (min, max); as with the harness's dictionaries, a later
line with the same key replaces an earlier one.

    python -m dygrass_amd.dataset prepare <graph dir>
    python -m dygrass_amd.dataset final <graph dir> [--kappa]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np


def _read(path: str, weighted: bool = True):
    """A 1-based harness file, "a b w" or "a b" per line (the reference's reader)."""
    import ctypes as C

    from ._lib import check, lib

    L = lib()
    h = C.c_void_p()
    check(L.dyg_mtx_read_edges(path.encode(), int(weighted), C.byref(h)))
    try:
        n = L.dyg_mtx_nlines(h)
        V = L.dyg_mtx_nv(h)
        if n == 0:
            return V, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32)
        u = np.ctypeslib.as_array(L.dyg_mtx_line_u(h), shape=(n,)).astype(np.int64)
        v = np.ctypeslib.as_array(L.dyg_mtx_line_v(h), shape=(n,)).astype(np.int64)
        w = np.ctypeslib.as_array(L.dyg_mtx_line_w(h), shape=(n,)).copy()
    finally:
        L.dyg_mtx_free(h)
    return V, u, v, w


def _write(path: str, u, v, w=None) -> None:
    """1-based lines "a b w" (or "a b"), weights with 9 significant digits."""
    with open(path, "w") as f:
        if w is None:
            f.writelines(f"{a + 1} {b + 1}\n" for a, b in zip(u.tolist(), v.tolist()))
        else:
            f.writelines(f"{a + 1} {b + 1} {x:.9g}\n"
                         for a, b, x in zip(u.tolist(), v.tolist(), w.tolist()))


def prepare_decremental(d: str) -> str:
    """dyGRASS.jl:408-427: updated_dense.mtx = dense.mtx lines + ext.mtx lines."""
    _, gu, gv, gw = _read(os.path.join(d, "dense.mtx"))
    _, eu, ev, ew = _read(os.path.join(d, "ext.mtx"))
    out = os.path.join(d, "updated_dense.mtx")
    _write(out, np.concatenate([gu, eu]), np.concatenate([gv, ev]), np.concatenate([gw, ew]))
    return out


def _dict(u, v, w) -> dict:
    return {(min(a, b), max(a, b)): x for a, b, x in zip(u.tolist(), v.tolist(), w.tolist())}


def final_graphs(d: str):
    """dyGRASS.jl:467-549: (V, final dense edges, final sparsifier edges) as
    (u, v, w) arrays, 0-based."""
    V1, gu, gv, gw = _read(os.path.join(d, "dense.mtx"))
    V2, eu, ev, ew = _read(os.path.join(d, "ext.mtx"))
    _, du, dv, _ = _read(os.path.join(d, "del.mtx"), weighted=False)
    V3, su, sv, sw = _read(os.path.join(d, "updated_adj_sparse.mtx"))
    G = _dict(np.concatenate([gu, eu]), np.concatenate([gv, ev]), np.concatenate([gw, ew]))
    H = _dict(su, sv, sw)
    for a, b in zip(du.tolist(), dv.tolist()):
        G.pop((min(a, b), max(a, b)), None)
        H.pop((min(a, b), max(a, b)), None)
    added = os.path.join(d, "added_edges.mtx")
    if os.path.exists(added) and os.path.getsize(added) > 0:
        _, au, av, _ = _read(added, weighted=False)
        for a, b in zip(au.tolist(), av.tolist()):
            k = (min(a, b), max(a, b))
            if k in G:
                H[k] = G[k]

    def arrays(D):
        if not D:
            return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32)
        k = np.array(list(D.keys()), np.int64)
        return k[:, 0], k[:, 1], np.array(list(D.values()), np.float32)

    return max(V1, V2, V3), arrays(G), arrays(H)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="file-set steps of the dyGRASS harness")
    ap.add_argument("step", choices=["prepare", "final"])
    ap.add_argument("dir")
    ap.add_argument("--kappa", action="store_true", help="final: also print kappa(G, H)")
    a = ap.parse_args(argv)
    if a.step == "prepare":
        print(prepare_decremental(a.dir))
        return 0
    V, G, H = final_graphs(a.dir)
    print(f"final dense edges: {len(G[0])}, final sparsifier edges: {len(H[0])}, "
          f"density: {len(H[0]) / V:.4f}")
    if a.kappa:
        from .kappa import kappa

        print(f"CND: {kappa(V, G, H):.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
