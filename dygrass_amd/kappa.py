"""Relative condition number kappa(L_G, L_H) in FP64 on the host.

Follows dyGRASS.jl `create_laplacian` (:54-75: graph Laplacian of the
weighted adjacency, duplicate edges summed, plus 1e-6 on the diagonal) and
the check `maximum(real(eigs(L_G, L_H)[1]))` (dyGRASS.jl:361, :543;
dynamic_update/cnd_analysis.jl).  The reference calls Arpack.jl `eigs`
(implicitly restarted Arnoldi, unpinned version) on the pencil; the pencil is
symmetric with L_H positive definite, so this tool runs ARPACK's symmetric
Lanczos driver (scipy.sparse.linalg.eigsh, largest algebraic eigenvalue) on
L_G x = lambda L_H x, with L_H^-1 applied through a sparse LU factorisation
(L_H is a spanning tree plus a few percent of edges, so its fill is small).

    python -m dygrass_amd.kappa dense.mtx sparse.mtx [--base 1]
"""
from __future__ import annotations

import argparse
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

SHIFT = 1e-6


def laplacian(V: int, u, v, w, shift: float = SHIFT) -> sp.csc_matrix:
    """L = D - A (+ shift I) in FP64; A symmetric with duplicate edges summed."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.asarray(w, np.float64)
    keep = u != v  # a self-loop adds nothing to a Laplacian
    u, v, w = u[keep], v[keep], w[keep]
    A = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([u, v]), np.concatenate([v, u]))),
                      shape=(V, V)).tocsr()
    A.sum_duplicates()
    d = np.asarray(A.sum(axis=1)).ravel()
    return (sp.diags(d + shift) - A).tocsc()


def kappa(V: int, g_edges, h_edges, tol: float = 1e-10) -> float:
    """lambda_max of the pencil (L_G, L_H), both with the 1e-6 diagonal shift."""
    LG = laplacian(V, *g_edges)
    LH = laplacian(V, *h_edges)
    if V <= 2000:  # small: dense generalized eigenproblem
        import scipy.linalg as la

        return float(la.eigh(LG.toarray(), LH.toarray(), eigvals_only=True)[-1])
    lu = sla.splu(LH, permc_spec="COLAMD")
    Minv = sla.LinearOperator((V, V), matvec=lu.solve, dtype=np.float64)
    vals = sla.eigsh(LG, k=1, M=LH, Minv=Minv, which="LA", tol=tol,
                     return_eigenvectors=False)
    return float(vals[-1])


def _read(path: str, base: int):
    from ._lib import lib, check
    import ctypes as C

    L = lib()
    h = C.c_void_p()
    check(L.dyg_mtx_read_edges(path.encode(), 1, C.byref(h)))
    try:
        n = L.dyg_mtx_nlines(h)
        u = np.ctypeslib.as_array(L.dyg_mtx_line_u(h), shape=(n,)).copy()
        v = np.ctypeslib.as_array(L.dyg_mtx_line_v(h), shape=(n,)).copy()
        w = np.ctypeslib.as_array(L.dyg_mtx_line_w(h), shape=(n,)).astype(np.float64)
    finally:
        L.dyg_mtx_free(h)
    if base == 0:  # the reader subtracts one from every id (uint32, so 0 wraps)
        u, v = u + np.uint32(1), v + np.uint32(1)
    return u.astype(np.int64), v.astype(np.int64), w


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="kappa(L_G, L_H) = lambda_max of the pencil")
    ap.add_argument("dense")
    ap.add_argument("sparse")
    ap.add_argument("--base", type=int, default=1, help="id base of both files (0 or 1)")
    a = ap.parse_args(argv)
    g = _read(a.dense, a.base)
    h = _read(a.sparse, a.base)
    V = int(max(g[0].max(), g[1].max(), h[0].max(), h[1].max())) + 1
    print(f"CND: {kappa(V, g, h):.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
