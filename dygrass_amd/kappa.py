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

    python -m dygrass_amd.kappa dense.mtx sparse.mtx

(files read with the tolerant reader: header, dimensions line and id base
are detected; both files must use the same id base).
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
    # canonical edge order: the same graph gives the same sums, bit for bit
    a, b = np.minimum(u, v), np.maximum(u, v)
    order = np.lexsort((w, b, a))
    u, v, w = a[order], b[order], w[order]
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
    v0 = np.random.default_rng(0).uniform(0.5, 1.5, V)  # fixed start: reproducible value
    vals = sla.eigsh(LG, k=1, M=LH, Minv=Minv, which="LA", tol=tol, v0=v0,
                     return_eigenvectors=False)
    return float(vals[-1])


def _read(path: str):
    from .core import read_mtx_auto

    _, V, u, v, w = read_mtx_auto(path)  # header, dimensions and id base detected
    return V, (u.astype(np.int64), v.astype(np.int64), w.astype(np.float64))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="kappa(L_G, L_H) = lambda_max of the pencil")
    ap.add_argument("dense")
    ap.add_argument("sparse")
    a = ap.parse_args(argv)
    vg, g = _read(a.dense)
    vh, h = _read(a.sparse)
    print(f"CND: {kappa(max(vg, vh), g, h):.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
