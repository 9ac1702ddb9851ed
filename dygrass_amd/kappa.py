"""Relative condition number kappa(L_G, L_H) in FP64 on the host.

Follows dyGRASS.jl `create_laplacian` (:54-75: graph Laplacian of the
weighted adjacency, duplicate edges summed, plus 1e-6 on the diagonal) and
the check `maximum(real(eigs(L_G, L_H)[1]))` (dyGRASS.jl:361, :543;
dynamic_update/cnd_analysis.jl).  The reference calls Arpack.jl `eigs`
(implicitly restarted Arnoldi, unpinned version) on the pencil; the pencil is
symmetric with L_H positive definite, so this tool runs ARPACK's symmetric
Lanczos driver (scipy.sparse.linalg.eigsh, largest algebraic eigenvalue) on
L_G x = lambda L_H x.  L_H^-1 is applied either through a sparse LU
factorisation (method "lu": L_H is a spanning tree plus a few percent of
edges, so a geometric sparsifier fills in little) or by conjugate gradients
preconditioned with L_H's diagonal, in FP64 on the host's cores
(libdygrass_kappa.so; method "pcg", with L_G's and L_H's products there too):
a sparsifier holding random long-range edges (the served workload's 3D grid
with shortcuts) fills a factorisation in, while those edges make it well
conditioned for CG.  Both give the same value to the solve tolerance.

    python -m dygrass_amd.kappa dense.mtx sparse.mtx [--method lu|pcg]

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


class _HostCSR:
    """A CSR matrix handed to libdygrass_kappa.so: products and CG solves on
    the host's cores in FP64."""

    def __init__(self, A: sp.spmatrix, threads: int):
        from ._lib import kappa_lib

        A = A.tocsr()
        self.n = A.shape[0]
        self.ip = np.ascontiguousarray(A.indptr, np.int64)
        self.ix = np.ascontiguousarray(A.indices, np.int32)
        self.a = np.ascontiguousarray(A.data, np.float64)
        self.threads = threads
        self.lib = kappa_lib()
        self.solves = 0
        self.iterations = 0

    def matvec(self, x):
        x = np.ascontiguousarray(x, np.float64).ravel()
        y = np.empty(self.n)
        self.lib.dyg_kappa_spmv(self.n, self.ip.ctypes.data, self.ix.ctypes.data,
                                self.a.ctypes.data, x.ctypes.data, y.ctypes.data, self.threads)
        return y

    def solve(self, b, tol: float, maxit: int = 100000):
        b = np.ascontiguousarray(b, np.float64).ravel()
        x = np.empty(self.n)
        it = self.lib.dyg_kappa_pcg(self.n, self.ip.ctypes.data, self.ix.ctypes.data,
                                    self.a.ctypes.data, b.ctypes.data, x.ctypes.data, tol, maxit,
                                    self.threads)
        if it < 0:
            raise RuntimeError(f"kappa: CG on L_H did not converge ({it})")
        self.solves += 1
        self.iterations += it
        return x


def host_threads() -> int:
    """The cores this process may use (OMP_NUM_THREADS when set: a GPU box
    allots its host cores per GPU that way)."""
    import os

    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def kappa(V: int, g_edges, h_edges, tol: float = 1e-10, method: str = "lu",
          threads: int = 0, info: dict | None = None, dense_limit: int = 2000,
          progress=None) -> float:
    """lambda_max of the pencil (L_G, L_H), both with the 1e-6 diagonal shift.
    Up to dense_limit vertices a dense generalized eigenproblem; above it
    Lanczos with method "lu" (sparse LU of L_H) or "pcg" (diagonal-
    preconditioned CG solves to 1e-3 * tol relative residual and the products
    on the host's cores; threads 0 = host_threads()).  `info` (optional)
    receives the solve counts; `progress` (optional) is called with them
    after every solve."""
    LG = laplacian(V, *g_edges)
    LH = laplacian(V, *h_edges)
    if V <= dense_limit:  # small: dense generalized eigenproblem
        import scipy.linalg as la

        return float(la.eigh(LG.toarray(), LH.toarray(), eigvals_only=True)[-1])
    v0 = np.random.default_rng(0).uniform(0.5, 1.5, V)  # fixed start: reproducible value
    if method == "lu":
        lu = sla.splu(LH, permc_spec="COLAMD")
        Minv = sla.LinearOperator((V, V), matvec=lu.solve, dtype=np.float64)
        vals = sla.eigsh(LG, k=1, M=LH, Minv=Minv, which="LA", tol=tol, v0=v0,
                         return_eigenvectors=False)
        return float(vals[-1])
    if method != "pcg":
        raise ValueError(f"kappa: unknown method {method!r}")
    nt = threads or host_threads()
    g, h = _HostCSR(LG, nt), _HostCSR(LH, nt)
    inner = max(tol * 1e-3, 1e-14)

    def solve(x):
        y = h.solve(x, inner)
        if progress is not None:
            progress(dict(solves=h.solves, cg_iterations=h.iterations))
        return y

    A = sla.LinearOperator((V, V), matvec=g.matvec, dtype=np.float64)
    M = sla.LinearOperator((V, V), matvec=h.matvec, dtype=np.float64)
    Minv = sla.LinearOperator((V, V), matvec=solve, dtype=np.float64)
    vals = sla.eigsh(A, k=1, M=M, Minv=Minv, which="LA", tol=tol, v0=v0,
                     return_eigenvectors=False)
    if info is not None:
        info.update(solves=h.solves, cg_iterations=h.iterations, threads=nt)
    return float(vals[-1])


def _read(path: str):
    from .core import read_mtx_auto

    _, V, u, v, w = read_mtx_auto(path)  # header, dimensions and id base detected
    return V, (u.astype(np.int64), v.astype(np.int64), w.astype(np.float64))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="kappa(L_G, L_H) = lambda_max of the pencil")
    ap.add_argument("dense")
    ap.add_argument("sparse")
    ap.add_argument("--method", choices=["lu", "pcg"], default="lu",
                    help="how L_H^-1 is applied: sparse LU, or CG on the host's cores")
    a = ap.parse_args(argv)
    vg, g = _read(a.dense)
    vh, h = _read(a.sparse)
    print(f"CND: {kappa(max(vg, vh), g, h, method=a.method):.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
