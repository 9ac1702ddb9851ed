"""FP64 kappa tool: dense generalized eigenvalues, closed forms (CPU only)."""
import numpy as np
import pytest
import scipy.linalg as la

from dygrass_amd.kappa import SHIFT, kappa


def _dense_lap(V, u, v, w):
    L = np.zeros((V, V))
    for a, b, x in zip(u, v, w):
        L[a, a] += x
        L[b, b] += x
        L[a, b] -= x
        L[b, a] -= x
    return L + SHIFT * np.eye(V)


def _cycle(n, w):
    u = np.arange(n)
    return u, (u + 1) % n, np.asarray(w, np.float64)


def _thomas(lo, di, up, rhs):
    """Tridiagonal solve (sub-, main-, super-diagonal), no pivoting."""
    n = len(di)
    c, d = np.zeros(n), np.zeros(n)
    c[0], d[0] = up[0] / di[0], rhs[0] / di[0]
    for i in range(1, n):
        m = di[i] - lo[i - 1] * c[i - 1]
        c[i] = up[i] / m if i < n - 1 else 0.0
        d[i] = (rhs[i] - lo[i - 1] * d[i - 1]) / m
    x = np.zeros(n)
    x[-1] = d[-1]
    for i in range(n - 2, -1, -1):
        x[i] = d[i] - c[i] * x[i + 1]
    return x


@pytest.mark.parametrize("n", [60, 2500])
def test_cycle_over_spanning_path_rank_one(n):
    # L_C = L_P + w_e b b^T with b = e_0 - e_{n-1}, so the pencil (L_C + eps I,
    # L_P + eps I) has eigenvalues 1 and 1 + w_e b^T (L_P + eps I)^-1 b;
    # the path matrix is tridiagonal and solved here by the Thomas algorithm
    rng = np.random.default_rng(n)
    w = rng.uniform(0.5, 2.0, n)
    u, v, ww = _cycle(n, w)
    wp = w[:-1]
    di = np.zeros(n)
    di[:-1] += wp
    di[1:] += wp
    di += SHIFT
    b = np.zeros(n)
    b[0], b[-1] = 1.0, -1.0
    x = _thomas(-wp, di, -wp, b)
    want = 1.0 + w[-1] * b @ x
    got = kappa(n, (u, v, ww), (u[:-1], v[:-1], ww[:-1]))
    assert got == pytest.approx(want, rel=1e-8)
    if n == 60:  # the shift is negligible here: 1 + w_e * effective resistance of the path
        assert got == pytest.approx(1.0 + w[-1] * np.sum(1.0 / wp), rel=1e-3)


def test_identical_graphs_give_one():
    u, v, w = _cycle(40, np.ones(40))
    assert kappa(40, (u, v, w), (u, v, w)) == pytest.approx(1.0, abs=1e-9)


@pytest.mark.parametrize("V", [200, 2100])
def test_matches_dense_generalized_eigenproblem(V):
    """Random connected graph vs its spanning tree plus a few edges; the dense
    reference is built entry by entry here and solved with LAPACK."""
    rng = np.random.default_rng(V)
    tu = np.arange(1, V)
    tv = rng.integers(0, tu)  # random recursive tree
    m = 3 * V
    eu, ev = rng.integers(0, V, m), rng.integers(0, V, m)
    keep = eu != ev
    gu = np.concatenate([tu, eu[keep]])
    gv = np.concatenate([tv, ev[keep]])
    gw = rng.uniform(0.1, 10.0, len(gu))
    extra = len(tu) + np.arange(0, keep.sum(), 20)
    hi = np.concatenate([np.arange(len(tu)), extra])
    G = (gu, gv, gw)
    H = (gu[hi], gv[hi], gw[hi])
    want = la.eigh(_dense_lap(V, *G), _dense_lap(V, *H), eigvals_only=True)[-1]
    assert kappa(V, G, H) == pytest.approx(want, rel=1e-7)


def test_kappa_cli_reads_both_files(tmp_path, capsys):
    from dygrass_amd.kappa import main

    n = 30
    u = np.arange(n)
    v = (u + 1) % n
    (tmp_path / "g.mtx").write_text("".join(f"{a + 1} {b + 1} 1\n" for a, b in zip(u, v)))
    (tmp_path / "h.mtx").write_text("%%MatrixMarket matrix coordinate real symmetric\n"
                                    f"{n} {n} {n - 1}\n" +
                                    "".join(f"{a + 1} {b + 1} 1\n" for a, b in zip(u[:-1], v[:-1])))
    assert main([str(tmp_path / "g.mtx"), str(tmp_path / "h.mtx")]) == 0
    got = float(capsys.readouterr().out.split()[-1])
    assert got == pytest.approx(n, rel=1e-3)  # 1 + resistance of the 29-edge path


def test_pcg_kappa_equals_dense_and_lu():
    """The CG path (L_H^-1 by diagonal-preconditioned conjugate gradients on
    the host's cores) against the dense generalized eigenproblem on a small
    3D grid with random shortcuts (the served workload's family) forced
    through Lanczos, and against the sparse-LU path; also independent of the
    thread count (block-ordered reductions)."""
    from dygrass_amd import synth

    wl = synth.config_served(12 ** 3 / 8e6, 0, count=4)  # 12^3 vertices
    G = (wl.G.u, wl.G.v, wl.G.w)
    H = (wl.H.u, wl.H.v, wl.H.w)
    V = wl.G.V
    assert V == 12 ** 3 and len(wl.H) < len(wl.G)
    dense = float(la.eigh(_dense_lap(V, *G), _dense_lap(V, *H), eigvals_only=True)[-1])
    info = {}
    k1 = kappa(V, G, H, tol=1e-12, method="pcg", threads=1, info=info, dense_limit=0)
    k4 = kappa(V, G, H, tol=1e-12, method="pcg", threads=4, dense_limit=0)
    klu = kappa(V, G, H, tol=1e-12, method="lu", dense_limit=0)
    assert info["solves"] > 0 and info["cg_iterations"] > 0
    assert abs(k1 - dense) <= 1e-9 * dense
    assert abs(klu - dense) <= 1e-9 * dense
    assert k1 == k4
    assert dense > 1.0


def test_pcg_kappa_cycle_closed_form():
    """kappa of a weighted cycle over its spanning path through the CG path
    (the closed form 1 + w_c * R_path, as for the sparse-LU path)."""
    n = 400
    rng = np.random.default_rng(7)
    w = rng.uniform(0.5, 2.0, n)
    u, v, ww = _cycle(n, w)
    path = (u[:-1], v[:-1], ww[:-1])
    k = kappa(n, (u, v, ww), path, tol=1e-12, method="pcg", dense_limit=0)
    # rank-one update on the path: lambda_max = 1 + w_c * b' L_P^-1 b, L_P shifted
    L = _dense_lap(n, *path)
    b = np.zeros(n)
    b[n - 1], b[0] = 1.0, -1.0
    want = 1.0 + ww[-1] * b @ np.linalg.solve(L, b)
    assert abs(k - want) <= 1e-7 * want
