"""The dual graph's two-hop recovery with real no-path deletions.

The graph is a chain of 10x10 grid clusters joined by single bridges.  Every
bridge lies in every spanning tree, so the sparsifier holds all of them;
deleting a bridge leaves no path between its endpoints, so its walks find
nothing and the two-hop recovery runs from both endpoints
(random_walk_decremental/main.cu:398-408, gpu_graph.cuh:577-615).  Cluster
deletions are mixed in, so the same run also recovers found paths.  Both the
dual graph through the C ABI (DualGraph.run) and the decremental_update CLI
are compared with the oracle's restatement of the whole run.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import csr

pytestmark = pytest.mark.gpu


def bridged_clusters(k=40, r=10, seed=0):
    """G: k r-by-r grids in a chain, cluster c joined to c+1 by one bridge
    between random vertices; weights log-uniform [0.5, 2].  H: a comb
    spanning tree of each grid (first row plus every column), every bridge,
    and 5% of the other grid edges.  Deletions (in random order): every bridge
    plus grid edges, no vertex losing more than one grid edge."""
    from dygrass_amd import synth

    rng = np.random.default_rng(seed)
    cu, cv, tree = [], [], []
    for c in range(k):
        base = c * r * r
        for i in range(r):
            for j in range(r):
                x = base + i * r + j
                if j + 1 < r:
                    cu.append(x); cv.append(x + 1); tree.append(i == 0)
                if i + 1 < r:
                    cu.append(x); cv.append(x + r); tree.append(True)
    ncl = len(cu)
    bu = [c * r * r + int(rng.integers(r * r)) for c in range(k - 1)]
    bv = [(c + 1) * r * r + int(rng.integers(r * r)) for c in range(k - 1)]
    u = np.array(cu + bu, np.uint32)
    v = np.array(cv + bv, np.uint32)
    w = np.exp(rng.uniform(np.log(0.5), np.log(2.0), len(u))).astype(np.float32)
    V = k * r * r
    G = synth.Edges(V, u, v, w)
    in_h = np.concatenate([np.array(tree), np.ones(k - 1, bool)])
    in_h[:ncl] |= rng.uniform(size=ncl) < 0.05
    H = synth.Edges(V, u[in_h], v[in_h], w[in_h])
    used = np.zeros(V, bool)
    pick = []
    for e in rng.permutation(ncl):
        if len(pick) == 300:
            break
        if not used[u[e]] and not used[v[e]]:
            used[u[e]] = used[v[e]] = True
            pick.append(e)
    dels = rng.permutation(np.array(pick + list(range(ncl, ncl + k - 1))))
    D = synth.Edges(V, u[dels], v[dels], w[dels])
    return synth.Workload(f"bridged{k}x{r}", G, H, dels=D)


def test_dual_graph_heuristic_matches_oracle(ctx):
    from dygrass_amd import DualGraph, params
    from oracle import pyoracle as po

    wl = bridged_clusters()
    rp, col, w = csr(wl.G)
    d = DualGraph(ctx, wl.G.V, rp, col, w, wl.H.u, wl.H.v)
    before = d.sparse_size
    log = d.run(wl.dels.u, wl.dels.v, params(512, 100, seed=0))
    added = d.added_edges()
    st, oadded = po.decremental_lines(wl.G.V, wl.G.u, wl.G.v, wl.G.w, wl.H.u, wl.H.v,
                                      wl.dels.u, wl.dels.v, seed=0, threads=8)
    n_bridges = 39
    assert st.no_path == n_bridges  # every deleted bridge was in H and has no path
    assert sum(b["no_path"] for b in log) == st.no_path
    assert sum(b["queued"] for b in log) == st.filtered
    assert st.added_edges > 0  # cluster deletions recovered paths too
    assert before == st.sparse_before
    assert d.sparse_size == st.sparse_after
    # the recovery inserted edges H lacked: H grew beyond what path recovery alone adds
    assert d.sparse_size > before - st.filtered + st.added_edges
    assert np.array_equal(added, oadded)
    d.close()


def test_decremental_cli_heuristic_matches_oracle():
    from dygrass_amd import BIN_DIR, synth
    from oracle import pyoracle as po

    wl = bridged_clusters(seed=1)
    with tempfile.TemporaryDirectory() as root:
        d = synth.write_dataset(root, wl)
        ref = os.path.join(root, "oracle_copy")
        shutil.copytree(d, ref)
        bindir = os.path.join(root, "bin")
        os.makedirs(bindir)
        p = subprocess.run([os.path.join(BIN_DIR, "decremental_update"), wl.name], cwd=bindir,
                           capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stdout + p.stderr
        log = os.path.join(root, "oracle.log")
        st = po.decremental_run(ref, seed=0, threads=8, log_path=log)
        assert st.no_path == 39

        def pairs(path):
            return {tuple(int(x) for x in ln.split()) for ln in open(path) if ln.strip()}

        assert pairs(os.path.join(d, "added_edges.mtx")) == pairs(
            os.path.join(ref, "added_edges.mtx"))
        drop = lambda t: [ln for ln in t.splitlines()  # noqa: E731
                          if not ln.startswith("Time taken:")]
        exp = open(log).read().replace(ref, os.path.join("..", "dataset", wl.name))
        assert drop(p.stdout) == drop(exp)
        assert any(ln.startswith("No path count: ") and ln != "No path count: 0"
                   for ln in p.stdout.splitlines())
