"""Drop-in CLIs against the oracle's restatement of the reference mains.

incremental_update: updated_adj_sparse.mtx must be byte-identical and the
standard output identical line by line (the timing line excepted).
decremental_update: added_edges.mtx equal as a set (the reference writes it
in hash order) and the standard output identical (timing line excepted).
"""
import os
import shutil
import subprocess
import tempfile

import pytest

from dygrass_amd import BIN_DIR, synth

pytestmark = pytest.mark.gpu


def _lines(text):
    return [ln for ln in text.splitlines() if not ln.startswith("Time taken:")]


def _run_cli(root, args):
    bindir = os.path.join(root, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(BIN_DIR, args[0])
    p = subprocess.run([exe, *args[1:]], cwd=bindir, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


@pytest.mark.parametrize("W,thr", [(512, 10.0), (64, 4.0)])
def test_incremental_cli_matches_oracle(grid_wl, W, thr):
    from oracle import pyoracle as po

    with tempfile.TemporaryDirectory() as root:
        d = synth.write_dataset(root, grid_wl)
        ref = os.path.join(root, "oracle_copy")
        shutil.copytree(d, ref)
        out = _run_cli(root, ["incremental_update", grid_wl.name, str(thr), str(W)])
        log = os.path.join(root, "oracle.log")
        po.incremental_run(ref, thr, W, seed=0, threads=8, log_path=log)
        a = open(os.path.join(d, "updated_adj_sparse.mtx"), "rb").read()
        b = open(os.path.join(ref, "updated_adj_sparse.mtx"), "rb").read()
        assert a == b
        exp = open(log).read().replace(ref, os.path.join("..", "dataset", grid_wl.name))
        assert _lines(out) == _lines(exp)


def test_incremental_cli_usage_and_bad_walkers(grid_wl):
    with tempfile.TemporaryDirectory() as root:
        synth.write_dataset(root, grid_wl)
        out = _run_cli(root, ["incremental_update", grid_wl.name])
        assert out.startswith("Input: ./incremental_update")
        exe = os.path.join(BIN_DIR, "incremental_update")
        os.makedirs(os.path.join(root, "bin"), exist_ok=True)
        p = subprocess.run([exe, grid_wl.name, "10", "48"], cwd=os.path.join(root, "bin"),
                           capture_output=True, text=True, timeout=120)
        assert p.returncode != 0 and "power of two" in p.stdout


def test_decremental_cli_matches_oracle(mesh_wl):
    from oracle import pyoracle as po

    with tempfile.TemporaryDirectory() as root:
        d = synth.write_dataset(root, mesh_wl)
        ref = os.path.join(root, "oracle_copy")
        shutil.copytree(d, ref)
        out = _run_cli(root, ["decremental_update", mesh_wl.name])
        log = os.path.join(root, "oracle.log")
        st = po.decremental_run(ref, seed=0, threads=8, log_path=log)

        def pairs(p):
            return {tuple(int(x) for x in ln.split()) for ln in open(p) if ln.strip()}

        got = pairs(os.path.join(d, "added_edges.mtx"))
        exp = pairs(os.path.join(ref, "added_edges.mtx"))
        assert got == exp and len(got) == st.added_edges
        exp_log = open(log).read().replace(ref, os.path.join("..", "dataset", mesh_wl.name))
        assert _lines(out) == _lines(exp_log)
        assert st.filtered > 0


def test_harness_pipeline_end_to_end(grid_wl):
    """The harness's whole flow (dyGRASS.jl:408-427, 467-549) on one dataset:
    incremental CLI, updated_dense.mtx, deletions, decremental CLI, final
    graphs and kappa -- the same files, graphs and kappa as the oracle's run
    of the same flow."""
    import numpy as np

    from dygrass_amd import dataset
    from dygrass_amd.kappa import kappa
    from oracle import pyoracle as po

    wl = grid_wl
    dels = synth.deletions(wl.G, 300, 11)
    with tempfile.TemporaryDirectory() as root:
        d = synth.write_dataset(root, wl)
        ref = os.path.join(root, "oracle_copy")
        shutil.copytree(d, ref)
        _run_cli(root, ["incremental_update", wl.name, "10", "128"])
        po.incremental_run(ref, 10.0, 128, seed=0, threads=8)
        for x in (d, ref):
            dataset.prepare_decremental(x)
            synth.write_mtx(os.path.join(x, "del.mtx"), dels.u, dels.v, None, base=1)
        _run_cli(root, ["decremental_update", wl.name])
        po.decremental_run(ref, seed=0, threads=8)
        V, G, H = dataset.final_graphs(d)
        V2, G2, H2 = dataset.final_graphs(ref)
        assert V == V2
        key = lambda E: sorted(zip(E[0].tolist(), E[1].tolist(), E[2].tolist()))  # noqa: E731
        assert key(G) == key(G2) and key(H) == key(H2)
        k = kappa(V, G, H)
        assert k == kappa(V2, G2, H2)
        assert 1.0 <= k < np.inf
