"""Full-size checks on the GPU (1M-node random geometric graph, 200k
insertions in 10 snapshot batches, threshold 100, 256 walkers).

Every one of the 200k decisions must equal the oracle's, and the final
sparsifier must equal the oracle's slot for slot.  Size-independent
properties are checked as well: H grows by exactly two slots per added edge,
a second run reproduces the first bit for bit.
"""
import os

import numpy as np
import pytest

from conftest import csr, f32bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rgg():
    from dygrass_amd import synth

    return synth.config_rgg_insert(1.0, 0)


def test_full_size_incremental_matches_oracle(ctx, rgg):
    from dygrass_amd import DeviceGraph, incremental_update, params
    from oracle import pyoracle as po

    rp, col, w = csr(rgg.H)
    p = params(rgg.walkers, 100, rgg.threshold, seed=0)
    H = DeviceGraph(ctx, rgg.H.V, rp, col, w, slack=2)
    _, n0, _ = H.info()
    found, st = incremental_update(H, rgg.ins.u, rgg.ins.v, rgg.ins.w, p)
    _, n1, _ = H.info()
    added = int((found == -1).sum())
    assert st["added"] == added and n1 - n0 == 2 * added
    assert np.all((found == -1) | ((found >= 1) & (found <= 99)))

    og = po.Graph(rgg.H.V, rp, col, w)
    threads = min(os.cpu_count() or 1, 16)
    ofound, ost = og.incremental(rgg.ins.u, rgg.ins.v, rgg.ins.w, rgg.threshold, rgg.walkers,
                                 seed=0, threads=threads)
    assert np.array_equal(found, ofound)
    grp, gcol, gw = H.download()
    orp, ocol, ow = og.export()
    assert np.array_equal(grp, orp) and np.array_equal(gcol, ocol)
    assert np.array_equal(f32bits(gw), f32bits(ow))
    assert H.check() == 0

    H2 = DeviceGraph(ctx, rgg.H.V, rp, col, w, slack=0)
    found2, _ = incremental_update(H2, rgg.ins.u, rgg.ins.v, rgg.ins.w, p)
    assert np.array_equal(found, found2)


def test_full_size_decremental_matches_oracle(ctx):
    """2M-node triangulated mesh, 100k uniform deletions in 10 snapshot batches,
    512 walkers: the recovered edge set, the sparsifier size and every counter
    equal the oracle's."""
    from dygrass_amd import DualGraph, params, synth
    from oracle import pyoracle as po

    wl = synth.config_mesh_delete(1.0, 0)
    rp, col, w = csr(wl.G)
    d = DualGraph(ctx, wl.G.V, rp, col, w, wl.H.u, wl.H.v)
    before = d.sparse_size
    log = d.run(wl.dels.u, wl.dels.v, params(512, 100, seed=0))
    added = d.added_edges()
    threads = min(os.cpu_count() or 1, 16)
    st, oadded = po.decremental_lines(wl.G.V, wl.G.u, wl.G.v, wl.G.w, wl.H.u, wl.H.v,
                                      wl.dels.u, wl.dels.v, seed=0, threads=threads)
    assert before == st.sparse_before
    assert d.sparse_size == st.sparse_after
    assert np.array_equal(added, oadded)
    assert sum(b["queued"] for b in log) == st.filtered
    assert sum(b["no_path"] for b in log) == st.no_path
    assert len(log) == 10 and st.filtered > 20000
