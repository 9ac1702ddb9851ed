"""File contract: parsing, batch layout and output text.

The oracle's parse restatement is pinned against the reference's own host
parsers (random_walk_incremental/functions.cpp and the decremental
read_targets), compiled from the reference sources into oracle/_ref by
`make -C oracle ref`; those tests skip where oracle/_ref was not built.  The
product's parser (libdygrass.so, host code) is pinned against the oracle, and
its weight printing against C++ `ostream << float` (the oracle's writer).
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from dygrass_amd import _lib as L
from oracle import pyoracle as po

REF_INC = os.path.join(po.REF_DIR, "ref_dump_inc")
REF_DEC = os.path.join(po.REF_DIR, "ref_dump_dec")
need_ref = pytest.mark.skipif(not os.path.exists(REF_INC) or not os.path.exists(REF_DEC),
                              reason="oracle/_ref (reference parsers) not built")


def sample_graph_text(rng, V=30, m=80, selfloop=True, dup=True):
    lines = []
    for i in range(m):
        a, b = (int(x) for x in rng.integers(0, V, 2))
        if a == b:
            b = (a + 1) % V
        w = float(np.float32(rng.uniform(0.01, 50)))
        sep = rng.choice([" ", "\t", "  "])
        lines.append(f"{a}{sep}{b}{sep}{w:.9g}")
    if selfloop:
        lines.insert(5, f"3 3 {0.5:.9g}")
    if dup:
        lines.append(lines[10])
    lines.append(f"{V - 1} 0 2.5")
    return "\n".join(lines) + "\n"


def product_graph(path, base):
    lib = L.lib()
    h = C.c_void_p()
    L.check(lib.dyg_mtx_read_graph(path.encode(), base, 1, C.byref(h)))
    V = lib.dyg_mtx_nv(h)
    rp = np.ctypeslib.as_array(lib.dyg_mtx_row_ptr(h), (V + 1,)).copy()
    n = int(rp[-1])
    col = np.ctypeslib.as_array(lib.dyg_mtx_col(h), (n,)).copy()
    w = np.ctypeslib.as_array(lib.dyg_mtx_w(h), (n,)).copy()
    lib.dyg_mtx_free(h)
    return V, rp, col, w


@need_ref
@pytest.mark.parametrize("seed", [0, 1])
def test_graph_parse_matches_reference_parser(seed):
    rng = np.random.default_rng(seed)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "adj_sparse.mtx")
        open(path, "w").write(sample_graph_text(rng))
        out = os.path.join(d, "ref.txt")
        subprocess.run([REF_INC, "graph", path, out], check=True, capture_output=True)
        ref = open(out).read().split("\n")
        Vr, Lr = (int(x) for x in ref[0].split())
        V, begin, adj, w = po.read_graph0(path)
        assert V == Vr
        for v in range(V):
            toks = ref[1 + v].split()
            assert int(toks[0]) == begin[v + 1] - begin[v]
            slots = [t.split(":") for t in toks[1:]]
            assert [int(a) for a, _ in slots] == adj[begin[v]:begin[v + 1]].tolist()
            assert [np.float32(float(x)) for _, x in slots] == w[begin[v]:begin[v + 1]].tolist()
        Vp, rp, col, pw = product_graph(path, 0)
        assert Vp == V and np.array_equal(rp, begin) and np.array_equal(col, adj)
        assert np.array_equal(pw.view(np.uint32), w.view(np.uint32))


@need_ref
@pytest.mark.parametrize("m", [1, 2, 5, 11, 12, 19, 2000, 2005])
def test_target_layout_matches_reference_parser(m):
    rng = np.random.default_rng(m)
    with tempfile.TemporaryDirectory() as d:
        ext = os.path.join(d, "ext.mtx")
        u = rng.integers(1, 500, m)
        v = rng.integers(1, 500, m)
        w = rng.uniform(0.1, 10, m).astype(np.float32)
        open(ext, "w").write("".join(f"{a} {b} {x:.9g}\n" for a, b, x in zip(u, v, w)))
        out = os.path.join(d, "ref.txt")
        subprocess.run([REF_INC, "targets", ext, out], check=True, capture_output=True)
        ref = open(out).read().split()
        assert int(ref[0]) == m
        rnodes = [int(x) for x in ref[1:1 + 2 * m]]
        rw = [np.float32(float(x)) for x in ref[1 + 2 * m:1 + 3 * m]]
        nodes, ow = po.read_targets(ext, True, 10)
        assert nodes.tolist() == rnodes and ow.tolist() == rw
        # the layout decodes to file order batch by batch
        st = po.batches(m, 10)
        for b0, b1 in zip(st[:-1], st[1:]):
            n = b1 - b0
            assert nodes[2 * b0:2 * b0 + n].tolist() == (u[b0:b1] - 1).tolist()
            assert nodes[2 * b0 + n:2 * b0 + 2 * n].tolist() == (v[b0:b1] - 1).tolist()
        # del.mtx: unweighted, same layout (decremental read_targets)
        dl = os.path.join(d, "del.mtx")
        open(dl, "w").write("".join(f"{a} {b}\n" for a, b in zip(u, v)))
        subprocess.run([REF_DEC, "targets", dl, out], check=True, capture_output=True)
        ref = [int(x) for x in open(out).read().split()]
        dn, _ = po.read_targets(dl, False, 10)
        assert ref[0] == m and dn.tolist() == ref[1:]


def test_product_edge_list_is_file_order():
    lib = L.lib()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ext.mtx")
        open(p, "w").write("3 4 1.5\n1\t2\t0.25\n10  9  7\n")
        h = C.c_void_p()
        L.check(lib.dyg_mtx_read_edges(p.encode(), 1, C.byref(h)))
        n = lib.dyg_mtx_nlines(h)
        u = np.ctypeslib.as_array(lib.dyg_mtx_line_u(h), (n,)).tolist()
        v = np.ctypeslib.as_array(lib.dyg_mtx_line_v(h), (n,)).tolist()
        w = np.ctypeslib.as_array(lib.dyg_mtx_line_w(h), (n,)).tolist()
        lib.dyg_mtx_free(h)
        assert (u, v, w) == ([2, 0, 9], [3, 1, 8], [1.5, 0.25, 7.0])


def test_product_parses_one_based_dense_like_oracle():
    rng = np.random.default_rng(3)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "dense.mtx")
        txt = sample_graph_text(rng, selfloop=False, dup=False)
        # shift to 1-based
        lines = []
        for ln in txt.strip().split("\n"):
            a, b, w = ln.split()
            lines.append(f"{int(a) + 1} {int(b) + 1} {w}")
        open(path, "w").write("\n".join(lines) + "\n")
        od = po.Dense(path)
        rp, col, w = od.graph.export()
        Vp, prp, pcol, pw = product_graph(path, 1)
        assert Vp == od.graph.V
        assert np.array_equal(prp, rp) and np.array_equal(pcol, col)
        assert np.array_equal(pw.view(np.uint32), w.view(np.uint32))


def test_weight_text_matches_ostream():
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        np.exp(rng.uniform(np.log(1e-9), np.log(1e9), 3000)),
        rng.uniform(0.1, 10, 2000),
        [0.1, 0.5, 1, 2, 10, 100, 1e6, 123456, 1234567, 999999.5, 1e-5, 3.4e38, 1e-40],
    ]).astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "o.mtx")
        n = len(vals)
        u = np.zeros(n, np.uint32)
        L.check(L.lib().dyg_mtx_write_weighted(p.encode(), L.ptr(u, L.u32p), L.ptr(u, L.u32p),
                                               L.ptr(vals, L.f32p), n, 0))
        got = [ln.split()[2] for ln in open(p).read().strip().split("\n")]
    assert got == [po.format_float(float(x)) for x in vals]


def test_abi_exports_every_declared_symbol():
    import re

    hdr = open(os.path.join(os.path.dirname(L.LIB_PATH), "..", "..", "include",
                            "dygrass.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(dyg_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 30
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_context_without_device_fails_cleanly():
    lib = L.lib()
    h = C.c_void_p()
    st = lib.dyg_ctx_create(0, C.byref(h))
    if st == 0:
        lib.dyg_ctx_destroy(h)
    else:
        assert st == -2 and lib.dyg_last_error(None)
