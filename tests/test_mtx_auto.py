"""Tolerant .mtx reader: the loader cases of the reference's dynamic_update
tests (dynamic_update/test.cpp:53-178 inputs, its test_matrix.mtx) with the
expected graphs worked out by hand (CPU only)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "standard": ("%%MatrixMarket matrix coordinate real general\n% This is a test adjacency "
                 "matrix\n% 4 vertices, 6 edges\n4 4 6\n1 2 0.5\n1 3 1.0\n2 3 0.8\n2 4 1.2\n"
                 "3 4 0.9\n4 1 0.7\n",
                 dict(has_header=1, has_dimensions=1, weighted=1, laplacian=0, triangular=1,
                      base=1), 4,
                 [(0, 1, 0.5), (0, 2, 1.0), (1, 2, 0.8), (1, 3, 1.2), (2, 3, 0.9), (3, 0, 0.7)]),
    "triangular": ("%%MatrixMarket matrix coordinate real symmetric\n% Symmetric matrix - only "
                   "upper triangular stored\n4 4 4\n1 2 0.5\n1 3 1.0\n2 3 0.8\n3 4 0.9\n",
                   dict(has_header=1, has_dimensions=1, weighted=1, triangular=1, base=1), 4,
                   [(0, 1, 0.5), (0, 2, 1.0), (1, 2, 0.8), (2, 3, 0.9)]),
    "no_dimensions": ("%%MatrixMarket matrix coordinate real general\n% No dimensions line\n"
                      "1 2 0.5\n2 3 1.0\n3 4 0.8\n4 1 1.2\n",
                      dict(has_header=1, has_dimensions=0, weighted=1, base=1), 4,
                      [(0, 1, 0.5), (1, 2, 1.0), (2, 3, 0.8), (3, 0, 1.2)]),
    "dimensions_only": ("4 4 4\n1 2 0.5\n2 3 1.0\n3 4 0.8\n4 1 1.2\n",
                        dict(has_header=0, has_dimensions=1, base=1), 4,
                        [(0, 1, 0.5), (1, 2, 1.0), (2, 3, 0.8), (3, 0, 1.2)]),
    "raw": ("1 2 0.5\n2 3 1.0\n3 4 0.8\n4 1 1.2\n",
            dict(has_header=0, has_dimensions=0, base=1), 4,
            [(0, 1, 0.5), (1, 2, 1.0), (2, 3, 0.8), (3, 0, 1.2)]),
    "zero_based": ("0 1 0.5\n1 2 1.0\n2 3 0.8\n3 0 1.2\n",
                   dict(has_dimensions=0, base=0), 4,
                   [(0, 1, 0.5), (1, 2, 1.0), (2, 3, 0.8), (3, 0, 1.2)]),
    "unweighted": ("%%MatrixMarket matrix coordinate pattern general\n4 4 4\n1 2\n2 3\n3 4\n"
                   "4 1\n", dict(has_header=1, has_dimensions=1, weighted=0, base=1), 4,
                   [(0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.0), (3, 0, 1.0)]),
    "laplacian": ("%%MatrixMarket matrix coordinate real general\n4 4 10\n1 1 2.0\n1 2 -0.5\n"
                  "1 4 -1.5\n2 2 1.3\n2 1 -0.5\n2 3 -0.8\n3 3 1.7\n3 2 -0.8\n3 4 -0.9\n"
                  "4 4 3.4\n", dict(has_dimensions=1, weighted=1, laplacian=1, triangular=0,
                                    base=1), 4,
                  [(0, 1, 0.5), (0, 3, 1.5), (1, 2, 0.8), (2, 3, 0.9)]),
    "single_edge": ("2 2 1\n1 2 1.0\n", dict(has_dimensions=1, base=1), 2, [(0, 1, 1.0)]),
    "large_ids": ("1000 1000 3\n1 500 0.5\n500 1000 1.0\n1000 1 0.8\n",
                  dict(has_dimensions=1, base=1), 1000,
                  [(0, 499, 0.5), (499, 999, 1.0), (999, 0, 0.8)]),
    "full_adjacency": ("%%MatrixMarket matrix coordinate real general\n3 3 6\n1 2 1.0\n"
                       "1 3 1.0\n2 1 1.0\n2 3 1.0\n3 1 1.0\n3 2 1.0\n",
                       dict(has_dimensions=1, triangular=0, base=1), 3,
                       [(0, 1, 1.0), (0, 2, 1.0), (1, 2, 1.0)]),
    "full_laplacian": ("%%MatrixMarket matrix coordinate real general\n3 3 3\n1 1 2.0\n"
                       "1 2 -1.0\n1 3 -1.0\n2 1 -1.0\n2 2 2.0\n2 3 -1.0\n3 1 -1.0\n3 2 -1.0\n"
                       "3 3 2.0\n", dict(has_dimensions=1, laplacian=1, triangular=0, base=1),
                       3, [(0, 1, 1.0), (0, 2, 1.0), (1, 2, 1.0)]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_loader_cases(tmp_path, name):
    from dygrass_amd import read_mtx_auto

    text, want_info, want_V, want_edges = CASES[name]
    f = tmp_path / f"{name}.mtx"
    f.write_text(text)
    info, V, u, v, w = read_mtx_auto(str(f))
    for k, x in want_info.items():
        assert info[k] == x, (k, info)
    assert V == want_V
    got = list(zip(u.tolist(), v.tolist(), np.round(w.astype(np.float64), 6).tolist()))
    assert got == want_edges


def test_reference_fixture_matrix():
    """dynamic_update/test_matrix.mtx (a 3-cycle behind a 13-line banner)."""
    from dygrass_amd import read_mtx_auto

    info, V, u, v, w = read_mtx_auto(os.path.join(HERE, "golden", "dyn_test_matrix.mtx"))
    assert info["has_header"] and info["has_dimensions"] and info["triangular"]
    assert V == 3 and info["entries"] == 3
    assert list(zip(u.tolist(), v.tolist())) == [(0, 1), (1, 2), (2, 0)]
    assert np.all(w == 1.0)


def test_bad_inputs(tmp_path):
    from dygrass_amd import DygError, read_mtx_auto

    for text in ("1 2 0.5\n% late comment\n2 3 1\n", "2 3 0.5\n3 4 1\n", "1 x 2\n", ""):
        f = tmp_path / "bad.mtx"
        f.write_text(text)
        with pytest.raises(DygError):
            read_mtx_auto(str(f))
