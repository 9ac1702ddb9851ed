import os
import sys
import threading

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    from dygrass_amd import Context

    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def grid_wl():
    from dygrass_amd import synth

    return synth.config_grid_insert(1.0, 0)


@pytest.fixture(scope="session")
def mesh_wl():
    from dygrass_amd import synth

    return synth.config_mesh_delete(0.004, 0)


def csr(E):
    from dygrass_amd import csr_from_edges

    return csr_from_edges(E.V, E.u, E.v, E.w)


def f32bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def serve_all(srv, reqs, timeout=300):
    """Submit every request from this thread while a second thread collects
    them in ticket order; returns the outcome codes and results per request."""
    tickets, got, res = [], [None] * len(reqs), [None] * len(reqs)

    def collect():
        for i in range(len(reqs)):
            while len(tickets) <= i:
                threading.Event().wait(0.0005)
            got[i], res[i] = srv.wait(tickets[i], len(reqs[i][1]))

    th = threading.Thread(target=collect)
    th.start()
    for op, u, v, w in reqs:
        tickets.append(srv.submit(op, u, v, w))
    th.join(timeout=timeout)
    assert not th.is_alive()
    assert tickets == sorted(tickets)
    return got, res


def mixed_batches(wl, seed, n_ins=300, n_del=200):
    """Insertions (new edges, existing G edges, in-batch repeats, a self-loop)
    and deletions (G edges, H edges, non-edges, repeats), alternating."""
    rng = np.random.default_rng(seed)
    G = wl.G
    out = []
    k = 0
    for b in range(4):
        if b % 2 == 0:
            u = list(wl.ins.u[k:k + n_ins])
            v = list(wl.ins.v[k:k + n_ins])
            w = list(wl.ins.w[k:k + n_ins])
            k += n_ins
            ex = rng.choice(len(G), 25, replace=False)  # existing edges: weight updates
            u += list(G.v[ex]); v += list(G.u[ex]); w += list(rng.uniform(0.1, 2, 25))
            rep = rng.choice(n_ins, 6, replace=False)  # repeats inside the batch
            u += [u[i] for i in rep]; v += [v[i] for i in rep]; w += list(rng.uniform(0.1, 2, 6))
            u.append(5); v.append(5); w.append(1.0)  # self-loop
            perm = rng.permutation(len(u))
            out.append(("ins", np.array(u)[perm], np.array(v)[perm],
                        np.array(w, np.float32)[perm]))
        else:
            idx = rng.choice(len(G), n_del, replace=False)
            hidx = rng.choice(len(wl.H), 30, replace=False)
            u = list(G.u[idx]) + list(wl.H.u[hidx]) + list(rng.integers(0, G.V, 10))
            v = list(G.v[idx]) + list(wl.H.v[hidx]) + list(rng.integers(0, G.V, 10))
            u += u[:5]; v += v[:5]  # repeats
            # a previously inserted edge, deleted again
            u.append(int(wl.ins.u[0])); v.append(int(wl.ins.v[0]))
            perm = rng.permutation(len(u))
            out.append(("del", np.array(u)[perm], np.array(v)[perm], None))
    return out
