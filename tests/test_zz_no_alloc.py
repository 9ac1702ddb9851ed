"""The no-allocation insertion walk (dyg_walk_incremental_noalloc) against
the allocating entry point it mirrors (dyg_walk_incremental), through the C
ABI: two calls in a row, the second with a smaller batch (its staging array
keeps the first call's contents), then a batch over the array's limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import csr, f32bits

pytestmark = pytest.mark.gpu


def _call(H, fn, src, dst, w, eid, p):
    from dygrass_amd import _lib as L
    from dygrass_amd._lib import check, ptr

    n = len(src)
    src, dst = np.ascontiguousarray(src, np.uint32), np.ascontiguousarray(dst, np.uint32)
    w, eid = np.ascontiguousarray(w, np.float32), np.ascontiguousarray(eid, np.uint64)
    f = np.full(max(n, 1), 7, np.int32)
    c = np.zeros(max(n, 1), np.float32)
    win = np.zeros(max(n, 1), np.uint32)
    hops = np.zeros(max(n, 1), np.uint32)
    st = L.WalkStats()
    check(getattr(H._lib, fn)(H.handle, ptr(src, L.u32p), ptr(dst, L.u32p), ptr(w, L.f32p),
                              ptr(eid, L.u64p), n, C.byref(p), ptr(f, L.i32p), ptr(c, L.f32p),
                              ptr(win, L.u32p), ptr(hops, L.u32p), C.byref(st)), H.ctx.handle)
    return f[:n], c[:n], win[:n], hops[:n], (st.hops, st.redirects, st.leaf_stops)


def test_noalloc_walk_matches_allocating_walk(ctx, grid_wl):
    from dygrass_amd import DeviceGraph, params

    wl = grid_wl
    rp, col, w = csr(wl.H)
    H = DeviceGraph(ctx, wl.H.V, rp, col, w)
    ins = wl.ins
    # pruned walks (the default): decisions, conductances and winners are
    # exact; how far the losing walkers get before they stop depends on timing,
    # so hop counts are compared on the unpruned walks only
    for prune in (True, False):
        p = params(512, 100, 10.0, seed=3, prune=prune)
        for lo, hi in ((0, 2000), (1200, 1500)):  # second call smaller
            args = (ins.u[lo:hi], ins.v[lo:hi], ins.w[lo:hi],
                    np.arange(lo, hi, dtype=np.uint64), p)
            want = _call(H, "dyg_walk_incremental", *args)
            got = _call(H, "dyg_walk_incremental_noalloc", *args)
            assert np.array_equal(got[0], want[0])
            assert np.array_equal(f32bits(got[1]), f32bits(want[1]))
            assert np.array_equal(got[2], want[2])
            if not prune:
                assert np.array_equal(got[3], want[3]) and got[4] == want[4]
            assert got[4][0] > 0
            assert np.any(got[0] == -1) and np.any(got[0] > 0)


def test_noalloc_walk_rejects_batches_over_its_limit(ctx, grid_wl):
    from dygrass_amd import DeviceGraph, DygError, params
    from dygrass_amd import _lib as L

    wl = grid_wl
    rp, col, w = csr(wl.H)
    H = DeviceGraph(ctx, wl.H.V, rp, col, w)
    one = np.zeros(1, np.uint32)
    p = params(64, 100, 10.0)
    n = L.DYG_NOALLOC_MAX_EDGES + 1
    with pytest.raises(DygError) as e:  # refused before any array is read
        _over(H, n, one, p)
    assert "DYG_NOALLOC_MAX_EDGES" in str(e.value) and e.value.status == -1


def _over(H, n, one, p):
    from dygrass_amd import _lib as L
    from dygrass_amd._lib import check, ptr

    f = np.zeros(1, np.int32)
    c = np.zeros(1, np.float32)
    w = np.ones(1, np.float32)
    check(H._lib.dyg_walk_incremental_noalloc(H.handle, ptr(one, L.u32p), ptr(one, L.u32p),
                                              ptr(w, L.f32p), None, n, C.byref(p),
                                              ptr(f, L.i32p), ptr(c, L.f32p), None, None, None),
          H.ctx.handle)
