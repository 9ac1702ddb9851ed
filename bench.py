#!/usr/bin/env python3
"""Benchmark of the insertion updater on one MI355X.

Default workload (the plain-at-scale configuration): a 1M-point 8-NN random
geometric graph (weights 1/distance, ~4.6M edges), sparsifier = maximum-weight
spanning tree + 2% off-tree edges, 200k insertions between vertices 2..5 hops
apart, processed as 10 snapshot batches of 20k; threshold 100, 256 walkers,
100 steps, seed 0.

A step is one whole run over the 200k insertions starting from the original
sparsifier (the sparsifier is restored before each step, outside the timed
window).  Each timed step is bracketed by device synchronisation; inputs are
resident in HBM when a step starts.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--mode plain|device] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "edge updates/s + p50/p99 request latency, served vs plain, at equal kappa"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats: dict, updates: int, walkers: int) -> float:
    """Bytes the walks must read.  Per walker the source's row record (8 B:
    first slot + degree); per neighbour draw one 16 B slot (neighbour 4 +
    weight 4 + the neighbour's first slot 4 + degree 4, i.e. the row-start,
    degree, column and weight words of a hop); +16 B for each
    non-backtracking redirect to the row's last slot.  A leaf is recognised
    from the degree already read.  Per edge 12 B of input and 8 B of output."""
    return (8.0 * updates * walkers + 16.0 * stats["hops"] + 16.0 * stats["redirects"]
            + 20.0 * updates)


def run_stream(ctx, wl, plain: bool, seed: int, block_edges: int = 0):
    """One pass of the whole batch list through a fresh stream; the stream is
    built (graphs uploaded, buffers reserved) outside the timed window."""
    from dygrass_amd import Stream, stream_params

    maxb = max(len(b[1]) for b in wl.batches)
    sp = stream_params(w_ins=wl.w_ins, w_del=wl.w_del, threshold=wl.threshold, seed=seed,
                       max_batch=maxb, max_inserts=wl.updates, plain=plain,
                       block_edges=block_edges)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp)
    ctx.sync()
    outs, lat, walk_ms, hops, agg = [], [], 0.0, 0, {}
    ts = time.perf_counter()
    for op, u, v, w in wl.batches:
        t0 = time.perf_counter()
        o, stt = st.apply(op, u, v, w)
        lat.append(time.perf_counter() - t0)
        outs.append(o)
        walk_ms += stt["walk_ms"]
        hops += stt["hops"]
        for k in ("weight_updates", "found", "added", "skipped", "g_only", "recovered",
                  "heuristic", "invalid"):
            agg[k] = agg.get(k, 0) + stt[k]
    ctx.sync()
    dt = time.perf_counter() - ts
    _, nh, _ = st.H.info()
    st.close()
    return dict(seconds=dt, outcomes=np.concatenate(outs), latencies=np.array(lat),
                walk_ms=walk_ms, hops=hops, counts=agg, h_slots=nh)


def leg_batched(args) -> dict:
    """Mixed insertion/deletion stream (64 uneven batches) on the 4M-point
    12-NN graph: device-resident form vs the plain form, same results."""
    from dygrass_amd import Context, synth

    t0 = time.time()
    wl = synth.config_batched(args.scale, args.seed)
    m = wl.updates
    log(f"workload {wl.name}: V={wl.G.V} E_G={len(wl.G)} E_H={len(wl.H)} batches="
        f"{len(wl.batches)} updates={m} ({time.time() - t0:.1f}s)")
    ctx = Context(0)
    for i in range(args.warmup):
        r = run_stream(ctx, wl, False, args.seed, args.block_edges)
        log(f"batched warmup {i}: {r['seconds'] * 1e3:.1f} ms")
    runs = []
    for i in range(args.steps):
        r = run_stream(ctx, wl, False, args.seed, args.block_edges)
        runs.append(r)
        log(f"batched step {i}: {r['seconds'] * 1e3:.1f} ms (walk kernels {r['walk_ms']:.1f} ms)")
    total = sum(r["seconds"] for r in runs)
    head = runs[-1]
    out = {
        "value": m * len(runs) / total,
        "ms_per_step": total / len(runs) * 1e3,
        "workload": f"{wl.name}: {len(wl.batches)} batches of log-uniform size 1..50000 "
                    f"({m} updates, 70% insertion batches), threshold {wl.threshold:g}, "
                    f"{wl.w_ins} insertion / {wl.w_del} deletion walkers, 100 steps "
                    "(device-resident stream: work-queue walk kernels, device edits)",
        "V": int(wl.G.V), "E_G": int(len(wl.G)), "E_H": int(len(wl.H)), "updates": int(m),
        "batches": len(wl.batches),
        "walk_kernel_ms_per_step": head["walk_ms"],
        "walk_steps_per_s": head["hops"] / head["seconds"],
        "batch_latency_ms": {"p50": float(np.percentile(head["latencies"], 50) * 1e3),
                             "p99": float(np.percentile(head["latencies"], 99) * 1e3)},
        "outcomes": head["counts"],
    }
    if not args.no_plain:
        pr = run_stream(ctx, wl, True, args.seed)
        log(f"batched plain: {pr['seconds'] * 1e3:.1f} ms (walk kernels {pr['walk_ms']:.1f} ms)")
        out["plain"] = {"value": m / pr["seconds"], "ms_per_step": pr["seconds"] * 1e3,
                        "walk_kernel_ms_per_step": pr["walk_ms"],
                        "batch_latency_ms": {
                            "p50": float(np.percentile(pr["latencies"], 50) * 1e3),
                            "p99": float(np.percentile(pr["latencies"], 99) * 1e3)},
                        "shape": "host edge maps and edit plans, one workgroup per edge, "
                                 "a wait per batch"}
        out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
        out["plain_outcomes_identical"] = bool(np.array_equal(pr["outcomes"], head["outcomes"]))
        out["plain_h_slots_identical"] = pr["h_slots"] == head["h_slots"]
    ctx.close()
    return out


def leg_served(args) -> dict:
    """Request loop on the 8M-vertex 3D grid + shortcuts: Poisson arrivals,
    admission by memory into slots, p50/p99 latency.  The plain form serves
    the first requests of the same trace one at a time (FIFO)."""
    import threading

    from dygrass_amd import Context, Server, Stream, stream_params, synth

    t0 = time.time()
    wl = synth.config_served(args.scale, args.seed, args.requests, args.rate)
    m = wl.updates
    log(f"workload {wl.name}: V={wl.G.V} E_G={len(wl.G)} E_H={len(wl.H)} requests="
        f"{len(wl.requests)} updates={m} rate={args.rate:g}/s ({time.time() - t0:.1f}s)")
    ctx = Context(0)
    maxr = max(len(r[1]) for r in wl.requests)

    def params(plain):
        return stream_params(w_ins=wl.w_ins, w_del=wl.w_del, threshold=wl.threshold,
                             seed=args.seed, max_batch=maxr, max_inserts=m, plain=plain)

    g_edges, h_edges = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)
    st = Stream(ctx, wl.G.V, g_edges, h_edges, params(False))
    srv = Server(st, slots=args.slots, max_request=maxr)
    n = len(wl.requests)
    tickets, lat, done_at = [], np.zeros(n), np.zeros(n)

    def collect():
        for i in range(n):
            while len(tickets) <= i:
                time.sleep(0.0002)
            _, res = srv.wait(tickets[i], len(wl.requests[i][1]))
            lat[i] = res["latency_ms"]
            done_at[i] = time.perf_counter()

    th = threading.Thread(target=collect)
    th.start()
    start = time.perf_counter()
    for i, (op, u, v, w) in enumerate(wl.requests):
        delay = start + wl.arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(op, u, v, w))
    th.join()
    span = done_at[-1] - start
    srv.close()
    st.close()
    sizes = np.array([len(r[1]) for r in wl.requests])
    small = sizes <= 1000
    out = {
        "value": m / span,
        "workload": f"{wl.name}: {n} requests, Poisson arrivals at {args.rate:g}/s, sizes "
                    f"log-uniform 1..10000 ({m} updates, 60% insertions), threshold "
                    f"{wl.threshold:g}, {wl.w_ins}/{wl.w_del} walkers, {args.slots} slots",
        "V": int(wl.G.V), "E_G": int(len(wl.G)), "E_H": int(len(wl.H)),
        "requests": n, "updates": int(m), "span_s": span,
        "latency_ms": {"p50": float(np.percentile(lat, 50)),
                       "p99": float(np.percentile(lat, 99))},
        "latency_ms_le_1k": {"p50": float(np.percentile(lat[small], 50)),
                             "p99": float(np.percentile(lat[small], 99)),
                             "requests": int(small.sum())},
    }
    log(f"served: {out['value']:.0f} updates/s, p50 {out['latency_ms']['p50']:.2f} ms, "
        f"p99 {out['latency_ms']['p99']:.2f} ms (<=1k: p99 "
        f"{out['latency_ms_le_1k']['p99']:.2f} ms)")
    if not args.no_plain and args.plain_requests > 0:
        k = min(n, args.plain_requests)
        pst = Stream(ctx, wl.G.V, g_edges, h_edges, params(True))
        plat, busy = [], 0.0
        start = time.perf_counter()
        for i in range(k):
            arr = start + wl.arrivals[i]
            now = time.perf_counter()
            if now < arr:
                time.sleep(arr - now)
            op, u, v, w = wl.requests[i]
            t1 = time.perf_counter()
            pst.apply(op, u, v, w)
            t2 = time.perf_counter()
            busy += t2 - t1
            plat.append((t2 - arr) * 1e3)
        pst.close()
        pm = int(sizes[:k].sum())
        out["plain"] = {"value": pm / busy, "requests": k, "updates": pm,
                        "latency_ms": {"p50": float(np.percentile(plat, 50)),
                                       "p99": float(np.percentile(plat, 99))},
                        "shape": "requests applied one at a time in arrival order, host edge "
                                 "maps and edit plans, one workgroup per edge",
                        "note": "value = updates / busy time over the first requests of the "
                                "same trace"}
        out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
        log(f"served plain ({k} requests): {out['plain']['value']:.0f} updates/s busy, "
            f"p99 {out['plain']['latency_ms']['p99']:.1f} ms")
    ctx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="workload scale (tests only)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mode", choices=["device", "plain"], default="device")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--block-edges", type=int, default=0,
                    help="edges per workgroup of the work-queue walk kernel (0 = automatic)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--workload", choices=["incremental", "batched", "served"],
                    default="incremental")
    ap.add_argument("--requests", type=int, default=1500, help="served: requests in the trace")
    ap.add_argument("--rate", type=float, default=1000.0, help="served: arrivals per second")
    ap.add_argument("--slots", type=int, default=8, help="served: requests admitted at once")
    ap.add_argument("--plain-requests", type=int, default=60,
                    help="served: requests of the trace the plain form serves")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (a known-size "
                         "dispatch for scaling hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); a multi-GPU run would "
              "replicate the graphs and split each snapshot batch's edges across GPUs, which "
              "this build does not implement", file=sys.stderr)
        return 2

    if args.workload == "served":
        b = leg_served(args)
        print(json.dumps({"metric": METRIC, "value": b["value"], "unit": "edge-updates/s",
                          "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": b["span_s"] * 1e3,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "f32", "data": "synthetic",
                          "config": {"workload": b.pop("workload")}, "served": b}), flush=True)
        return 0
    if args.workload == "batched":
        b = leg_batched(args)
        print(json.dumps({"metric": METRIC, "value": b["value"], "unit": "edge-updates/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": b["ms_per_step"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                          "data": "synthetic", "config": {"workload": b.pop("workload")},
                          "batched": b}), flush=True)
        return 0

    from dygrass_amd import (Context, DeviceArray, DeviceGraph, csr_from_edges,
                             incremental_update, incremental_update_dev, params)
    from dygrass_amd import synth

    t0 = time.time()
    wl = synth.config_rgg_insert(args.scale, args.seed)
    log(f"workload {wl.name}: V={wl.G.V} E_G={len(wl.G)} E_H={len(wl.H)} "
        f"updates={len(wl.ins)} ({time.time() - t0:.1f}s)")
    rp, col, w = csr_from_edges(wl.H.V, wl.H.u, wl.H.v, wl.H.w)
    ctx = Context(0)
    m = len(wl.ins)

    def run_mode(mode: str, steps: int, warmup: int) -> dict:
        plain = mode == "plain"
        p = params(wl.walkers, 100, wl.threshold, seed=args.seed, per_edge=plain,
                   host_edits=plain, block_edges=args.block_edges)
        I = (m + 9) // 10
        if not plain:
            # inputs resident in HBM before the timed window
            t_src = DeviceArray.from_numpy(ctx, wl.ins.u)
            t_dst = DeviceArray.from_numpy(ctx, wl.ins.v)
            t_w = DeviceArray.from_numpy(ctx, wl.ins.w)
            t_found = DeviceArray(ctx, m, np.int32)
            t_cond = DeviceArray(ctx, m, np.float32)

        def one_step():
            H = DeviceGraph(ctx, wl.H.V, rp, col, w, slack=2)  # restore (untimed)
            if not plain:
                H.reserve(m, I)
            ctx.sync()
            ts = time.perf_counter()
            if plain:
                found, st = incremental_update(H, wl.ins.u, wl.ins.v, wl.ins.w, p)
            else:
                st = incremental_update_dev(H, t_src.ptr, t_dst.ptr, t_w.ptr, m, t_found.ptr,
                                            t_cond.ptr, p)
            ctx.sync()
            te = time.perf_counter()
            if not plain:
                found = t_found.numpy()
            H.close()
            return te - ts, found, st

        for i in range(warmup):
            dt, _, _ = one_step()
            log(f"{mode} warmup {i}: {dt * 1e3:.2f} ms")
        times, walk_ms, upd_ms, st, found = [], [], [], None, None
        for i in range(steps):
            dt, found, st = one_step()
            times.append(dt)
            walk_ms.append(st["walk_ms"])
            upd_ms.append(st["update_ms"])
            log(f"{mode} step {i}: {dt * 1e3:.2f} ms (walk kernels {st['walk_ms']:.2f} ms, "
                f"graph edits {st['update_ms']:.2f} ms, added {st['added']})")
        total = float(np.sum(times))
        return dict(found=found, stats=st, total=total, steps=steps,
                    value=m * steps / total, ms_per_step=total / steps * 1e3,
                    walk_ms=float(np.mean(walk_ms)), edit_ms=float(np.mean(upd_ms)))

    if args.copy_calibration_mb:
        nb = args.copy_calibration_mb << 20
        a_src = DeviceArray(ctx, nb // 4, np.uint32)
        a_dst = DeviceArray(ctx, nb // 4, np.uint32)
        import ctypes

        ctx._lib.dyg_copy_on_device(ctx.handle, ctypes.c_void_p(a_dst.ptr),
                                    ctypes.c_void_p(a_src.ptr), nb)
        log(f"calibration copy: {args.copy_calibration_mb} MiB device to device")
        a_src.close()
        a_dst.close()
    plain = (run_mode("plain", max(1, min(args.steps, 3)), 1)
             if args.mode == "device" and args.steps > 1 else None)
    head = run_mode(args.mode, args.steps, args.warmup)
    stats = head["stats"]
    found = head["found"]
    launches = int(stats["batches"])
    bytes_run = algorithmic_bytes(stats, m, wl.walkers)
    walk_s = head["walk_ms"] / 1e3
    achieved = bytes_run / walk_s / 1e9 if walk_s > 0 else 0.0
    kernel = (f"k_walk_pool<W={wl.walkers}, insert, prune>" if args.mode == "device"
              else f"k_walk<W={wl.walkers}, insert, prune>")
    shape = ("work-queue walk kernel, graph edits on the device, no host round trip between "
             "batches" if args.mode == "device"
             else "one workgroup per edge, graph edits planned on the host between batches")

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "edge-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{wl.name}: incremental update, {launches} snapshot batches of "
                        f"{(m + 9) // 10} insertions, threshold {wl.threshold:g}, "
                        f"{wl.walkers} walkers, 100 steps ({shape})",
            "V": int(wl.G.V), "E_G": int(len(wl.G)), "E_H": int(len(wl.H)), "updates": int(m),
            "walkers": int(wl.walkers), "threshold": float(wl.threshold), "mode": args.mode,
            "seed": args.seed, "scale": args.scale,
        },
        "walk_steps_per_s": stats["hops"] * head["steps"] / head["total"],
        "no_path_fraction": float(stats["added"]) / m,
        "walk_kernel_ms_per_step": head["walk_ms"],
        "graph_edit_ms_per_step": head["edit_ms"],
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": kernel,
            "bytes_per_launch": bytes_run / launches,
            "avg_launch_ms": head["walk_ms"] / launches,
        },
        "cpu_baseline": None,
    }
    if plain is not None:
        out["plain"] = {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                        "walk_kernel_ms_per_step": plain["walk_ms"],
                        "graph_edit_ms_per_step": plain["edit_ms"],
                        "steps": plain["steps"],
                        "shape": "one workgroup per edge, graph edits planned on the host"}
        out["speedup_vs_plain"] = head["value"] / plain["value"]
        out["plain_decisions_identical"] = bool(np.array_equal(plain["found"], found))

    if not args.no_cpu:
        from oracle import pyoracle as po

        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        og = po.Graph(wl.H.V, rp, col, w)
        ts = time.perf_counter()
        cfound, cst = og.incremental(wl.ins.u, wl.ins.v, wl.ins.w, wl.threshold, wl.walkers,
                                     seed=args.seed, threads=threads)
        cdt = time.perf_counter() - ts
        out["cpu_baseline"] = {
            "value": m / cdt,
            "unit": "edge-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"the whole workload ({m} insertions, {int(cst.n_batches)} snapshot "
                      f"batches) through the oracle, OpenMP over each batch's edges",
            "walk_seconds": cst.walk_seconds,
        }
        out["parity_with_cpu"] = bool(np.array_equal(found, cfound))
        log(f"cpu baseline: {m / cdt:.0f} edge-updates/s on {threads} threads; "
            f"decisions identical: {out['parity_with_cpu']}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
