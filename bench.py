#!/usr/bin/env python3
"""Benchmark of the random-walk sparsifier updaters on one MI355X.

Five workloads, all seeded synthetic graphs, every result checked against a
second implementation in the same run:

batched (the headline `value`): a 4M-point 12-NN random geometric graph
  (w = 1/distance, ~27M edges) and its sparsifier (maximum-weight spanning
  tree + 2%); 64 batches of log-uniform size 1..50000 (~380k updates), 70%
  insertion batches, 30% deletion batches, applied as one stream with
  per-batch snapshot semantics; walk settings from the committed autotuning.
  A step is the whole stream from the original graphs (rebuilt outside the
  timed window), every batch already in HBM and the list applied in one call
  (dyg_stream_apply_many_dev); the same stream from host arrays, batch by
  batch (PCIe inside the window), is reported beside it.
  Compared with the plain form (host edge maps and edit plans, one
  workgroup per edge, a wait per batch) on the same batches.
served: an 8M-vertex 3D grid plus 16M shortcuts (~40M edges) and a ~10M-edge
  sparsifier; 1500 requests with Poisson arrivals (1000/s), sizes log-uniform
  1..10000, 60/40 insert/delete; admission into slots by memory; p50/p99
  latency, plus the capacity with every request arriving at once.
incremental: a 1M-point 8-NN random geometric graph, 200k insertions as 10
  snapshot batches (the plain-at-scale configuration); the walk kernel's
  roofline and the FP64 kappa check come from this leg.
decremental: a 2M-node triangulated mesh, 100k deletions as 10 snapshot
  batches (the plain-decremental configuration); deletion-only stream (the
  ten batches in one call) vs the dual graph.
grid: the 128x128 grid with 2,000 insertions in 10 snapshot batches,
  threshold 10, 512 walkers (the plain incremental configuration that also
  runs through the CPU walk path), GPU beside the oracle's whole run.

CPU baselines (the oracle, OpenMP, on the cores this process may use, in a
child process with its threads bound to neighbouring cores) run after every
GPU leg.

  python bench.py [--steps K] [--warmup W]
                  [--workload all|batched|served|incremental|decremental|grid]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "edge updates/s + p50/p99 request latency, served vs plain, at equal kappa"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
OUT_FOUND, OUT_ADDED, OUT_RECOVERED, OUT_HEURISTIC = 2, 3, 6, 7  # DYG_OUT_* (dygrass.h)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(a, q) -> float:
    return float(np.percentile(np.asarray(a), q)) if len(a) else 0.0


def algorithmic_bytes(hops: int, redirects: int, edges: int, per_edge: float = 20.0) -> float:
    """Bytes a walk launch must move: 16 B per neighbour draw, the first
    included (row start 4 + degree 4 of the row drawn from, column 4 + weight 4
    of the drawn slot); 8 B per non-backtracking redirect (column + weight of
    the row's last slot); per edge 12 B in and 8 B out for insertions (16 B
    for deletions, whose winner path of 4 B per vertex is not counted).  Hop
    and redirect counts come from the kernel's own counters."""
    return 16.0 * hops + 8.0 * redirects + per_edge * edges


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(kernel, {}).get(workload)
    return None if e is None else float(e["bytes_per_launch"])


# ---------------------------------------------------------------------------
def leg_incremental(args, ctx, grid: bool = False) -> dict:
    """Snapshot-batched insertions through dyg_incremental_update_dev (device
    edits) against the plain form and the oracle's whole run on the host's
    cores: the 1M-point 8-NN graph (200k insertions, threshold 100, 256
    walkers) or, with grid, the 128x128 grid (2,000 insertions, threshold 10,
    512 walkers)."""
    from dygrass_amd import (DeviceArray, DeviceGraph, csr_from_edges, incremental_update,
                             incremental_update_dev, params, synth)

    t0 = time.time()
    wl = (synth.config_grid_insert if grid else synth.config_rgg_insert)(args.scale, args.seed)
    leg = "grid" if grid else "incremental"
    m = len(wl.ins)
    log(f"{leg}: {wl.name} V={wl.G.V} E_H={len(wl.H)} updates={m} "
        f"({time.time() - t0:.1f}s)")
    rp, col, w = csr_from_edges(wl.H.V, wl.H.u, wl.H.v, wl.H.w)
    I = (m + 9) // 10

    def run_mode(plain: bool, steps: int, warmup: int) -> dict:
        p = params(wl.walkers, 100, wl.threshold, seed=args.seed, per_edge=plain,
                   host_edits=plain, block_edges=args.block_edges)
        if not plain:  # inputs resident in HBM before the timed window
            d = [DeviceArray.from_numpy(ctx, a) for a in (wl.ins.u, wl.ins.v, wl.ins.w)]
            d_found, d_cond = DeviceArray(ctx, m, np.int32), DeviceArray(ctx, m, np.float32)

        final = {}

        def one_step():
            H = DeviceGraph(ctx, wl.H.V, rp, col, w, slack=2)  # original sparsifier (untimed)
            if not plain:
                H.reserve(m, I)
            ctx.sync()
            ts = time.perf_counter()
            if plain:
                found, st = incremental_update(H, wl.ins.u, wl.ins.v, wl.ins.w, p)
            else:
                st = incremental_update_dev(H, d[0].ptr, d[1].ptr, d[2].ptr, m, d_found.ptr,
                                            d_cond.ptr, p)
            ctx.sync()
            te = time.perf_counter()
            if not plain:
                found = d_found.numpy()
            final["H"] = H.download()  # updated sparsifier (untimed)
            H.close()
            return te - ts, found, st

        for _ in range(warmup):
            one_step()
        times, walk_ms, upd_ms = [], [], []
        for i in range(steps):
            dt, found, st = one_step()
            times.append(dt)
            walk_ms.append(st["walk_ms"])
            upd_ms.append(st["update_ms"])
            log(f"{leg} {'plain' if plain else 'device'} step {i}: {dt * 1e3:.2f} ms "
                f"(walk kernels {st['walk_ms']:.2f} ms, graph edits {st['update_ms']:.2f} ms)")
        total = float(np.sum(times))
        return dict(found=found, stats=st, total=total, steps=steps, value=m * steps / total,
                    ms_per_step=total / steps * 1e3, walk_ms=float(np.mean(walk_ms)),
                    edit_ms=float(np.mean(upd_ms)), H=final["H"])

    head = run_mode(False, args.steps, args.warmup)
    stats = head["stats"]
    launches = int(stats["batches"])
    bytes_run = algorithmic_bytes(stats["hops"], stats["redirects"], m)
    achieved = bytes_run / (head["walk_ms"] / 1e3) / 1e9
    traffic = measured_traffic(f"k_walk_pool<{int(np.log2(wl.walkers))},true,true,true>",
                               wl.name)
    out = {
        "value": head["value"], "ms_per_step": head["ms_per_step"],
        "workload": f"{wl.name}: {launches} snapshot batches of {I} insertions, threshold "
                    f"{wl.threshold:g}, {wl.walkers} walkers, 100 steps (work-queue walk "
                    "kernel, graph edits on the device)",
        "V": int(wl.G.V), "E_H": int(len(wl.H)), "updates": int(m),
        "walk_kernel_ms_per_step": head["walk_ms"], "graph_edit_ms_per_step": head["edit_ms"],
        "walk_steps_per_s": stats["hops"] * head["steps"] / head["total"],
        "no_path_fraction": float(stats["added"]) / m,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"k_walk_pool<W={wl.walkers}, insert, prune>",
                     "workload": wl.name, "launches": launches,
                     "bytes_per_launch": bytes_run / launches,
                     "avg_launch_ms": head["walk_ms"] / launches,
                     "edges_per_launch": m / launches,
                     "hops_per_launch": stats["hops"] / launches,
                     "redirects_per_launch": stats["redirects"] / launches,
                     "bytes_basis": "16 B per neighbour draw + 8 B per redirect + 20 B per "
                                    "edge"},
    }
    if not args.no_plain:
        plain = run_mode(True, max(1, min(args.steps, 3)), 1)
        out["plain"] = {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                        "walk_kernel_ms_per_step": plain["walk_ms"],
                        "graph_edit_ms_per_step": plain["edit_ms"],
                        "shape": "one workgroup per edge, graph edits planned on the host"}
        out["speedup_vs_plain"] = head["value"] / plain["value"]
        out["plain_decisions_identical"] = bool(np.array_equal(plain["found"], head["found"]))
        out["plain_sparsifier_identical"] = all(
            np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                           b.view(np.uint32) if b.dtype == np.float32 else b)
            for a, b in zip(plain["H"], head["H"]))
    if not args.no_kappa:
        from dygrass_amd.kappa import kappa

        hrp, hcol, hw = head["H"]
        rows = np.repeat(np.arange(wl.H.V, dtype=np.int64), np.diff(hrp).astype(np.int64))
        once = hcol > rows  # each undirected edge once
        dense = (np.concatenate([wl.G.u, wl.ins.u]), np.concatenate([wl.G.v, wl.ins.v]),
                 np.concatenate([wl.G.w, wl.ins.w]))
        ts = time.perf_counter()
        k_after = kappa(wl.G.V, dense, (rows[once], hcol[once], hw[once]), tol=1e-8)
        k_before = kappa(wl.G.V, dense, (wl.H.u, wl.H.v, wl.H.w), tol=1e-8)
        out["kappa"] = {"before": k_before, "after": k_after,
                        "edges_H_after": int(once.sum()),
                        "definition": "lambda_max(L_G+ins + 1e-6 I, L_H + 1e-6 I), FP64 "
                                      "Lanczos on the host; plain and device H identical, so "
                                      "kappa is equal by construction",
                        "seconds": time.perf_counter() - ts}
        log(f"{leg} kappa: {k_before:.1f} -> {k_after:.1f}")
    if not args.no_cpu:
        out["_cpu"] = ("grid" if grid else "incremental", head["found"])
    return out


# ---------------------------------------------------------------------------
def tuned(name: str, w_ins: int, w_del: int):
    """The autotuner's setting for workload `name` (dygrass_amd/tuning/<name>.json,
    made by `python -m dygrass_amd.autotune` on this GPU) and a summary for the
    JSON line; without a committed tuning, the fixed walker counts given."""
    from dygrass_amd import autotune

    t = autotune.load(name)
    if t is None:
        s = autotune.Setting(w_ins=w_ins, w_del=w_del)
        return s, {"source": None, "setting": s.asdict(), "note": "no committed tuning"}
    s = autotune.setting_of(t)
    return s, {"source": f"dygrass_amd/tuning/{name}.json", "setting": s.asdict(),
               "anchor": t["anchor"]["setting"],
               "growth_anchor": t["anchor"]["growth"], "growth": t["chosen"]["growth"],
               "kappa_anchor": t["anchor"]["kappa"], "kappa": t["chosen"]["kappa"],
               "tolerances": t["tolerances"], "calibration_updates": t["updates"],
               "command": t["command"]}


def sp_of(setting, threshold: float, seed: int, max_batch: int, max_inserts: int, plain: bool):
    """Stream parameters of a tuned setting; the plain form takes its walker
    counts and step budgets (same results), not its device-only shape knobs."""
    from dygrass_amd import stream_params

    s = setting
    return stream_params(w_ins=s.w_ins, w_del=s.w_del, threshold=threshold, seed=seed,
                         max_batch=max_batch, max_inserts=max_inserts, plain=plain,
                         steps_ins=s.steps_ins, steps_del=s.steps_del,
                         block_edges_ins=0 if plain else s.block_ins,
                         block_edges_del=0 if plain else s.block_del,
                         dense_no_desc=s.dense_no_desc and not plain,
                         graph=s.graph and not plain, small_1024=s.small_1024)


def upload_batches(ctx, batches) -> list:
    """Every batch's arrays in device memory (DeviceArray u, v, w or None)."""
    from dygrass_amd import DeviceArray

    out = []
    for op, u, v, w in batches:
        du = DeviceArray.from_numpy(ctx, np.ascontiguousarray(u, np.uint32))
        dv = DeviceArray.from_numpy(ctx, np.ascontiguousarray(v, np.uint32))
        dw = (DeviceArray.from_numpy(ctx, np.ascontiguousarray(w, np.float32))
              if w is not None else None)
        out.append((op, du, dv, dw))
    ctx.sync()
    return out


def upload_list(ctx, batches) -> tuple:
    """The whole batch list as one set of device arrays: (ops, offsets, u, v, w)
    (w zero at deletion updates)."""
    from dygrass_amd import DeviceArray

    off = np.concatenate([[0], np.cumsum([len(b[1]) for b in batches])]).astype(np.uint64)
    u = np.concatenate([np.asarray(b[1], np.uint32) for b in batches])
    v = np.concatenate([np.asarray(b[2], np.uint32) for b in batches])
    w = np.concatenate([np.asarray(b[3], np.float32) if b[3] is not None
                        else np.zeros(len(b[1]), np.float32) for b in batches])
    out = ([b[0] for b in batches], off, DeviceArray.from_numpy(ctx, u),
           DeviceArray.from_numpy(ctx, v), DeviceArray.from_numpy(ctx, w))
    ctx.sync()
    return out


def run_stream(ctx, wl, plain: bool, seed: int, setting, csr=None, dev_list=None) -> dict:
    """One pass of the whole batch list through a fresh stream; the stream is
    built (graphs uploaded, buffers reserved) outside the timed window.  With
    dev_list (upload_list) every batch is already in HBM when the timed window
    opens and the list goes in with one call (dyg_stream_apply_many_dev);
    otherwise batch after batch from host arrays through pinned staging
    (dyg_stream_apply, PCIe inside the window)."""
    from dygrass_amd import Stream

    maxb = max(len(b[1]) for b in wl.batches)
    sp = sp_of(setting, wl.threshold, seed, maxb, wl.updates, plain)
    st = Stream(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w), sp, csr=csr)
    ctx.sync()
    outs, lat, walk_ms, hops, agg, walks = [], [], 0.0, 0, {}, []
    ts = time.perf_counter()
    if dev_list is not None:
        o, sts = st.apply_many_dev(*dev_list)
        results = [(op, None, s_) for op, s_ in zip(dev_list[0], sts)]
        outs.append(o)
    else:
        results = []
        for op, u, v, w in wl.batches:
            t0 = time.perf_counter()
            o, stt = st.apply(op, u, v, w)
            results.append((op, time.perf_counter() - t0, stt))
            outs.append(o)
    ctx.sync()
    dt = time.perf_counter() - ts  # the work only: the bookkeeping below is not timed
    for op, dt_b, stt in results:
        lat.append(dt_b if dt_b is not None else stt["total_ms"] / 1e3)
        walk_ms += stt["walk_ms"]
        hops += stt["hops"]
        walks.append({"op": op, "kept": int(stt["found"] + stt["added"]), "hops": int(stt["hops"]),
                      "redirects": int(stt["redirects"]), "walk_ms": float(stt["walk_ms"])})
        for k in ("weight_updates", "found", "added", "skipped", "g_only", "recovered",
                  "heuristic", "invalid"):
            agg[k] = agg.get(k, 0) + stt[k]
    H = st.H.download()  # final sparsifier (untimed)
    st.close()
    return dict(seconds=dt, outcomes=np.concatenate(outs), latencies=np.array(lat) * 1e3,
                walk_ms=walk_ms, hops=hops, counts=agg, H=H, walks=walks)


def stream_roofline(runs, walkers: int, workload: str) -> dict:
    """Roofline of the batched stream's dominant kernel, the insertion walk,
    over every insertion batch's walk launch in the timed passes:
    algorithmic bytes (algorithmic_bytes: 16 B per neighbour draw, 8 B per
    redirect, 20 B per edge; counts from the kernel's own counters) over the
    launches' HIP-event durations.  `per_launch` lists the last pass's
    launches as [edges, hops, redirects, microseconds], so the figure can be
    recomputed from them (and from a kernel trace's durations)."""
    nbytes, ms, launches = 0.0, 0.0, 0
    hops = redirects = edges = 0
    for r in runs:
        for b in r["walks"]:
            if b["op"] != "ins" or b["kept"] == 0:
                continue
            nbytes += algorithmic_bytes(b["hops"], b["redirects"], b["kept"])
            hops += b["hops"]
            redirects += b["redirects"]
            edges += b["kept"]
            ms += b["walk_ms"]
            launches += 1
    achieved = nbytes / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    kernel = f"k_walk_pool<{int(np.log2(walkers))},true,true,true>"
    last = [[b["kept"], b["hops"], b["redirects"], round(b["walk_ms"] * 1e3, 2)]
            for b in runs[-1]["walks"] if b["op"] == "ins" and b["kept"] > 0]
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kernel, workload),
            "kernel": f"k_walk_pool<W={walkers}, insert, prune, decision only>",
            "workload": workload, "launches": launches,
            "bytes_per_launch": nbytes / max(launches, 1),
            "avg_launch_ms": ms / max(launches, 1),
            "hops_per_launch": hops / max(launches, 1),
            "redirects_per_launch": redirects / max(launches, 1),
            "edges_per_launch": edges / max(launches, 1),
            "bytes_basis": "16 B per neighbour draw + 8 B per redirect + 20 B per edge",
            "per_launch": last}


def leg_batched(args, ctx) -> dict:
    from dygrass_amd import csr_from_edges, synth

    t0 = time.time()
    wl = synth.config_batched(args.scale, args.seed)
    m = wl.updates
    log(f"batched: {wl.name} V={wl.G.V} E_G={len(wl.G)} E_H={len(wl.H)} batches="
        f"{len(wl.batches)} updates={m} ({time.time() - t0:.1f}s)")
    setting, tuning = tuned(wl.name, wl.w_ins, wl.w_del)
    log(f"batched setting: {setting}")
    csr = (csr_from_edges(wl.G.V, wl.G.u, wl.G.v, wl.G.w),
           csr_from_edges(wl.G.V, wl.H.u, wl.H.v, wl.H.w))
    dev_list = upload_list(ctx, wl.batches)  # inputs resident in HBM
    for _ in range(args.warmup):
        run_stream(ctx, wl, False, args.seed, setting, csr, dev_list)
    runs = []
    for i in range(args.steps):
        r = run_stream(ctx, wl, False, args.seed, setting, csr, dev_list)
        runs.append(r)
        log(f"batched step {i}: {r['seconds'] * 1e3:.1f} ms (insertion walk kernels "
            f"{r['walk_ms']:.1f} ms)")
    hr = run_stream(ctx, wl, False, args.seed, setting, csr)  # host arrays: PCIe included
    log(f"batched from host arrays: {hr['seconds'] * 1e3:.1f} ms")
    total = sum(r["seconds"] for r in runs)
    head = runs[-1]
    out = {
        "value": m * len(runs) / total, "ms_per_step": total / len(runs) * 1e3,
        "workload": f"{wl.name}: {len(wl.batches)} batches of log-uniform size 1..50000 "
                    f"({m} updates, 70% insertion batches), threshold {wl.threshold:g}, "
                    f"{setting.w_ins} insertion / {setting.w_del} deletion walkers, "
                    f"{setting.steps_ins}/{setting.steps_del} steps (autotuned; device-resident "
                    "stream: work-queue walk kernels, device edits)",
        "V": int(wl.G.V), "E_G": int(len(wl.G)), "E_H": int(len(wl.H)), "updates": int(m),
        "batches": len(wl.batches),
        # a list times only its insertion walks (events on the deletion
        # walks' dispatch would cost more than those walks' share of the time)
        "insertion_walk_kernel_ms_per_step": head["walk_ms"],
        "walk_steps_per_s": head["hops"] / head["seconds"],
        "batch_latency_ms": {"p50": pct(head["latencies"], 50),
                             "p99": pct(head["latencies"], 99)},
        "outcomes": head["counts"],
        "roofline": stream_roofline(runs, setting.w_ins, wl.name),
        "tuning": tuning,
        "growth_H": int(len(head["H"][1]) // 2 - len(wl.H)),
        "inputs": "every batch in HBM before the timed window, the list applied in one call "
                  "(dyg_stream_apply_many_dev)",
        "host_inputs": {"value": m / hr["seconds"], "ms_per_step": hr["seconds"] * 1e3,
                        "same_outcomes": bool(np.array_equal(hr["outcomes"], head["outcomes"])),
                        "note": "batches from host arrays through pinned staging "
                                "(dyg_stream_apply): PCIe copies inside the timed window"},
    }
    if tuning.get("source"):
        # the stream is deterministic: this run's growth is the tuning run's
        out["growth_matches_tuning"] = out["growth_H"] == tuning["growth"]
        # the same stream at the reference's walker counts and step budgets
        # (the tuning's anchor: 512 walkers, 100 steps for both kinds), the
        # tuned launch shape kept: what the chosen budgets are measured against
        from dygrass_amd import autotune

        anc = dict(tuning["anchor"])
        ref = autotune.Setting(**{**setting.asdict(), **{k: anc[k] for k in
                                                        ("w_ins", "w_del", "steps_ins",
                                                         "steps_del")}})
        run_stream(ctx, wl, False, args.seed, ref, csr, dev_list)  # warm-up
        rr = [run_stream(ctx, wl, False, args.seed, ref, csr, dev_list) for _ in range(2)]
        rt = sum(r["seconds"] for r in rr) / len(rr)
        log(f"batched at the reference's budgets: {rt * 1e3:.1f} ms")
        out["reference_budgets"] = {
            "value": m / rt, "ms_per_step": rt * 1e3, "setting": ref.asdict(),
            "insertion_walk_kernel_ms_per_step": rr[-1]["walk_ms"],
            "growth_H": int(len(rr[-1]["H"][1]) // 2 - len(wl.H)),
            "growth_matches_anchor": int(len(rr[-1]["H"][1]) // 2 - len(wl.H))
            == tuning["growth_anchor"]}
    if not args.no_plain:
        pr = run_stream(ctx, wl, True, args.seed, setting, csr)
        log(f"batched plain: {pr['seconds'] * 1e3:.1f} ms")
        out["plain"] = {"value": m / pr["seconds"], "ms_per_step": pr["seconds"] * 1e3,
                        "walk_kernel_ms_per_step": pr["walk_ms"],
                        "batch_latency_ms": {"p50": pct(pr["latencies"], 50),
                                             "p99": pct(pr["latencies"], 99)},
                        "shape": "host edge maps and edit plans, one workgroup per edge, "
                                 "a wait per batch"}
        out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
        out["plain_outcomes_identical"] = bool(np.array_equal(pr["outcomes"], head["outcomes"]))
        out["plain_sparsifier_identical"] = all(
            np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                           b.view(np.uint32) if b.dtype == np.float32 else b)
            for a, b in zip(pr["H"], head["H"]))
    if not args.no_cpu:
        out["_cpu"] = ("batched", head["outcomes"])
    return out


def cpu_threads(args) -> int:
    """Host threads for the CPU baselines: --cpu-threads, else the cores this
    process may use (OMP_NUM_THREADS when the environment sets it: a GPU box
    allots its cores per GPU that way), else nproc.  Read at start-up, before
    any OpenMP runtime binds a thread."""
    if args.cpu_threads:
        return args.cpu_threads
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline_child(args, leg: str, gpu_outcomes) -> dict:
    """Runs one leg's CPU baseline in a child process whose OpenMP threads
    are bound to neighbouring cores (OMP_PROC_BIND=close, OMP_PLACES=cores,
    OMP_NUM_THREADS = the cores counted at start-up): the binding stays out of
    this process, whose served loop and GPU threads must not inherit it.  The
    child regenerates the same seeded workload and compares its outcomes with
    the GPU's (passed through a temporary file)."""
    import subprocess
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "gpu_outcomes.npy")
        np.save(path, np.asarray(gpu_outcomes))
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores",
                   OMP_NUM_THREADS=str(args.threads_at_start))
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", leg, "--gpu-outcomes",
               path, "--seed", str(args.seed), "--scale", str(args.scale),
               "--cpu-threads", str(args.threads_at_start)]
        p = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=900)
        if p.returncode != 0:
            raise RuntimeError(f"CPU baseline of {leg} failed ({p.returncode})")
        return json.loads(p.stdout.strip().splitlines()[-1])


def cpu_leg_main(args) -> int:
    """--cpu-leg: one leg's CPU baseline on this process's host cores (no GPU
    work), printed as one JSON line."""
    from dygrass_amd import csr_from_edges, synth

    gpu = np.load(args.gpu_outcomes) if args.gpu_outcomes else None
    threads = args.cpu_threads = cpu_threads(args)  # before any OpenMP runtime starts
    if args.cpu_leg == "batched":
        wl = synth.config_batched(args.scale, args.seed)
        setting, _ = tuned(wl.name, wl.w_ins, wl.w_del)
        csr = (csr_from_edges(wl.G.V, wl.G.u, wl.G.v, wl.G.w),
               csr_from_edges(wl.G.V, wl.H.u, wl.H.v, wl.H.w))
        out = cpu_stream_batched(wl, args, setting, csr, gpu)
    else:
        from oracle import pyoracle as po

        grid = args.cpu_leg == "grid"
        wl = (synth.config_grid_insert if grid else synth.config_rgg_insert)(args.scale,
                                                                             args.seed)
        m = len(wl.ins)
        rp, col, w = csr_from_edges(wl.H.V, wl.H.u, wl.H.v, wl.H.w)
        og = po.Graph(wl.H.V, rp, col, w)
        ts = time.perf_counter()
        cfound, _ = og.incremental(wl.ins.u, wl.ins.v, wl.ins.w, wl.threshold, wl.walkers,
                                   seed=args.seed, threads=threads)
        cdt = time.perf_counter() - ts
        log(f"{args.cpu_leg} cpu: {m / cdt:.0f} edge-updates/s on {threads} threads")
        out = {"value": m / cdt, "unit": "edge-updates/s", "cores": threads,
               "nproc": os.cpu_count(), "kind": "port",
               "sample": "the whole workload (every snapshot batch's walks and the sparsifier "
                         "appends) through the oracle, OpenMP over each batch's edges, "
                         f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', 'unset')}",
               "decisions_identical": (bool(np.array_equal(gpu, cfound))
                                       if gpu is not None else None)}
    print(json.dumps(out), flush=True)
    return 0


def cpu_stream_batched(wl, args, setting, csr, gpu_outcomes) -> dict:
    """The whole batched stream through the C++ stream oracle on the host's
    cores: every batch's classification, G and H edits, walks (OpenMP over the
    batch's edges; insertion walks decide only, as on the GPU) and recovery,
    from the original graphs, batch after batch while the first 30 s last.
    Its outcome codes are compared with the GPU pass's."""
    from oracle import pyoracle as po

    threads = cpu_threads(args)
    s = setting
    orc = po.StreamCpp(wl.G.V, thr=wl.threshold, w_ins=s.w_ins, w_del=s.w_del, seed=args.seed,
                       steps_ins=s.steps_ins, steps_del=s.steps_del, threads=threads, csr=csr)
    outs, nb = [], 0
    for op, u, v, w in wl.batches:
        outs.append(orc.apply(op, u, v, w))
        nb += 1
        if orc.stats()["total_seconds"] > 30.0:
            break
    st = orc.stats()
    orc.close()
    got = np.concatenate(outs)
    n = len(got)
    same = (bool(np.array_equal(got, gpu_outcomes[:n])) if gpu_outcomes is not None else None)
    log(f"batched cpu: {n / st['total_seconds']:.0f} edge-updates/s on {threads} threads "
        f"({nb} of {len(wl.batches)} batches, outcomes "
        f"{'not compared' if same is None else 'equal' if same else 'DIFFER'})")
    return {"value": n / st["total_seconds"], "unit": "edge-updates/s", "cores": threads,
            "nproc": os.cpu_count(), "kind": "port",
            "sample": f"the stream's first {nb} of {len(wl.batches)} batches ({n} updates: "
                      "classification, G and H edits, walks and recovery) through the C++ "
                      f"stream oracle, OpenMP over each batch's walks, OMP_PROC_BIND="
                      f"{os.environ.get('OMP_PROC_BIND', 'unset')}; "
                      f"{st['walk_seconds']:.1f} of {st['total_seconds']:.1f} s in walks",
            "outcomes_identical": same, "batches": nb,
            "walk_seconds": st["walk_seconds"], "total_seconds": st["total_seconds"]}


# ---------------------------------------------------------------------------
def leg_served(args, ctx) -> dict:
    from dygrass_amd import Server, Stream, stream_params, synth

    t0 = time.time()
    wl = synth.config_served(args.scale, args.seed, args.requests, args.rate)
    m = wl.updates
    n = len(wl.requests)
    log(f"served: {wl.name} V={wl.G.V} E_G={len(wl.G)} E_H={len(wl.H)} requests={n} "
        f"updates={m} rate={args.rate:g}/s ({time.time() - t0:.1f}s)")
    maxr = max(len(r[1]) for r in wl.requests)
    sizes = np.array([len(r[1]) for r in wl.requests])
    g_edges, h_edges = (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w)

    setting, tuning = tuned(wl.name, wl.w_ins, wl.w_del)
    log(f"served setting: {setting}")

    def sp(plain):
        return sp_of(setting, wl.threshold, args.seed, maxr, m, plain)

    def serve(arrivals):
        st = Stream(ctx, wl.G.V, g_edges, h_edges, sp(False))
        srv = Server(st, slots=args.slots, max_request=maxr)
        tickets, lat, done_at = [], np.zeros(n), np.zeros(n)
        got = [None] * n

        def collect():
            for i in range(n):
                while len(tickets) <= i:
                    time.sleep(0.0002)
                got[i] = srv.wait(tickets[i], int(sizes[i]))
                lat[i] = got[i][1]["latency_ms"]
                done_at[i] = time.perf_counter()

        th = threading.Thread(target=collect)
        th.start()
        start = time.perf_counter()
        for i, (op, u, v, w) in enumerate(wl.requests):
            delay = start + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(op, u, v, w))
        th.join()
        span = done_at[-1] - start
        srv.close()
        st.close()
        walks = np.zeros((n, 4))  # per request: walked edges, hops, redirects, walk ms
        for i, (codes, res) in enumerate(got):
            walked = np.isin(codes, [OUT_FOUND, OUT_ADDED] if wl.requests[i][0] == "ins"
                             else [OUT_RECOVERED, OUT_HEURISTIC])
            walks[i] = (walked.sum(), res["hops"], res["redirects"], res["walk_ms"])
        return span, lat, walks

    span, lat, walks = serve(wl.arrivals)
    small = sizes <= 1000
    out = {
        "value": m / span,
        "workload": f"{wl.name}: {n} requests, Poisson arrivals at {args.rate:g}/s, sizes "
                    f"log-uniform 1..10000 ({m} updates, 60% insertions), threshold "
                    f"{wl.threshold:g}, {setting.w_ins}/{setting.w_del} walkers, "
                    f"{setting.steps_ins}/{setting.steps_del} steps (autotuned), {args.slots} "
                    "slots",
        "V": int(wl.G.V), "E_G": int(len(wl.G)), "E_H": int(len(wl.H)), "requests": n,
        "updates": int(m), "span_s": span, "tuning": tuning,
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
        "latency_ms_le_1k": {"p50": pct(lat[small], 50), "p99": pct(lat[small], 99),
                             "requests": int(small.sum())},
    }
    log(f"served at {args.rate:g}/s: {out['value']:.0f} updates/s, p50 {pct(lat, 50):.2f} ms, "
        f"p99 {pct(lat, 99):.2f} ms (<=1k edges: p99 {pct(lat[small], 99):.2f} ms)")
    # the deletion walks in the dense graph (the leg's dominant kernel): one
    # launch per deletion request that walks, timed by events taken with its
    # dispatch, bytes from its own hop and redirect counters
    dl = np.array([r[0] == "del" for r in wl.requests]) & (walks[:, 0] > 0)
    if dl.any():
        e, h, r_, ms = walks[dl].sum(axis=0)
        nb = algorithmic_bytes(h, r_, e, per_edge=16.0)
        achieved = nb / (ms * 1e-3) / 1e9
        kernel = f"k_walk_pool<{int(np.log2(setting.w_del))},false,true,true>"
        out["roofline_delete_walks"] = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kernel, wl.name),
            "kernel": f"k_walk_pool<W={setting.w_del}, delete, prune>", "workload": wl.name,
            "launches": int(dl.sum()), "bytes_per_launch": nb / dl.sum(),
            "avg_launch_ms": ms / dl.sum(), "edges_per_launch": e / dl.sum(),
            "hops_per_launch": h / dl.sum(), "redirects_per_launch": r_ / dl.sum(),
            "bytes_basis": "16 B per neighbour draw + 8 B per redirect + 16 B per edge "
                           "(winner path not counted)"}
        log(f"served deletion walks: {achieved:.0f} GB/s algorithmic over {int(dl.sum())} "
            f"launches, {ms / dl.sum() * 1e3:.1f} us per launch")
    cspan, _, _ = serve(np.zeros(n))  # every request waiting at t=0: capacity
    out["capacity"] = {"value": m / cspan, "span_s": cspan,
                       "requests_per_s": n / cspan,
                       "note": "same trace, all requests arriving at once"}
    log(f"served capacity: {m / cspan:.0f} updates/s ({n / cspan:.0f} requests/s)")
    curve = []
    for rate in [float(x) for x in args.rates.split(",") if x.strip()]:
        # the same trace with its arrival times compressed to `rate`
        rspan, rlat, _ = serve(wl.arrivals * (args.rate / rate))
        curve.append({"rate": rate, "updates_per_s": m / rspan,
                      "p50_ms": pct(rlat, 50), "p99_ms": pct(rlat, 99),
                      "p50_ms_le_1k": pct(rlat[small], 50), "p99_ms_le_1k": pct(rlat[small], 99)})
        log(f"served at {rate:g}/s: p50 {pct(rlat, 50):.2f} ms, p99 {pct(rlat, 99):.2f} ms "
            f"(<=1k edges: p99 {pct(rlat[small], 99):.2f} ms)")
    out["latency_curve"] = curve
    if not args.no_plain and args.plain_requests > 0:
        k = min(n, args.plain_requests)
        pst = Stream(ctx, wl.G.V, g_edges, h_edges, sp(True))
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
                        "latency_ms": {"p50": pct(plat, 50), "p99": pct(plat, 99)},
                        "shape": "requests applied one at a time in arrival order, host edge "
                                 "maps and edit plans, one workgroup per edge",
                        "note": "value = updates / busy time over the first requests of the "
                                "same trace"}
        out["speedup_vs_plain"] = out["capacity"]["value"] / out["plain"]["value"]
        log(f"served plain ({k} requests): {pm / busy:.0f} updates/s busy, "
            f"p99 {pct(plat, 99):.1f} ms")
    return out


# ---------------------------------------------------------------------------
def leg_decremental(args, ctx) -> dict:
    """The plain-decremental configuration: 2M-node triangulated mesh, 100k
    deletions in 10 snapshot batches, 512 walkers.  Device form: a
    deletion-only stream (G and H on the device); plain form: the dual graph
    (G with the host edge map, H as a host key set, one workgroup per edge)."""
    from dygrass_amd import DualGraph, Stream, csr_from_edges, params, stream_params, synth

    t0 = time.time()
    wl = synth.config_mesh_delete(args.scale, args.seed)
    G, H, D = wl.G, wl.H, wl.dels
    M = len(D)
    I = (M + 9) // 10
    log(f"decremental: {wl.name} V={G.V} E_G={len(G)} E_H={len(H)} deletions={M} "
        f"({time.time() - t0:.1f}s)")
    sp = stream_params(w_ins=512, w_del=512, threshold=0.0, seed=args.seed, max_batch=I,
                       max_inserts=4 * M, dense_no_desc=True)
    batches = [("del", D.u[b0:b0 + I], D.v[b0:b0 + I], None) for b0 in range(0, M, I)]
    dev_list = upload_list(ctx, batches)  # resident in HBM, applied in one call
    times, counts, h_list = [], None, None
    for i in range(args.warmup + args.steps):
        st = Stream(ctx, G.V, (G.u, G.v, G.w), (H.u, H.v, H.w), sp)
        ctx.sync()
        ts = time.perf_counter()
        _, sts = st.apply_many_dev(*dev_list)
        ctx.sync()
        dt = time.perf_counter() - ts
        if i >= args.warmup:
            times.append(dt)
            log(f"decremental device step {i - args.warmup}: {dt * 1e3:.2f} ms")
        counts = {k: int(sum(x[k] for x in sts))
                  for k in ("skipped", "g_only", "recovered", "heuristic")}
        h_list = st.H.download()
        st.close()
    # the deletion walks' durations: a list takes no events on deletion walks,
    # so one more pass applies the batches one call each, every walk timed by
    # events taken with its dispatch (same batches, same results)
    dev_batches = upload_batches(ctx, batches)
    st = Stream(ctx, G.V, (G.u, G.v, G.w), (H.u, H.v, H.w), sp)
    ctx.sync()
    walks = []
    ts = time.perf_counter()
    for _, du, dv, _ in dev_batches:
        _, stt = st.apply_dev("del", du, dv)
        walks.append((stt["recovered"] + stt["heuristic"], stt["hops"], stt["redirects"],
                      stt["walk_ms"]))
    ctx.sync()
    per_batch_s = time.perf_counter() - ts
    rp, col, _ = st.H.download()
    st.close()
    same_h = bool(np.array_equal(rp, h_list[0]) and np.array_equal(col, h_list[1]))
    rows = np.repeat(np.arange(G.V, dtype=np.int64), np.diff(rp).astype(np.int64))
    h_dev = set(zip(rows[col > rows].tolist(), col[col > rows].tolist()))
    total = float(np.sum(times))
    out = {"value": M * len(times) / total, "ms_per_step": total / len(times) * 1e3,
           "workload": f"{wl.name}: {M} deletions in 10 snapshot batches, 512 walkers, 100 "
                       "steps (deletion-only stream, G and H on the device, G without "
                       "neighbour descriptors, batches resident in HBM, the list applied in "
                       "one call)",
           "V": int(G.V), "E_G": int(len(G)), "E_H": int(len(H)), "deletions": M,
           "outcomes": counts, "H_after": len(h_dev),
           "batch_by_batch": {"ms_per_step": per_batch_s * 1e3, "same_H": same_h,
                              "note": "the same batches one dyg_stream_apply_dev call each "
                                      "(a wait per batch), every deletion walk timed: the "
                                      "roofline below"}}
    nbytes = sum(algorithmic_bytes(h, r, k, per_edge=16.0) for k, h, r, _ in walks)
    wms = sum(x[3] for x in walks)
    achieved = nbytes / (wms / 1e3) / 1e9 if wms > 0 else 0.0
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS,
                       "traffic": measured_traffic("k_walk_pool<9,false,true,false>", wl.name),
                       "kernel": "k_walk_pool<W=512, delete, prune, row table per hop>",
                       "workload": wl.name, "launches": len(walks),
                       "bytes_per_launch": nbytes / max(len(walks), 1),
                       "avg_launch_ms": wms / max(len(walks), 1),
                       "hops_per_launch": sum(x[1] for x in walks) / max(len(walks), 1),
                       "bytes_basis": "16 B per neighbour draw + 8 B per redirect + 16 B per "
                                      "edge (winner path not counted)"}
    if not args.no_plain:
        grp, gcol, gw = csr_from_edges(G.V, G.u, G.v, G.w)
        dual = DualGraph(ctx, G.V, grp, gcol, gw, H.u, H.v)
        p = params(512, 100, 0.0, seed=args.seed, per_edge=True)
        ts = time.perf_counter()
        dual.run(D.u, D.v, p)
        ctx.sync()
        dt = time.perf_counter() - ts
        added = dual.added_edges()
        out["plain"] = {"value": M / dt, "ms_per_step": dt * 1e3,
                        "shape": "dual graph: host edge map and key set, host-planned deletes, "
                                 "one workgroup per edge"}
        out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
        out["plain_same_sparsifier_size"] = int(dual.sparse_size) == len(h_dev)
        gone = set(zip(np.minimum(D.u, D.v).tolist(), np.maximum(D.u, D.v).tolist()))
        out["plain_recovered_edges_agree"] = all(
            ((min(a, b), max(a, b)) in h_dev) == ((min(a, b), max(a, b)) not in gone)
            for a, b in added.tolist())
        dual.close()
        log(f"decremental plain: {dt * 1e3:.1f} ms")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="workload scale (tests only)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="all",
                    help="all, or a comma-separated list of batched, served, incremental, "
                         "decremental, grid (the first is the headline)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparisons")
    ap.add_argument("--no-kappa", action="store_true", help="skip the FP64 kappa check")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = the cores this process may use (OMP_NUM_THREADS when set)")
    ap.add_argument("--block-edges", type=int, default=0,
                    help="edges per workgroup of the work-queue walk kernel (0 = automatic)")
    ap.add_argument("--requests", type=int, default=1500, help="served: requests in the trace")
    ap.add_argument("--rate", type=float, default=1000.0, help="served: arrivals per second")
    ap.add_argument("--slots", type=int, default=8, help="served: requests admitted at once")
    ap.add_argument("--rates", default="2000,3000,4000,5000",
                    help="served: further arrival rates (per second) for the latency curve")
    ap.add_argument("--plain-requests", type=int, default=40,
                    help="served: requests of the trace the plain form serves")
    ap.add_argument("--cpu-leg", choices=["batched", "incremental", "grid"],
                    help=argparse.SUPPRESS)  # child process of a CPU baseline
    ap.add_argument("--gpu-outcomes", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (a known-size "
                         "dispatch for scaling hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); a multi-GPU run would "
              "replicate the graphs and split each snapshot batch's edges across GPUs, which "
              "this build does not implement", file=sys.stderr)
        return 2

    if args.cpu_leg:
        return cpu_leg_main(args)
    # the host cores for the CPU baselines, counted before any OpenMP runtime
    # (the workload generators use one) binds a thread
    args.threads_at_start = cpu_threads(args)
    from dygrass_amd import Context, DeviceArray

    ctx = Context(0)
    if args.copy_calibration_mb:
        import ctypes

        nb = args.copy_calibration_mb << 20
        a, b = DeviceArray(ctx, nb // 4, np.uint32), DeviceArray(ctx, nb // 4, np.uint32)
        ctx._lib.dyg_copy_on_device(ctx.handle, ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr),
                                    nb)
        ctx.sync()
        log(f"calibration copy: {args.copy_calibration_mb} MiB device to device")
        a.close()
        b.close()
    known = ["batched", "served", "incremental", "decremental", "grid"]
    want = known if args.workload == "all" else [x for x in args.workload.split(",") if x]
    if not want or any(x not in known for x in want):
        print(f"unknown workload in {args.workload!r}; choose from {known}", file=sys.stderr)
        return 2
    legs = {}
    for name in want:
        legs[name] = {"batched": leg_batched, "served": leg_served,
                      "incremental": leg_incremental,
                      "decremental": leg_decremental,
                      "grid": lambda a, c: leg_incremental(a, c, grid=True)}[name](args, ctx)
    for name, leg in legs.items():  # CPU baselines, after all GPU work
        job = leg.pop("_cpu", None)
        if job is not None:
            leg["cpu"] = cpu_baseline_child(args, *job)
            leg["speedup_vs_cpu"] = leg["value"] / leg["cpu"]["value"]
    head_name = want[0]
    head = legs[head_name]
    inc = legs.get("incremental")
    cpu = head.pop("cpu", None)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "edge-updates/s",
        "n_gpus": 1,
        "steps": args.steps if head_name != "served" else 1,
        "warmup": args.warmup if head_name != "served" else 0,
        "ms_per_step": head.get("ms_per_step", head.get("span_s", 0.0) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": head.pop("workload"), "form": head_name},
        "roofline": (head.pop("roofline", None)
                     or (inc.get("roofline") if inc is not None else None)),
        "cpu_baseline": cpu,
    }
    for name, leg in legs.items():
        out[name] = leg
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
