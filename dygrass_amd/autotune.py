"""Autotuner of the stream forms' walk settings (walkers per edge, step
budget, workgroup shape, dense-graph layout).

The reference has no tuner: its updaters take the walker count on the
command line (random_walk_incremental/main.cu:240-262, `<#walkers>` at most
512) or fix it with the step bound (random_walk_decremental/main.cu:332-333:
512 walkers, 100 steps; random_walk_incremental/main.cu:44: 100 steps).
Those fixed values are the tuner's *anchor*: the quality the tuned setting
must keep.

Nine knobs, in two kinds:

* Shape knobs change no result, only the time, because every draw is keyed
  by (edge, walker, step) and never by where a walker runs: edges per
  workgroup of the work-queue walk kernel for insertion walks and for
  deletion walks, whether the dense graph G keeps neighbour descriptors in
  its slots, whether one-workgroup batch steps run as captured hipGraphs,
  and whether the one-workgroup path stops at 1024 updates.  They are chosen by timing alone, and every shape's outcome
  codes and final sparsifier must equal the first shape's (checked; a
  mismatch is an error, not a tuning result).
* Quality knobs change which paths the walks find: walkers per edge and
  step budget, for insertion walks and for deletion walks.  A candidate is
  accepted when, over the whole calibration stream, (1) the sparsifier grows
  by at most (1 + tol_growth) x the anchor's growth, and (2) the final
  kappa(L_G, L_H) lies within tol_kappa of the anchor's.  Growth is exact and
  free (the edge count after the stream); kappa is an FP64 Lanczos solve on
  the host (a minute or more on millions of vertices), so every candidate is
  screened on growth and kappa is computed once per knob, at the lowest value
  that passed the screen, stepping back up the knob's values while it misses.

The served workload's sparsifier holds random shortcuts, which fill a sparse
LU factorisation in: its kappa applies L_H^-1 by conjugate gradients on the
host's cores instead (kappa method "pcg", about 100 s at 8M vertices).
`--check-kappa` records kappa for the anchor and the chosen setting of an
existing tuning (the chosen setting must lie within tol_kappa of the
anchor), for a tuning made while kappa was out of reach there.

Search: coordinate descent from the anchor.  Each quality knob walks down
its candidate list (fewer walkers, shorter budgets) while candidates pass
the growth screen and its walk kernels do not get slower (beyond
`time_slack`, the timing noise), then takes its kappa check,
with the knobs set so far kept; then each shape knob takes its fastest value
at the chosen quality setting (edges per workgroup by the walk kernels' own
time, the other shape knobs by the whole stream's).  Times are the
minimum over `repeats` passes of the whole stream from the original graphs
(host clock around the batch loop, inputs from host arrays, as the batched
benchmark times them).

The search is written against two callables, so the same code drives the
device (`StreamTrials`) and the host-only tests:
  run(setting) -> Trial     one pass of the stream with that setting
  kappa(setting) -> float   kappa of the final graphs of that setting's pass

    python -m dygrass_amd.autotune --workload batched --out tuning.json
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class Setting:
    """Walk settings of a stream: quality knobs (w_*, steps_*) and shape
    knobs (block_*, dense_no_desc)."""
    w_ins: int = 512
    w_del: int = 512
    steps_ins: int = 100
    steps_del: int = 100
    block_ins: int = 0        # edges per workgroup, insertion walks (0 = automatic)
    block_del: int = 0        # edges per workgroup, deletion walks (0 = automatic)
    dense_no_desc: bool = False
    graph: bool = False       # one-workgroup batch steps as captured hipGraphs
    small_1024: bool = False  # one-workgroup path for batches of at most 1024 updates

    def quality(self) -> tuple:
        return (self.w_ins, self.w_del, self.steps_ins, self.steps_del)

    def asdict(self) -> dict:
        return dataclasses.asdict(self)


ANCHOR = Setting()  # 512 walkers and 100 steps for both walks, automatic shape

QUALITY_GRID = {
    "w_ins": (512, 256, 128, 64),
    "w_del": (512, 256, 128, 64),
    "steps_del": (100, 80, 64, 48, 32),
    "steps_ins": (100, 80, 64, 48, 32),
}
SHAPE_GRID = {
    "dense_no_desc": (False, True),
    "graph": (False, True),
    "small_1024": (False, True),
    "block_ins": (0, 1, 2, 4, 8, 16),
    "block_del": (0, 1, 2, 4, 8, 16),
}


# shape knobs that change only the walk kernels' launch
WALK_SHAPE_KNOBS = ("block_ins", "block_del")


@dataclass
class Trial:
    """One pass of the calibration stream."""
    seconds: float            # the whole stream, host clock
    walk_ms: float            # walk kernels (HIP events), summed over batches
    edges_h: int              # sparsifier edges after the stream
    digest: str               # outcome codes + final sparsifier, for shape checks
    outcomes: dict = field(default_factory=dict)


@dataclass
class Record:
    setting: Setting
    knob: str
    seconds: float
    walk_ms: float
    growth: int
    accepted: bool
    reason: str

    def asdict(self) -> dict:
        d = dataclasses.asdict(self)
        d["setting"] = self.setting.asdict()
        return d


class TuneError(RuntimeError):
    pass


def tune(run, kappa, edges_h0: int, *, anchor: Setting = ANCHOR, quality_grid=None,
         shape_grid=None, tol_growth: float = 0.02, tol_kappa: float = 0.02, repeats: int = 2,
         time_slack: float = 0.02, log=None) -> dict:
    """Coordinate-descent search (module docstring).  `edges_h0` is the
    sparsifier's edge count before the stream.  With an empty quality_grid
    only the shape knobs are tuned, at the anchor's walker counts and step
    budgets; with kappa None the quality knobs answer to the growth screen
    alone (the result records kappa as None)."""
    qg = QUALITY_GRID if quality_grid is None else quality_grid
    sg = SHAPE_GRID if shape_grid is None else shape_grid
    say = log or (lambda m: None)
    records: list[Record] = []

    def timed(s: Setting) -> Trial:
        # the minimum over the repeats, of the whole stream and of its walks
        best, walk = None, None
        for _ in range(max(1, repeats)):
            t = run(s)
            if best is not None and t.digest != best.digest:
                raise TuneError(f"two passes of {s} disagree: the stream is not deterministic")
            if best is None or t.seconds < best.seconds:
                best = t
            walk = t.walk_ms if walk is None else min(walk, t.walk_ms)
        return dataclasses.replace(best, walk_ms=walk)

    a = timed(anchor)
    g_anchor = a.edges_h - edges_h0
    limit = g_anchor * (1.0 + tol_growth) if g_anchor > 0 else tol_growth * max(edges_h0, 1)
    records.append(Record(anchor, "anchor", a.seconds, a.walk_ms, g_anchor, True, "anchor"))
    say(f"anchor {anchor.quality()}: {a.seconds * 1e3:.1f} ms (walks {a.walk_ms:.2f} ms), "
        f"growth {g_anchor}")

    # kappa of the anchor (skipped for a shape-only tuning, where no quality
    # knob moves, and when no kappa is given: the growth screen alone decides)
    use_kappa = bool(qg) and kappa is not None
    k_anchor = float(kappa(anchor)) if use_kappa else None
    kappa_checks = [(anchor.quality(), k_anchor)] if use_kappa else []
    if use_kappa:
        say(f"kappa anchor {k_anchor:.6g}")

    def kappa_ok(s: Setting) -> bool:
        if not use_kappa:
            return True
        k = float(kappa(s))
        kappa_checks.append((s.quality(), k))
        ok = abs(k - k_anchor) <= tol_kappa * abs(k_anchor)
        say(f"kappa {s.quality()}: {k:.6g} ({'within' if ok else 'outside'} {tol_kappa:.0%})")
        return ok

    cur, cur_t, k_cur = anchor, a, k_anchor
    for knob, values in qg.items():
        start = getattr(cur, knob)
        passed = [(start, cur_t)]  # values that passed the growth screen, in order
        for val in values:
            if val >= start:
                continue  # the grid walks down from the current value
            cand = dataclasses.replace(cur, **{knob: val})
            t = timed(cand)
            growth = t.edges_h - edges_h0
            ok_growth = growth <= limit
            # fewer walkers or steps never add walk work: a step down is taken
            # unless its walk kernels (device-timed, steadier than the whole
            # stream's host-clock time) measure slower by more than the noise
            not_slower = t.walk_ms <= passed[-1][1].walk_ms * (1.0 + time_slack)
            ok = ok_growth and not_slower
            reason = ("passed the growth screen" if ok else
                      f"growth {growth} > {limit:.0f}" if not ok_growth else "slower")
            records.append(Record(cand, knob, t.seconds, t.walk_ms, growth, ok, reason))
            say(f"{knob}={val}: {t.seconds * 1e3:.1f} ms (walks {t.walk_ms:.2f} ms), "
                f"growth {growth} -> {reason}")
            if not ok:
                break
            passed.append((val, t))
        # kappa at the knob's lowest passing value; step back up while it misses
        while len(passed) > 1:
            val, t = passed[-1]
            cand = dataclasses.replace(cur, **{knob: val})
            if kappa_ok(cand):
                cur, cur_t = cand, t
                k_cur = kappa_checks[-1][1] if use_kappa else None
                break
            passed.pop()

    # shape knobs at the chosen quality setting: results must not move.  The
    # walk kernels' launch shape (edges per workgroup) changes only the walk
    # kernels, so it is judged by their device-timed total, which is steadier
    # than the whole stream's host-clock time; the other shape knobs by the
    # whole stream
    ref = cur_t
    for knob, values in sg.items():
        best, best_t = cur, cur_t
        walk_only = knob in WALK_SHAPE_KNOBS
        for val in values:
            if val == getattr(cur, knob):
                continue
            cand = dataclasses.replace(cur, **{knob: val})
            t = timed(cand)
            if t.digest != ref.digest:
                raise TuneError(f"shape knob {knob}={val} changed the results")
            better = (t.walk_ms < best_t.walk_ms) if walk_only else (t.seconds < best_t.seconds)
            records.append(Record(cand, knob, t.seconds, t.walk_ms, t.edges_h - edges_h0, better,
                                  "faster" if better else "slower"))
            say(f"{knob}={val}: {t.seconds * 1e3:.1f} ms (walks {t.walk_ms:.2f} ms)")
            if better:
                best, best_t = cand, t
        cur, cur_t = best, best_t

    return {
        "anchor": {"setting": anchor.asdict(), "seconds": a.seconds, "walk_ms": a.walk_ms,
                   "growth": g_anchor, "kappa": k_anchor, "outcomes": a.outcomes},
        "chosen": {"setting": cur.asdict(), "seconds": cur_t.seconds, "walk_ms": cur_t.walk_ms,
                   "growth": cur_t.edges_h - edges_h0, "kappa": k_cur,
                   "outcomes": cur_t.outcomes},
        "tolerances": {"growth": tol_growth, "kappa": tol_kappa},
        "kappa_checks": [{"quality": list(q), "kappa": k} for q, k in kappa_checks],
        "trials": [r.asdict() for r in records],
    }


# ---------------------------------------------------------------------------
# Device trials
def stream_params_for(s: Setting, threshold: float, seed: int, max_batch: int, max_inserts: int):
    from .core import stream_params

    return stream_params(w_ins=s.w_ins, w_del=s.w_del, threshold=threshold, seed=seed,
                         max_batch=max_batch, max_inserts=max_inserts,
                         steps_ins=s.steps_ins, steps_del=s.steps_del,
                         block_edges_ins=s.block_ins, block_edges_del=s.block_del,
                         dense_no_desc=s.dense_no_desc, graph=s.graph,
                         small_1024=s.small_1024)


class StreamTrials:
    """Runs the calibration stream (G, H, batches) on the device, one fresh
    Stream per pass (the CSR arrays are built once)."""

    def __init__(self, ctx, V: int, g_edges, h_edges, batches, threshold: float, seed: int = 0,
                 kappa_method: str = "lu"):
        from .core import csr_from_edges

        self.kappa_method = kappa_method
        self.ctx, self.V, self.batches = ctx, int(V), batches
        self.threshold, self.seed = float(threshold), seed
        self.g_edges = tuple(np.asarray(x) for x in g_edges)
        self.csr = (csr_from_edges(V, *g_edges), csr_from_edges(V, *h_edges))
        self.edges_h0 = len(h_edges[0])
        self.max_batch = max(len(b[1]) for b in batches)
        self.updates = sum(len(b[1]) for b in batches)
        self.final = None  # (G, H) edge lists of the last pass run with keep=True

    def run(self, s: Setting, keep: bool = False) -> Trial:
        import hashlib

        from .core import Stream

        sp = stream_params_for(s, self.threshold, self.seed, self.max_batch, self.updates)
        st = Stream(self.ctx, self.V, None, None, sp, csr=self.csr)
        try:
            self.ctx.sync()
            walk_ms, counts, outs = 0.0, {}, []
            t0 = time.perf_counter()
            for op, u, v, w in self.batches:
                o, stt = st.apply(op, u, v, w)
                outs.append(o)
                walk_ms += stt["walk_ms"]
                for k in ("found", "added", "recovered", "heuristic", "g_only", "skipped",
                          "weight_updates"):
                    counts[k] = counts.get(k, 0) + int(stt[k])
            self.ctx.sync()
            dt = time.perf_counter() - t0
            H = st.H.download()
            h = hashlib.sha256()
            h.update(np.concatenate(outs).tobytes())
            for a in H:
                h.update(np.ascontiguousarray(a).tobytes())
            if keep:
                self.final = (_edges_once(st.G.download()), _edges_once(H))
            return Trial(dt, walk_ms, int(len(H[1]) // 2), h.hexdigest(), counts)
        finally:
            st.close()

    def kappa(self, s: Setting, progress=None) -> float:
        from .kappa import kappa

        self.run(s, keep=True)  # the final graphs of this setting (deterministic pass)
        G, H = self.final
        self.final = None
        # tolerance far below the 2% bound it is compared at
        return kappa(self.V, G, H, tol=1e-4, method=self.kappa_method, progress=progress)


def _edges_once(csr):
    """(u, v, w) with each undirected edge once, from a (row_ptr, col, w) download."""
    rp, col, w = csr
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp).astype(np.int64))
    once = col > rows
    return rows[once], col[once].astype(np.int64), w[once].astype(np.float64)


def load(name: str, path: str | None = None) -> dict | None:
    """The committed tuning of workload `name` (dygrass_amd/tuning/<name>.json), or None."""
    import os

    p = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", f"{name}.json")
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def setting_of(tuning: dict, which: str = "chosen") -> Setting:
    return Setting(**tuning[which]["setting"])


def retune_shapes(run, edges_h0: int, tuning: dict, *, shape_grid=None, repeats: int = 2,
                  log=None) -> dict:
    """The shape knobs again, from a tuning's chosen setting (its quality
    knobs kept): a shape knob changes no result (every trial's outcome codes
    and final sparsifier must equal the chosen setting's), so the tuning's
    growth and kappa still hold and only the shape fields of `chosen` and its
    times change.  For a shape knob added after a tuning was made, without
    repeating the quality search and its kappa solves."""
    import copy

    base = setting_of(tuning)
    res = tune(run, None, edges_h0, anchor=base, quality_grid={}, shape_grid=shape_grid,
               repeats=repeats, log=log)
    growth = res["chosen"]["growth"]
    if growth != tuning["chosen"]["growth"]:
        raise TuneError(f"the chosen setting grows the sparsifier by {growth}, the tuning "
                        f"recorded {tuning['chosen']['growth']}: a different stream")
    out = copy.deepcopy(tuning)
    out["chosen"]["setting"] = res["chosen"]["setting"]
    out["chosen"]["seconds"] = res["chosen"]["seconds"]
    out["chosen"]["walk_ms"] = res["chosen"]["walk_ms"]
    out["shape_retune"] = {"from": base.asdict(), "base_seconds": res["anchor"]["seconds"],
                           "trials": res["trials"]}
    return out


def check_kappa(tr: StreamTrials, tuning: dict, method: str, log,
                which_settings=("anchor", "chosen")) -> dict:
    """kappa of the anchor's and the chosen setting's final graphs on the
    tuning's calibration stream, written into a copy of the tuning; each
    setting's pass must reproduce the growth the tuning recorded.
    `which_settings` limits the work to some of them (the others keep the
    kappa the tuning file holds, so a check can be split over runs)."""
    if tuning is None:
        raise TuneError("check-kappa: tuning file not found")
    if tr.updates != tuning["updates"]:
        raise TuneError(f"check-kappa: calibration stream of {tr.updates} updates, the tuning "
                        f"was made on {tuning['updates']}")
    out = dict(tuning)
    checks = []
    for which in ("anchor", "chosen"):
        s = setting_of(tuning, which)
        if which not in which_settings:
            checks.append({"quality": list(s.quality()), "kappa": tuning[which]["kappa"]})
            continue
        t0 = time.time()

        def progress(p, which=which, t0=t0):
            if p["solves"] % 5 == 0:
                log(f"kappa {which}: {p['solves']} solves, {p['cg_iterations']} CG iterations, "
                    f"{time.time() - t0:.0f} s")

        k = tr.kappa(s, progress=progress)
        t = tr.run(s)
        growth = t.edges_h - tr.edges_h0
        if growth != tuning[which]["growth"]:
            raise TuneError(f"check-kappa: {which} grew H by {growth}, the tuning recorded "
                            f"{tuning[which]['growth']}")
        out[which] = dict(tuning[which], kappa=k)
        checks.append({"quality": list(s.quality()), "kappa": k})
        log(f"kappa {which} {s.quality()}: {k:.6g} ({time.time() - t0:.0f} s, {method})")
    out["kappa_checks"] = checks
    out["kappa_method"] = method
    ka, kc = out["anchor"]["kappa"], out["chosen"]["kappa"]
    tol = tuning["tolerances"]["kappa"]
    if ka is not None and kc is not None:
        out["kappa_within_tolerance"] = abs(kc - ka) <= tol * abs(ka)
        log(f"chosen vs anchor kappa: {kc:.6g} vs {ka:.6g} "
            f"({'within' if out['kappa_within_tolerance'] else 'OUTSIDE'} {tol:.0%})")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="tune the stream form's walk settings")
    ap.add_argument("--workload", choices=["batched", "served"], default="batched")
    ap.add_argument("--requests", type=int, default=400,
                    help="served: calibration requests (the trace's first, applied in order)")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--tol-growth", type=float, default=0.02)
    ap.add_argument("--tol-kappa", type=float, default=0.02)
    ap.add_argument("--out", required=True)
    ap.add_argument("--check-kappa", metavar="TUNING",
                    help="instead of a search: kappa of the anchor and the chosen setting of "
                         "this tuning file, on its calibration stream (written to --out)")
    ap.add_argument("--kappa-of", choices=["anchor", "chosen", "both"], default="both",
                    help="check-kappa: which settings to compute (others keep the file's)")
    ap.add_argument("--retune-shapes", metavar="TUNING",
                    help="re-run only the shape knobs from this tuning's chosen setting "
                         "(its quality knobs, growth and kappa kept)")
    ap.add_argument("--kappa-method", choices=["lu", "pcg"], default=None,
                    help="how kappa applies L_H^-1 (default: lu for batched, pcg for served)")
    a = ap.parse_args(argv)
    method = a.kappa_method or ("lu" if a.workload == "batched" else "pcg")

    from . import synth
    from .core import Context

    def log(m):
        print(f"[autotune] {m}", flush=True)

    t0 = time.time()
    if a.workload == "batched":
        wl = synth.config_batched(a.scale, a.seed)
        batches, name = wl.batches, wl.name
    else:
        wl = synth.config_served(a.scale, a.seed)
        batches, name = wl.requests[:a.requests], wl.name
    updates = sum(len(b[1]) for b in batches)
    log(f"{name}: V={wl.G.V} E_G={len(wl.G)} E_H={len(wl.H)} batches={len(batches)} "
        f"updates={updates} ({time.time() - t0:.1f}s)")
    ctx = Context(0)
    tr = StreamTrials(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w),
                      batches, wl.threshold, a.seed, kappa_method=method)
    if a.retune_shapes:
        sg = SHAPE_GRID if a.workload == "batched" else \
            {k: v for k, v in SHAPE_GRID.items() if k != "graph"}
        res = retune_shapes(tr.run, tr.edges_h0, load(name, a.retune_shapes), shape_grid=sg,
                            repeats=a.repeats, log=log)
        res["shape_retune"]["command"] = ("python -m dygrass_amd.autotune "
                                          + " ".join(argv or sys.argv[1:]))
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
        log(f"chosen {res['chosen']['setting']}")
        ctx.close()
        return 0
    if a.check_kappa:
        which = ("anchor", "chosen") if a.kappa_of == "both" else (a.kappa_of,)
        res = check_kappa(tr, load(name, a.check_kappa), method, log, which)
        res["kappa_command"] = ("python -m dygrass_amd.autotune "
                                + " ".join(argv or sys.argv[1:]))
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
        ctx.close()
        return 0
    if a.workload == "batched":
        res = tune(tr.run, tr.kappa, tr.edges_h0, repeats=a.repeats, tol_growth=a.tol_growth,
                   tol_kappa=a.tol_kappa, log=log)
    else:
        # kappa on the served sparsifier (a tree plus random shortcuts) runs
        # through CG solves on the host (about 100 s per evaluation at 8M
        # vertices); step graphs are not a knob here (the served loop
        # launches directly)
        res = tune(tr.run, tr.kappa, tr.edges_h0, repeats=a.repeats, tol_growth=a.tol_growth,
                   tol_kappa=a.tol_kappa, log=log,
                   shape_grid={k: v for k, v in SHAPE_GRID.items() if k != "graph"})
        res["kappa_method"] = method
    res.update({"workload": name, "V": int(wl.G.V), "E_G": int(len(wl.G)),
                "E_H": int(len(wl.H)), "updates": int(updates),
                "batches": len(batches), "threshold": wl.threshold, "seed": a.seed,
                "command": "python -m dygrass_amd.autotune " + " ".join(argv or sys.argv[1:]),
                "seconds": time.time() - t0})
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    log(f"chosen {res['chosen']['setting']} ({res['chosen']['seconds'] * 1e3:.1f} ms vs anchor "
        f"{res['anchor']['seconds'] * 1e3:.1f} ms; kappa {res['chosen']['kappa']} vs "
        f"{res['anchor']['kappa']})")
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
