"""The autotuner's search rules on a synthetic model (host only).

The model: time falls with fewer walkers and shorter budgets, and with the
shape knobs at block_ins=4, block_del=2, dense_no_desc=True; the sparsifier
grows by 100 edges at the anchor, +60 when w_ins < 128, +1 when w_del < 256,
+10 more when w_del < 128; kappa is 10, or 10.5 when steps_del < 48.  Worked
by hand with the 2% bounds (growth limit 102):
  w_ins 256, 128 pass the screen, 64 grows by 160 -> kappa(128) = 10: w_ins = 128
  w_del 256 (100), 128 (101) pass, 64 grows by 111 -> kappa ok: w_del = 128
  steps_del 80, 64, 48, 32 pass the screen; kappa(32) = 10.5 misses (5%),
    kappa(48) = 10 holds: steps_del = 48
  steps_ins 80, 64, 48, 32 pass, kappa ok: steps_ins = 32
  shapes: dense_no_desc True, block_ins 4, block_del 2.
"""
import dataclasses

import pytest

from dygrass_amd.autotune import ANCHOR, Setting, Trial, TuneError, tune

H0 = 1000


def model_run(s: Setting, digest_of=None) -> Trial:
    sec = 1e-3 * (s.w_ins / 512 + 2 * s.w_del / 512 + s.steps_del / 100 + 0.5 * s.steps_ins / 100)
    sec += 1e-5 * (abs(s.block_ins - 4) + abs(s.block_del - 2)) + (0 if s.dense_no_desc else 5e-5)
    growth = 100 + (60 if s.w_ins < 128 else 0) + (1 if s.w_del < 256 else 0) + \
        (10 if s.w_del < 128 else 0)
    digest = (digest_of or (lambda x: str(x.quality())))(s)
    return Trial(sec, sec * 500, H0 + growth, digest)


def model_kappa(s: Setting) -> float:
    return 10.0 * (1.05 if s.steps_del < 48 else 1.0)


def test_search_follows_the_rules():
    calls = []

    def kappa(s):
        calls.append(s.quality())
        return model_kappa(s)

    res = tune(model_run, kappa, H0, repeats=1)
    chosen = Setting(**res["chosen"]["setting"])
    assert chosen == Setting(w_ins=128, w_del=128, steps_ins=32, steps_del=48, block_ins=4,
                             block_del=2, dense_no_desc=True)
    assert res["anchor"]["growth"] == 100 and res["chosen"]["growth"] == 101
    assert res["chosen"]["kappa"] == 10.0
    # kappa: anchor, w_ins=128, w_del=128, steps_del=32 (miss), steps_del=48, steps_ins=32
    assert calls == [(512, 512, 100, 100), (128, 512, 100, 100), (128, 128, 100, 100),
                     (128, 128, 100, 32), (128, 128, 100, 48), (128, 128, 32, 48)]
    rejected = [r for r in res["trials"] if not r["accepted"] and r["knob"] in ("w_ins", "w_del")]
    assert [r["setting"][r["knob"]] for r in rejected] == [64, 64]
    assert all("growth" in r["reason"] for r in rejected)


def test_nothing_passes_keeps_the_anchor():
    def run(s):
        t = model_run(s)
        return dataclasses.replace(t, edges_h=t.edges_h + (0 if s.quality() == ANCHOR.quality()
                                                          else 50))

    res = tune(run, model_kappa, H0, repeats=1, shape_grid={})
    assert Setting(**res["chosen"]["setting"]) == ANCHOR


def test_slower_candidate_is_not_taken():
    def run(s):
        t = model_run(s)
        slow = 1.0 if s.w_del == 256 else 0.0
        return dataclasses.replace(t, seconds=t.seconds + slow, walk_ms=t.walk_ms + 1e3 * slow)

    res = tune(run, model_kappa, H0, repeats=1, shape_grid={})
    assert res["chosen"]["setting"]["w_del"] == 512  # the first step down was slower: stop


def test_shape_knob_that_changes_results_is_an_error():
    def run(s):
        return model_run(s, digest_of=lambda x: f"{x.quality()}{x.block_del}")

    with pytest.raises(TuneError, match="block_del"):
        tune(run, model_kappa, H0, repeats=1,
             quality_grid={}, shape_grid={"block_del": (0, 2)})


def test_nondeterministic_pass_is_an_error():
    n = [0]

    def run(s):
        n[0] += 1
        return model_run(s, digest_of=lambda x: str(n[0]))

    with pytest.raises(TuneError, match="not deterministic"):
        tune(run, model_kappa, H0, repeats=2)


def test_shape_only_tuning_needs_no_kappa():
    anchor = Setting(w_ins=256, w_del=512)
    res = tune(model_run, None, H0, anchor=anchor, quality_grid={}, repeats=1)
    assert Setting(**res["chosen"]["setting"]) == Setting(w_ins=256, w_del=512, block_ins=4,
                                                          block_del=2, dense_no_desc=True)
    assert res["anchor"]["kappa"] is None and res["kappa_checks"] == []


def test_committed_tunings_hold_their_bounds():
    """Every committed tuning file: the chosen setting is a Setting, its growth
    and kappa satisfy the recorded tolerances against the anchor (when kappa
    was measured), every kappa check of an accepted quality step is within
    bounds, and every shape trial (results must not move) shows the chosen
    setting's growth exactly."""
    import glob
    import json
    import os

    from dygrass_amd.autotune import SHAPE_GRID, load, setting_of

    root = os.path.join(os.path.dirname(__file__), "..", "dygrass_amd", "tuning")
    files = sorted(glob.glob(os.path.join(root, "*.json")))
    assert files
    for path in files:
        t = json.load(open(path))
        name = os.path.basename(path)[:-5]
        assert load(name) == t and t["workload"] == name
        chosen = setting_of(t)
        a, c, tol = t["anchor"], t["chosen"], t["tolerances"]
        if a["kappa"] is not None:
            assert c["growth"] <= a["growth"] * (1 + tol["growth"]), name
            assert abs(c["kappa"] - a["kappa"]) <= tol["kappa"] * abs(a["kappa"]), name
            chosen_q = [chosen.w_ins, chosen.w_del, chosen.steps_ins, chosen.steps_del]
            assert any(k["quality"] == chosen_q for k in t["kappa_checks"]), name
        else:  # growth screen only (or a shape-only tuning)
            assert c["kappa"] is None and t["kappa_checks"] == [], name
            assert c["growth"] <= a["growth"] * (1 + tol["growth"]), name
        shapes = [r for r in t["trials"] if r["knob"] in SHAPE_GRID]
        assert shapes, name
        assert all(r["growth"] == c["growth"] for r in shapes), name
        assert c["seconds"] <= a["seconds"] * 1.05, name


def test_growth_only_tuning():
    """kappa None with quality knobs: the growth screen alone decides; by hand,
    w_ins 128 and w_del 128 pass (64 fails each), every budget passes."""
    res = tune(model_run, None, H0, repeats=1, shape_grid={})
    assert Setting(**res["chosen"]["setting"]) == Setting(w_ins=128, w_del=128, steps_ins=32,
                                                          steps_del=32)
    assert res["chosen"]["kappa"] is None and res["kappa_checks"] == []


def test_retune_shapes_keeps_quality_growth_and_kappa():
    """A shape knob added after a tuning: only the shape knobs run again, from
    the tuning's chosen setting; its quality knobs, growth and kappa stay, and
    a tuning whose growth the stream does not reproduce is refused."""
    import copy

    from dygrass_amd.autotune import retune_shapes

    res = tune(model_run, model_kappa, H0, repeats=1)
    old = copy.deepcopy(res)
    old["chosen"]["setting"]["block_ins"] = 0  # an older shape choice

    def run(s):  # the newer knob is faster in this model
        t = model_run(s)
        return dataclasses.replace(t, seconds=t.seconds - (1e-6 if s.small_1024 else 0.0))

    new = retune_shapes(run, H0, old, repeats=1)
    chosen = Setting(**new["chosen"]["setting"])
    assert chosen.quality() == Setting(**old["chosen"]["setting"]).quality()
    assert chosen.block_ins == 4 and chosen.small_1024 and chosen.dense_no_desc
    assert new["chosen"]["kappa"] == old["chosen"]["kappa"]
    assert new["chosen"]["growth"] == old["chosen"]["growth"]
    assert new["anchor"] == old["anchor"]
    assert new["shape_retune"]["from"] == old["chosen"]["setting"]
    bad = copy.deepcopy(old)
    bad["chosen"]["growth"] += 1
    with pytest.raises(TuneError, match="different stream"):
        retune_shapes(run, H0, bad, repeats=1)
