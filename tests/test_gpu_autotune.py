"""The autotuner on the device, on a small 12-NN stream (40k points, 64
batches of up to 2000 updates): every shape knob must leave the outcome
codes and the final sparsifier unchanged (the tuner raises otherwise), and
the chosen setting must keep the anchor's growth and kappa bounds."""
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_small_stream(ctx):
    from dygrass_amd import synth
    from dygrass_amd.autotune import StreamTrials, setting_of, tune

    wl = synth.config_batched(0.01, 0)
    tr = StreamTrials(ctx, wl.G.V, (wl.G.u, wl.G.v, wl.G.w), (wl.H.u, wl.H.v, wl.H.w),
                      wl.batches, wl.threshold, 0)
    res = tune(tr.run, tr.kappa, tr.edges_h0, repeats=1,
               quality_grid={"w_ins": (512, 256, 64), "steps_del": (100, 48)},
               shape_grid={"dense_no_desc": (False, True), "block_ins": (0, 3),
                           "block_del": (0, 5)})
    a, c = res["anchor"], res["chosen"]
    assert c["growth"] <= max(a["growth"] * 1.02, 0.02 * tr.edges_h0)
    assert abs(c["kappa"] - a["kappa"]) <= 0.02 * a["kappa"]
    s = setting_of(res)
    assert s.w_ins in (512, 256, 64) and s.steps_del in (100, 48)
    # every shape trial ran (and matched the reference digest, or tune() raised)
    shapes = [r for r in res["trials"] if r["knob"] in ("dense_no_desc", "block_ins", "block_del")]
    assert len(shapes) == 3
