"""The learned-scheduling workload at full size: 100,000 uneven windows drawn as in the
10,000-window workload (5-60 reads, lognormal template length, 5-15% error) from a seed of
their own, scheduled by the trained network and run as a graph replay. Every window's output
is identical to the one the read-length-estimate schedule gives through eager launches (the
predictions never change a result), and every 10th window (10,000 of them) matches the oracle."""
from __future__ import annotations

import os

import pytest

import oracle_lib
from poa_amd import api, sched, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.timeout(900)
@pytest.mark.skipif(not os.path.exists(sched.MODEL), reason="no committed model")
def test_100k_windows_learned_equals_estimate_and_oracle(ctx):
    params = api.Params()
    _say("config5: generating 100,000 windows")
    b = synth.config5()
    n = b.n_windows
    model = sched.Model.load()
    net = sched.Network(0)
    net.set_weights(model.weights)
    _say("config5: predictions")
    pred = sched.predict(model, net, params, b.win_off, b.seq, b.seq_off)
    net.close()
    _say("config5: learned schedule, graph replay")
    g = api.Batch(ctx, params, b.win_off, b.seq, b.seq_off)
    g.set_predictions(pred)
    g.execute()
    r = g.results()
    g.free()
    assert r.stats.graph_launches == 1
    assert r.stats.reruns <= n // 100
    _say(f"config5: {r.stats.kernel_ms:.0f} ms, {r.stats.reruns} re-runs; read-length "
         "estimate, eager launches")
    h = api.Batch(ctx, params, b.win_off, b.seq, b.seq_off, tuned=False)
    h.set_graph(False)
    h.execute()
    rh = h.results()
    h.free()
    assert rh.stats.graph_launches == 0
    assert r.texts == rh.texts
    sample = range(0, n, 10)
    _say("config5: oracle on", len(sample), "windows")
    sub = b.subset(sample)
    want = oracle_lib.batch(params, sub.win_off, sub.seq, sub.seq_off,
                            threads=min(16, os.cpu_count() or 1))
    assert [r.texts[w] for w in sample] == want
