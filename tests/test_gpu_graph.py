"""The graph replay of the batched pass: every replay's output equals the eagerly enqueued
pass and the oracle, with eager device work and synchronisations between the replays; the
graph is built once and rebuilt only when a launch's arguments change; overflow re-runs after
a replay stay exact."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib
from poa_amd import api, sched, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _batch(ctx, b, graph):
    bb = api.Batch(ctx, api.Params(), b.win_off, b.seq, b.seq_off)
    bb.set_graph(graph)
    return bb


def test_replays_match_eager_and_oracle(ctx):
    b = synth.config3(seed=91, n_windows=300)
    sub = b.subset(range(0, 300, 15))
    want = oracle_lib.batch(api.Params(), sub.win_off, sub.seq, sub.seq_off)
    _say("graph: eager pass")
    eager = _batch(ctx, b, False)
    eager.execute()
    r0 = eager.results()
    assert r0.stats.graph_launches == 0
    assert [r0.texts[w] for w in range(0, 300, 15)] == want
    other = synth.config3(seed=92, n_windows=40)
    g = _batch(ctx, b, True)
    for rep in range(4):
        _say("graph: replay", rep)
        g.execute()
        r = g.results()
        assert r.texts == r0.texts
        assert r.stats.graph_launches == 1 and r.stats.graph_builds == 1
        assert r.stats.launches == r0.stats.launches
        # eager device work of the program's own between replays
        oe = _batch(ctx, other, False)
        oe.execute()
        oe.results()
        oe.free()
    # new launch arguments: one rebuild, same output
    _say("graph: rebuilt after a tile change")
    g.set_tile(64, 2)
    g.execute()
    r = g.results()
    assert r.texts == r0.texts and r.stats.graph_builds == 2
    g.free()
    eager.free()


def test_reruns_after_replay_stay_exact(ctx):
    b = synth.config3(seed=93, n_windows=24)
    want = oracle_lib.batch(api.Params(), b.win_off, b.seq, b.seq_off)
    low = sched.estimate_windows(api.Params(), b.win_off, b.seq_off).astype(np.float32)
    low[:, 1] *= 0.02
    low[:, 3] = 10
    g = _batch(ctx, b, True)
    g.set_predictions(low)
    for rep in range(2):
        _say("graph: underestimated, replay", rep)
        g.execute()
        r = g.results()
        assert r.texts == want
        assert r.stats.reruns >= 20 and r.stats.graph_launches == 1
    g.free()
