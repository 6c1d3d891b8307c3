"""The scheduling network's device kernels against the float64 numpy checker (forward, loss,
every gradient), training on the device, and the rule that predictions never change a result:
mispredicted capacities overflow on the device and the window is re-run exactly."""
from __future__ import annotations

import numpy as np
import pytest

import mlp_ref
import oracle_lib
from poa_amd import api, sched, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def test_forward_and_gradients_match_cpu_checker():
    rng = np.random.default_rng(11)
    net = sched.Network(0, seed=99)
    P = mlp_ref.unpack(net.get_weights())
    for n in (1, 63, 64, 200, 1000):
        _say("mlp gradients", n)
        Z = rng.standard_normal((n, 8)).astype(np.float32)
        T = rng.standard_normal((n, 2)).astype(np.float32)
        y_ref, _ = mlp_ref.forward(P, Z.astype(np.float64))
        y = net.forward(Z)
        # FP32 accumulation of FP16-valued weights against float64
        assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-4)
        g, loss = net.gradients(Z, T)
        L_ref, g_ref = mlp_ref.loss_and_grad(P, Z.astype(np.float64), T.astype(np.float64))
        assert abs(loss - L_ref) <= 1e-4 * max(1.0, L_ref)
        scale = np.abs(g_ref).max()
        assert np.max(np.abs(g - g_ref)) <= 1e-4 * scale


def test_training_on_device_fits_a_smooth_function():
    rng = np.random.default_rng(12)
    Z = rng.standard_normal((8192, 8)).astype(np.float32)
    T = np.stack([np.sin(Z[:, 0]) + 0.5 * Z[:, 1], Z[:, 2] * Z[:, 3] * 0.5], 1).astype(np.float32)
    net = sched.Network(0, seed=5)
    _say("mlp training")
    losses = net.train(Z, T, epochs=30, batch=512, lr=3e-3)
    assert losses[-1] < 0.2 * losses[0]
    # the trained FP16 weights give the same predictions through the checker
    P = mlp_ref.unpack(net.get_weights())
    y_ref, _ = mlp_ref.forward(P, Z[:100].astype(np.float64))
    assert np.allclose(net.forward(Z[:100]), y_ref, rtol=1e-4, atol=1e-4)


def _texts(ctx, b, pred=None, mode=api.MODE_BATCHED):
    bb = api.Batch(ctx, api.Params(), b.win_off, b.seq, b.seq_off)
    if pred is not None:
        bb.set_predictions(pred)
    bb.set_mode(mode)
    bb.execute()
    r = bb.results()
    bb.free()
    return r.texts, r.stats


def test_mispredictions_rerun_exactly(ctx):
    b = synth.config3(seed=77, n_windows=24)
    want = oracle_lib.batch(api.Params(), b.win_off, b.seq, b.seq_off)
    _say("heuristic")
    t0, s0 = _texts(ctx, b)
    assert t0 == want and s0.reruns == 0
    est = sched.estimate_windows(api.Params(), b.win_off, b.seq_off).astype(np.float32)
    # Slots are sized by the largest prediction of their launch, so every window is
    # underestimated: far too small DP arenas and graph capacities.
    low = est.copy()
    low[:, 1] *= 0.02
    low[:, 3] = 10
    _say("underestimated")
    t1, s1 = _texts(ctx, b, low)
    assert t1 == want
    assert s1.reruns >= 20


def test_plain_form_matches(ctx):
    b = synth.config3(seed=78, n_windows=6)
    want = oracle_lib.batch(api.Params(), b.win_off, b.seq, b.seq_off)
    _say("plain form")
    t, st = _texts(ctx, b, mode=api.MODE_PLAIN)
    assert t == want
    assert st.launches == 6


@pytest.mark.skipif(not __import__("os").path.exists(sched.MODEL), reason="no committed model")
def test_learned_predictions_keep_results(ctx):
    b = synth.config3(seed=79, n_windows=64)
    m = sched.Model.load()
    net = sched.Network(0)
    net.set_weights(m.weights)
    pred = sched.predict(m, net, api.Params(), b.win_off, b.seq, b.seq_off)
    want = oracle_lib.batch(api.Params(), b.win_off, b.seq, b.seq_off)
    _say("learned predictions")
    t, st = _texts(ctx, b, pred)
    assert t == want
    assert st.reruns <= 4
