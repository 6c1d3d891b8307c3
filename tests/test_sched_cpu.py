"""Host side of the learned scheduling, no GPU: the feature extractor, the numpy checker of
the network (against finite differences), aggregation, and what the committed model gains
over the read-length estimate on held-out traces (costliest-call DP sizing and the makespan of
a longest-first queue)."""
from __future__ import annotations

import os

import numpy as np
import pytest

import mlp_ref
from poa_amd import api, sched, synth

HOLDOUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "sched_holdout.npz")


def test_features_shape_and_sanity():
    b = synth.config3(seed=321, n_windows=30)
    X = sched.features(api.Params(), b.win_off, b.seq, b.seq_off)
    assert X.shape == (int(sched.calls_per_window(b.win_off).sum()), 8)
    assert np.all(np.isfinite(X))
    lens = np.diff(b.seq_off)
    first = np.concatenate([[0], np.cumsum(sched.calls_per_window(b.win_off))])[:-1]
    # feature 0 of call (w, r=1) is the length of read 1 of window w
    assert np.array_equal(X[first, 0], lens[b.win_off[:-1] + 1].astype(np.float32))
    assert np.all((X[:, 3] >= 0) & (X[:, 3] <= 1)) and np.all(X[:, 7] == 4)
    # reads of one template share most minimizers with the first read
    assert np.median(X[:, 3]) > 0.05
    assert np.all((X[:, 6] > 0.0) & (X[:, 6] < 0.3))
    # the same reads give the same features
    assert np.array_equal(X, sched.features(api.Params(), b.win_off, b.seq, b.seq_off))


def test_mlp_ref_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w16 = (rng.standard_normal(mlp_ref.N_PARAM) * 0.2).astype(np.float16).view(np.uint16)
    P = mlp_ref.unpack(w16)
    X = rng.standard_normal((40, 8))
    T = rng.standard_normal((40, 2))
    L, g = mlp_ref.loss_and_grad(P, X, T)
    flat = np.concatenate([P[n].ravel() for n, _ in mlp_ref.SHAPES])
    eps = 1e-6
    for k in rng.choice(len(flat), 60, replace=False):
        f2 = flat.copy()
        f2[k] += eps
        o, Q = 0, {}
        for name, shape in mlp_ref.SHAPES:
            n = int(np.prod(shape))
            Q[name] = f2[o:o + n].reshape(shape)
            o += n
        L2, _ = mlp_ref.loss_and_grad(Q, X, T)
        assert abs((L2 - L) / eps - g[k]) < 1e-4 * max(1.0, abs(g[k]))


def test_aggregate():
    win_off = np.array([0, 3, 4, 7], np.int64)          # windows of 3, 1 and 3 reads
    seq_off = np.arange(0, 8 * 10, 10, dtype=np.int64)  # 7 reads of 10 bases
    logy = np.log(np.array([[100, 1000], [120, 3000], [90, 900], [200, 4000]], float))
    out = sched.aggregate(logy, win_off, seq_off)
    assert out.shape == (3, 4)
    assert np.allclose(out[0], [4000, 3000, 120, 25 + 10])
    assert np.allclose(out[1], [0, 0, 0, 10])
    assert np.allclose(out[2, :3], [4900, 4000, 200])


def _makespan(cost_true: np.ndarray, order: np.ndarray, slots: int) -> float:
    import heapq
    h = [0.0] * slots
    for w in order:
        t = heapq.heappop(h)
        heapq.heappush(h, t + cost_true[w])
    return max(h)


@pytest.mark.skipif(not (os.path.exists(HOLDOUT) and os.path.exists(sched.MODEL)),
                    reason="no committed model or held-out traces")
def test_trained_model_gains_over_estimate():
    """Held-out windows (their own seed): the network's per-window predictions against the
    read-length estimate the library uses without it."""
    z = np.load(HOLDOUT, allow_pickle=False)
    X, truth, win_off, seq_off = z["X"], z["truth"], z["win_off"], z["seq_off"]
    est = z["estimate"]
    m = sched.Model.load()
    P = mlp_ref.unpack(m.weights)
    y, _ = mlp_ref.forward(P, m.norm_x(X).astype(np.float64))
    pred = sched.aggregate(m.denorm_y(y), win_off, seq_off, m.margin_cells, m.margin_nodes)
    ok = truth[:, 1] > 0
    # DP arena sizing: capacity / need of the costliest call, and how often it would overflow
    over_est = est[ok, 1] / truth[ok, 1]
    over_mlp = pred[ok, 1] / truth[ok, 1]
    miss = float(np.mean(over_mlp < 1.0))
    assert miss < 0.02
    assert np.median(over_mlp) < 0.8 * np.median(over_est)
    # cost ranking: Spearman correlation with the true cost
    def rank(a):
        return np.argsort(np.argsort(a))
    r_mlp = np.corrcoef(rank(pred[:, 0]), rank(truth[:, 0]))[0, 1]
    r_est = np.corrcoef(rank(est[:, 0]), rank(truth[:, 0]))[0, 1]
    assert r_mlp > r_est
    # longest-first queue on 1/6 as many slots as windows (the 10,000-window run's ratio)
    slots = max(1, len(truth) // 6)
    t_mlp = _makespan(truth[:, 0], np.argsort(-pred[:, 0], kind="stable"), slots)
    t_est = _makespan(truth[:, 0], np.argsort(-est[:, 0], kind="stable"), slots)
    t_fifo = _makespan(truth[:, 0], np.arange(len(truth)), slots)
    assert t_mlp <= t_est * 1.001 and t_mlp < t_fifo
