"""Learned scheduling: the network that predicts each DP call's band extent and cell count.

Pipeline:
  features (C++, mpoa_sched_features)  ->  transform (log / standardise)  ->
  network on the GPU (mpoa_mlp_forward, FP16 weights, FP32 accumulation)  ->
  per-window aggregation (cost, costliest call, widest row, node count) with a calibrated
  safety margin  ->  mpoa_batch_set_predictions.

Training (``python -m poa_amd.sched train``) runs the plain method over windows drawn as in the
uneven-window workload from a seed of their own, records every DP call's band extent and cells
(mpoa_batch_traces), and fits the network on the GPU with the library's own forward/backward
kernels (mpoa_mlp_train). The model file holds the FP16 weights, the feature/target
normalisation and the margins. Predictions never change a result: a window that outgrows a
predicted capacity is re-run on the device at worst-case capacities.
"""
from __future__ import annotations

import argparse
import ctypes
import os
import time

import numpy as np

from . import api

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "sched_mlp.npz")
N_FEAT, N_OUT = 8, 2
LOG_COLS = (0, 1, 2, 4, 5)  # counts and lengths enter as log1p


def _lib():
    L = api.lib()
    if not getattr(L, "_sched_bound", False):
        P, I32, I64, VP, F = ctypes.POINTER, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_float
        L.mpoa_sched_features.argtypes = [P(api.Params), I32, VP, VP, VP, VP, I64, P(I64)]
        L.mpoa_mlp_create.argtypes = [ctypes.c_int, ctypes.c_uint64, P(VP)]
        L.mpoa_mlp_destroy.argtypes = [VP]
        L.mpoa_mlp_destroy.restype = None
        L.mpoa_mlp_n_params.restype = I32
        L.mpoa_mlp_get_weights.argtypes = [VP, VP, I32]
        L.mpoa_mlp_set_weights.argtypes = [VP, VP, I32]
        L.mpoa_mlp_forward.argtypes = [VP, VP, I64, VP]
        L.mpoa_mlp_gradients.argtypes = [VP, VP, VP, I64, VP, P(F)]
        L.mpoa_mlp_train.argtypes = [VP, VP, VP, I64, I32, I32, F, ctypes.c_uint64, VP]
        L.mpoa_mlp_error.restype = ctypes.c_char_p
        L._sched_bound = True
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise api.MpoaError(rc, _lib().mpoa_mlp_error().decode() or "scheduling call failed")


def features(params: api.Params, win_off, seq, seq_off) -> np.ndarray:
    """Raw features, one row per DP call in trace order (window, read 1..n-1)."""
    L = _lib()
    win_off = np.ascontiguousarray(win_off, np.int32)
    seq = np.ascontiguousarray(seq, np.uint8)
    seq_off = np.ascontiguousarray(seq_off, np.int64)
    nw = len(win_off) - 1
    n = ctypes.c_int64()
    _check(L.mpoa_sched_features(ctypes.byref(params), nw, _p(win_off), _p(seq), _p(seq_off),
                                 None, 0, ctypes.byref(n)))
    X = np.zeros((max(n.value, 1), N_FEAT), np.float32)
    _check(L.mpoa_sched_features(ctypes.byref(params), nw, _p(win_off), _p(seq), _p(seq_off),
                                 _p(X), n.value, ctypes.byref(n)))
    return X[:n.value]


def transform(X: np.ndarray) -> np.ndarray:
    Z = X.astype(np.float64).copy()
    for c in LOG_COLS:
        Z[:, c] = np.log1p(Z[:, c])
    return Z


def targets(tr: np.ndarray) -> np.ndarray:
    """log band extent and log cells of each traced call."""
    return np.stack([np.log(np.maximum(tr["width"], 1)), np.log(np.maximum(tr["cells"], 1))],
                    axis=1)


class Network:
    """The device network (mpoa_mlp_*)."""

    def __init__(self, device: int = 0, seed: int = 1):
        h = ctypes.c_void_p()
        _check(_lib().mpoa_mlp_create(device, seed, ctypes.byref(h)))
        self.h = h
        self.n_params = _lib().mpoa_mlp_n_params()

    def close(self):
        if self.h:
            _lib().mpoa_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.n_params, np.uint16)
        _check(_lib().mpoa_mlp_get_weights(self.h, _p(w), self.n_params))
        return w

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.uint16)
        _check(_lib().mpoa_mlp_set_weights(self.h, _p(w), len(w)))

    def forward(self, Z: np.ndarray) -> np.ndarray:
        Z = np.ascontiguousarray(Z, np.float32)
        Y = np.zeros((len(Z), N_OUT), np.float32)
        _check(_lib().mpoa_mlp_forward(self.h, _p(Z), len(Z), _p(Y)))
        return Y

    def gradients(self, Z: np.ndarray, T: np.ndarray):
        Z = np.ascontiguousarray(Z, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        _check(_lib().mpoa_mlp_gradients(self.h, _p(Z), _p(T), len(Z), _p(g), ctypes.byref(loss)))
        return g, loss.value

    def train(self, Z, T, epochs=40, batch=1024, lr=2e-3, seed=7) -> np.ndarray:
        Z = np.ascontiguousarray(Z, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        losses = np.zeros(epochs, np.float32)
        _check(_lib().mpoa_mlp_train(self.h, _p(Z), _p(T), len(Z), epochs, batch, lr, seed,
                                     _p(losses)))
        return losses


class Model:
    """Weights plus normalisation and margins (the model file)."""

    def __init__(self, weights, x_mu, x_sd, y_mu, y_sd, margin_cells, margin_nodes, meta=""):
        self.weights = np.asarray(weights, np.uint16)
        self.x_mu, self.x_sd = np.asarray(x_mu), np.asarray(x_sd)
        self.y_mu, self.y_sd = np.asarray(y_mu), np.asarray(y_sd)
        self.margin_cells = float(margin_cells)
        self.margin_nodes = float(margin_nodes)
        self.meta = str(meta)

    def save(self, path: str = MODEL) -> None:
        np.savez(path, weights=self.weights, x_mu=self.x_mu, x_sd=self.x_sd, y_mu=self.y_mu,
                 y_sd=self.y_sd, margin_cells=self.margin_cells, margin_nodes=self.margin_nodes,
                 meta=np.array(self.meta))

    @staticmethod
    def load(path: str = MODEL) -> "Model":
        with np.load(path, allow_pickle=False) as z:
            return Model(z["weights"], z["x_mu"], z["x_sd"], z["y_mu"], z["y_sd"],
                         z["margin_cells"], z["margin_nodes"], str(z["meta"]))

    def norm_x(self, X: np.ndarray) -> np.ndarray:
        return ((transform(X) - self.x_mu) / self.x_sd).astype(np.float32)

    def denorm_y(self, Y: np.ndarray) -> np.ndarray:
        return Y.astype(np.float64) * self.y_sd + self.y_mu


def calls_per_window(win_off) -> np.ndarray:
    return np.maximum(np.diff(np.asarray(win_off, np.int64)) - 1, 0)


def aggregate(logy: np.ndarray, win_off, seq_off, margin_cells=1.0, margin_nodes=1.0):
    """Per-window (cost, cells of the costliest call, widest row, final nodes) from per-call
    log predictions (trace order)."""
    win_off = np.asarray(win_off, np.int64)
    seq_off = np.asarray(seq_off, np.int64)
    nw = len(win_off) - 1
    per = calls_per_window(win_off)
    width = np.exp(logy[:, 0])
    cells = np.exp(logy[:, 1])
    rows = cells / np.maximum(width, 1.0)
    out = np.zeros((nw, 4), np.float64)
    qlen = np.diff(seq_off)
    last_q = qlen[np.maximum(win_off[1:] - 1, 0)] if nw else np.zeros(0)
    starts = np.concatenate([[0], np.cumsum(per)])[:-1]
    has = per > 0
    if len(cells):
        idx = starts[has]
        out[has, 0] = np.add.reduceat(cells, idx)
        out[has, 1] = np.maximum.reduceat(cells, idx)
        out[has, 2] = np.maximum.reduceat(width, idx)
        out[has, 3] = np.maximum.reduceat(rows, idx)
    out[:, 1] *= margin_cells
    out[:, 3] = out[:, 3] * margin_nodes + last_q
    return out.astype(np.float32)


def predict(model: Model, net: Network, params: api.Params, win_off, seq, seq_off):
    """Per-window predictions for mpoa_batch_set_predictions."""
    X = features(params, win_off, seq, seq_off)
    logy = model.denorm_y(net.forward(model.norm_x(X)))
    return aggregate(logy, win_off, seq_off, model.margin_cells, model.margin_nodes)


def window_truth(tr: np.ndarray, win_off) -> np.ndarray:
    """Per-window (cost, costliest call cells, widest row, rows of the biggest call) from
    traces, the quantities `aggregate` predicts."""
    per = calls_per_window(win_off)
    nw = len(per)
    out = np.zeros((nw, 4), np.float64)
    starts = np.concatenate([[0], np.cumsum(per)])[:-1]
    has = per > 0
    c = tr["cells"].astype(np.float64)
    idx = starts[has]
    out[has, 0] = np.add.reduceat(c, idx)
    out[has, 1] = np.maximum.reduceat(c, idx)
    out[has, 2] = np.maximum.reduceat(tr["width"].astype(np.float64), idx)
    out[has, 3] = np.maximum.reduceat(tr["rows"].astype(np.float64), idx)
    return out


def estimate_windows(params: api.Params, win_off, seq_off) -> np.ndarray:
    """The library's read-length estimate per window (used when no predictions are given,
    poa_amd/csrc/mpoa.cpp estimate()): (cost, costliest-call DP cells, band width, nodes)."""
    win_off = np.asarray(win_off, np.int64)
    lens = np.diff(np.asarray(seq_off, np.int64))
    nw = len(win_off) - 1
    out = np.zeros((nw, 4))
    for w in range(nw):
        q = lens[win_off[w]:win_off[w + 1]]
        qmax, qtot, nr = int(q.max()), int(q.sum()), len(q)
        n = min(2 + qmax + int(np.ceil(0.16 * (qtot - qmax))), 2 + qtot)
        m = qmax + 1
        full = (m // 8 + 2) * 8
        width = min(2 * (params.band_b + m // params.band_f) + 64, full) if params.band_f > 0 else full
        out[w] = (nr * n * width, n * width, width, n)
    return out


def collect(ctx: api.Context, params: api.Params, batch, log=print):
    """Runs the plain method over `batch` and returns (raw features, traces)."""
    X = features(params, batch.win_off, batch.seq, batch.seq_off)
    b = api.Batch(ctx, params, batch.win_off, batch.seq, batch.seq_off)
    b.set_traces(True)
    t0 = time.perf_counter()
    b.execute()
    log(f"traced {batch.n_windows} windows in {time.perf_counter() - t0:.2f} s")
    tr = b.traces()
    b.free()
    assert len(tr) == len(X)
    return X, tr


def fit(X: np.ndarray, tr: np.ndarray, win_off, seq_off, epochs=40, device=0, log=print,
        holdout=0.1, coverage=0.995):
    """Trains a Model on the GPU. The margins are the `coverage` quantiles of true/predicted
    on held-out windows (costliest-call cells, node count)."""
    Z = transform(X)
    T = targets(tr)
    nw = len(win_off) - 1
    per = calls_per_window(win_off)
    win_of_call = np.repeat(np.arange(nw), per)
    n_hold = max(1, int(nw * holdout))
    hold = win_of_call >= nw - n_hold
    x_mu, x_sd = Z[~hold].mean(0), Z[~hold].std(0) + 1e-6
    y_mu, y_sd = T[~hold].mean(0), T[~hold].std(0) + 1e-6
    Zn = ((Z - x_mu) / x_sd).astype(np.float32)
    Tn = ((T - y_mu) / y_sd).astype(np.float32)
    net = Network(device, seed=12345)
    t0 = time.perf_counter()
    losses = net.train(Zn[~hold], Tn[~hold], epochs=epochs)
    log(f"trained {int((~hold).sum())} calls x {epochs} epochs in {time.perf_counter() - t0:.2f} s; "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    m = Model(net.get_weights(), x_mu, x_sd, y_mu, y_sd, 1.0, 1.0)
    logy = m.denorm_y(net.forward(Zn[hold]))
    wo = np.asarray(win_off, np.int64)
    sub_off = wo[nw - n_hold:] - wo[nw - n_hold]
    sub_seq_off = np.asarray(seq_off, np.int64)[wo[nw - n_hold]:] - np.asarray(seq_off)[wo[nw - n_hold]]
    pred = aggregate(logy, sub_off, sub_seq_off).astype(np.float64)
    truth = window_truth(tr[hold], sub_off)
    ok = truth[:, 1] > 0
    r_cells = truth[ok, 1] / np.maximum(pred[ok, 1], 1)
    last_q = np.diff(sub_seq_off)[np.maximum(sub_off[1:] - 1, 0)]
    r_nodes = truth[ok, 3] / np.maximum(pred[ok, 3] - last_q[ok], 1)
    m.margin_cells = float(np.quantile(r_cells, coverage))
    m.margin_nodes = float(np.quantile(r_nodes, coverage))
    err = np.abs(logy - T[hold])
    log(f"held-out calls: mean |log error| width {err[:, 0].mean():.4f}, cells {err[:, 1].mean():.4f}; "
        f"margins cells x{m.margin_cells:.3f}, nodes x{m.margin_nodes:.3f}")
    net.close()
    return m, losses


def main() -> int:
    from . import synth
    ap = argparse.ArgumentParser(description="train the scheduling network on the GPU")
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--windows", type=int, default=6500, help="~31 DP calls per window")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--out", default=MODEL)
    ap.add_argument("--holdout-out", default=None)
    ap.add_argument("--holdout-windows", type=int, default=1000)
    args = ap.parse_args()
    params = api.Params()
    batch = synth.config3(seed=synth.SEEDS["train"], n_windows=args.windows)
    ctx = api.Context(0)
    X, tr = collect(ctx, params, batch)
    print(f"{len(tr)} traces", flush=True)
    m, losses = fit(X, tr, batch.win_off, batch.seq_off, epochs=args.epochs)
    m.meta = (f"trained on {len(tr)} DP-call traces of {args.windows} windows "
              f"(uneven-window workload, seed {synth.SEEDS['train']:#x}), {args.epochs} epochs")
    m.save(args.out)
    print(f"saved {args.out}", flush=True)
    if args.holdout_out:
        # held-out traces of windows from a seed of their own, for the CPU gain test
        hb = synth.config3(seed=synth.SEEDS["train"] + 1, n_windows=args.holdout_windows)
        hX, htr = collect(ctx, params, hb)
        np.savez_compressed(args.holdout_out, X=hX, truth=window_truth(htr, hb.win_off),
                            win_off=hb.win_off.astype(np.int64),
                            seq_off=hb.seq_off.astype(np.int64),
                            estimate=estimate_windows(params, hb.win_off, hb.seq_off))
        print(f"saved {args.holdout_out}", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
