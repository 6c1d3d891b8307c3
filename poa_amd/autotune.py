"""Autotuner for the batched launch: the small tile's shape (threads per window workgroup and
DP columns per thread, i.e. wavefronts per window and tile width) and how many of the costliest
small-tile windows join the medium-tile launch. It times every candidate on a sample of
uneven windows drawn from a seed of its own (the same distribution as the 10,000-window
workload), checks that every candidate's outputs are identical, and writes the fastest to
poa_amd/tuned_gfx950.json, which ``api.Batch`` applies by default.

    python -m poa_amd.autotune [--windows 10000] [--repeat 2]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "tuned_gfx950.json")

SMALL_TILES = [(64, 2), (64, 4), (64, 8)]
MOVE_TOP = [0, 32, 64, 96, 128, 192, 256]


def load() -> dict | None:
    if not os.path.exists(TUNED):
        return None
    with open(TUNED) as f:
        return json.load(f)


def main() -> int:
    from . import api, sched, synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--windows", type=int, default=10000)
    ap.add_argument("--repeat", type=int, default=2)
    ap.add_argument("--out", default=TUNED)
    args = ap.parse_args()
    batch = synth.config3(seed=synth.SEEDS["train"] + 7, n_windows=args.windows)
    params = api.Params()
    ctx = api.Context(0)
    pred = None
    if os.path.exists(sched.MODEL):
        model = sched.Model.load()
        net = sched.Network(0)
        net.set_weights(model.weights)
        pred = sched.predict(model, net, params, batch.win_off, batch.seq, batch.seq_off)
        net.close()
    results, ref = [], None
    for (nt, cpl), top in itertools.product(SMALL_TILES, MOVE_TOP):
        b = api.Batch(ctx, params, batch.win_off, batch.seq, batch.seq_off, tuned=False)
        if pred is not None:
            b.set_predictions(pred)
        b.set_tile(nt, cpl)
        b.set_schedule(move_top=top)
        b.execute()  # warm-up (allocates the slot pools)
        ms = []
        for _ in range(args.repeat):
            t0 = time.perf_counter()
            b.execute()
            ms.append((time.perf_counter() - t0) * 1e3)
        texts = b.results().texts
        b.free()
        if ref is None:
            ref = texts
        same = texts == ref
        results.append({"threads": nt, "cols": cpl, "move_top": top,
                        "ms": round(float(np.median(ms)), 2), "identical": same})
        print(f"tile {nt}x{cpl} move_top {top}: {np.median(ms):.1f} ms"
              f"{'' if same else '  OUTPUT MISMATCH'}", flush=True)
    ok = [r for r in results if r["identical"]]
    best = min(ok, key=lambda r: r["ms"])
    out = {"device": "gfx950", "workload": f"{args.windows} uneven windows "
           f"(seed {synth.SEEDS['train'] + 7:#x})", "small_tile": [best["threads"], best["cols"]],
           "move_top": best["move_top"], "candidates": results}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(f"best: tile {best['threads']}x{best['cols']}, move_top {best['move_top']} "
          f"({best['ms']} ms); wrote {args.out}", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
