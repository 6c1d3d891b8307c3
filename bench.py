#!/usr/bin/env python3
"""Benchmark: batched uneven POA windows on one MI355X, bit-exact to minipoa.

Workload (default): 10,000 independent POA jobs, each 5-60 reads over a template whose length
is lognormal (median 500 bp, sigma 0.5, clipped to 150-2,000 bp), per-job error uniform in
5-15% (sub:ins:del 1:1:1), minipoa's default scoring and band (-M2 -X-4 -O-4 -E-2 -b100
-f40), consensus output. A step is one pass of the whole per-window pipeline (DP, backtrack,
add-to-graph, re-sort, consensus) over all windows, inputs resident in HBM.

Prints one JSON line: windows/s (value), GCUPS, the dominant kernel's roofline figures and the
upstream minipoa binary timed on the host cores over a bounded sample of the same windows.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "minipoa")


def cpu_baseline(batch, gpu_texts, sample: int, cores: int) -> dict:
    """Upstream minipoa (oracle/_ref/minipoa, built from its own sources) on `sample` windows
    spread over the batch, `cores` processes at a time; the CPU oracle if the binary is absent.
    Also checks the GPU outputs of those windows against the CPU outputs."""
    n = batch.n_windows
    idx = np.linspace(0, n - 1, num=min(sample, n)).round().astype(int)
    idx = np.unique(idx)
    if os.path.exists(REF_BIN):
        tmp = tempfile.mkdtemp(prefix="mpoa_cpu_")
        try:
            paths = []
            for w in idx:
                pth = os.path.join(tmp, f"w{w}.fa")
                batch.window(int(w)).to_fasta(pth)
                paths.append(pth)

            def run(pth):
                return subprocess.run([REF_BIN, pth], check=True, capture_output=True).stdout

            t0 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as ex:
                outs = list(ex.map(run, paths))
            wall = time.perf_counter() - t0
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        texts = [o.decode() for o in outs]
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        from poa_amd import api
        sub = batch.subset(idx)
        t0 = time.perf_counter()
        texts = oracle_lib.batch(api.Params(), sub.win_off, sub.seq, sub.seq_off, threads=cores)
        wall = time.perf_counter() - t0
        kind = "port"
    same = all(texts[i] == gpu_texts[int(w)] for i, w in enumerate(idx))
    return {
        "value": round(len(idx) / wall, 2),
        "unit": "windows/s",
        "cores": cores,
        "kind": kind,
        "sample": f"{len(idx)} of the {n} windows (evenly spaced), one minipoa process per "
                  f"window, {cores} at a time; wall {wall:.2f} s",
        "outputs_identical_to_gpu": bool(same),
    }


def plain_form(ctx, params, batch, sample: int, gpu_texts, batched_wps: float, log) -> dict:
    """The plain GPU form (one launch of one workgroup per window, windows one after another)
    on `sample` evenly spaced windows of the same batch; its rate against the batched one."""
    from poa_amd import api
    n = batch.n_windows
    idx = np.unique(np.linspace(0, n - 1, num=min(sample, n)).round().astype(int))
    sub = batch.subset(idx)
    pb = api.Batch(ctx, params, sub.win_off, sub.seq, sub.seq_off)
    pb.set_mode(api.MODE_PLAIN)
    t0 = time.perf_counter()
    pb.execute()
    wall = time.perf_counter() - t0
    texts = pb.results().texts
    pb.free()
    wps = len(idx) / wall
    log(f"plain form: {len(idx)} windows in {wall:.2f} s = {wps:.1f} windows/s")
    return {"value": round(wps, 2), "unit": "windows/s", "windows": int(len(idx)),
            "sample": "evenly spaced windows of the same batch",
            "batched_over_plain": round(batched_wps / wps, 2),
            "outputs_identical": all(texts[i] == gpu_texts[int(w)] for i, w in enumerate(idx))}


SINGLE = [
    # (name, generator, flags, workload): the single-window configurations, run through
    # mpoa_align_reads (minipoa's whole main loop on one read set)
    ("config1", "config1", ["-r", "0"],
     "consensus: 20 reads of a 16,569 bp template at 1% sub, 0.5% ins, 0.5% del"),
    ("config2", "config2", ["-r", "1"],
     "row-column MSA: 100 reads of a 2,000 bp template at 10% error (4:3:3, geometric indels)"),
    ("config2_S", "config2", ["-S", "-r", "1"], "the same reads, -S seeding (README form)"),
    ("config4_S", "config4", ["-S", "-k", "15", "-w", "10", "-r", "0"],
     "50 reads of a 20,000 bp template at 12% error (2:4:4), -S k=15 w=10, graph past 100k "
     "nodes"),
]


def single_window(ctx, cores: int, with_cpu: bool, log) -> dict:
    """The single-window configurations: device wall time of mpoa_align_reads (best of two
    runs), GCUPS over band cells, algorithmic HBM rate over that wall time, and the upstream
    binary on the same FASTA on the host cores (-t 1 without -S, where minipoa has no other
    parallelism; -t cores with -S), outputs compared."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import golden_util as gu
    from poa_amd import api, synth
    out = {}
    for name, gen, args, what in SINGLE:
        reads = getattr(synth, gen)()
        tmp = tempfile.mkdtemp(prefix="mpoa_single_")
        try:
            path = os.path.join(tmp, "r.fa")
            reads.to_fasta(path)
            with open(path) as f:
                seq, off, names, comments = gu.parse_fasta(f.read())
            params = api.Params(**gu.parse_args(args))
            walls = []
            for _ in range(2):
                t0 = time.perf_counter()
                text, stats = ctx.align_reads(params, seq, off, names, comments, with_stats=True)
                walls.append(time.perf_counter() - t0)
            wall = min(walls)
            cells = sum(s.cells for s in stats)
            ecells = sum(s.edge_cells for s in stats)
            logical = sum(s.logical_cells for s in stats)
            bpc = 4 * (3 + 2 * ecells / max(cells, 1))
            ach = cells * bpc / wall / 1e9
            rec = {"workload": what, "flags": " ".join(args), "dp_calls": len(stats),
                   "band_cells": int(cells), "logical_cells": int(logical),
                   "gpu_wall_s": round(wall, 4), "gcups": round(cells / wall / 1e9, 4),
                   "roofline": {"bound": "hbm", "achieved": round(ach, 2),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(ach / HBM_PEAK_GBS, 6),
                                "bytes_per_cell": round(bpc, 3),
                                "note": "algorithmic bytes over the call's wall time"}}
            if with_cpu and os.path.exists(REF_BIN):
                threads = cores if "-S" in args else 1
                t0 = time.perf_counter()
                p = subprocess.run([REF_BIN, path, *args, "-t", str(threads)],
                                   capture_output=True, text=True, check=True)
                cw = time.perf_counter() - t0
                rec["cpu_baseline"] = {"value": round(cw, 4), "unit": "s", "cores": threads,
                                       "kind": "reference", "gcups": round(cells / cw / 1e9, 4),
                                       "outputs_identical_to_gpu": p.stdout == text}
                rec["gpu_over_cpu"] = round(cw / wall, 3)
            log(f"single {name}: GPU {wall:.3f} s ({rec['gcups']} GCUPS)" +
                (f", CPU {rec['cpu_baseline']['value']} s on {rec['cpu_baseline']['cores']} "
                 f"cores, identical {rec['cpu_baseline']['outputs_identical_to_gpu']}"
                 if "cpu_baseline" in rec else ""))
            out[name] = rec
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    return out


def seeded_batch(ctx, batch, with_cpu: bool, cpu_sample: int, cores: int, log) -> dict:
    """The same windows under -S through the batch API (k_seeded: one workgroup per window
    seeds every read itself and runs its sub-windows in turn), best of two executes; upstream
    minipoa -S on 1,000 evenly spaced windows, one process per window (-t 1), `cores` at a time
    (windows it fails on count, with their exit status); the oracle on the first cpu_sample
    windows on one host core; outputs and failures compared."""
    import collections
    from poa_amd import api
    params = api.Params(seeding=1, result=0)
    n = len(batch.win_off) - 1
    b = api.Batch(ctx, params, batch.win_off, batch.seq, batch.seq_off)
    best = None
    for _ in range(2):
        t0 = time.perf_counter()
        try:
            b.execute()
        except api.MpoaError:
            pass  # windows minipoa does not finish report their own status
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    ws, _ = b.window_stats()
    status = [int(x.status) for x in ws]
    texts = b.results().texts
    st = b.stats()
    b.free()
    rec = {"flags": "-S -r 0", "windows": n, "windows_per_s": round(n / best, 1),
           "kernel_ms": round(st.kernel_ms, 1), "launches": int(st.launches),
           "reruns": int(st.reruns),
           "statuses": {str(k): v for k, v in sorted(collections.Counter(status).items())}}
    if with_cpu and cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        m = min(cpu_sample, n)
        same = diverged = 0
        t0 = time.perf_counter()
        for w in range(m):
            win = batch.window(w)
            try:
                want, rc = oracle_lib.window(params, win.seq, win.off)[0], 0
            except RuntimeError as e:
                want, rc = None, int(str(e).split()[-1])
            if rc == status[w] and (rc != 0 or want == texts[w]):
                same += 1
            else:
                diverged += 1
        cw = time.perf_counter() - t0
        rec["oracle_sample"] = {"value": round(m / cw, 2), "unit": "windows/s", "cores": 1,
                                "kind": "port", "sample": f"first {m} windows, -S -r 0",
                                "windows_identical": same, "windows_differing": diverged}
    if with_cpu and os.path.exists(REF_BIN):
        idx = np.unique(np.linspace(0, n - 1, num=min(1000, n)).round().astype(int))
        tmp = tempfile.mkdtemp(prefix="mpoa_cpu_s_")
        try:
            paths = []
            for w in idx:
                pth = os.path.join(tmp, f"w{w}.fa")
                batch.window(int(w)).to_fasta(pth)
                paths.append(pth)

            def run(pth):
                return subprocess.run([REF_BIN, pth, "-S", "-t", "1"], capture_output=True)

            t0 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as ex:
                outs = list(ex.map(run, paths))
            wall = time.perf_counter() - t0
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        ok = [o.returncode == 0 for o in outs]
        same = sum(1 for i, w in enumerate(idx)
                   if ok[i] and status[int(w)] == 0 and outs[i].stdout.decode() == texts[int(w)])
        rec["cpu_baseline"] = {
            "value": round(len(idx) / wall, 2), "unit": "windows/s", "cores": cores,
            "kind": "reference",
            "sample": f"{len(idx)} of the {n} windows (evenly spaced), one minipoa -S -t 1 "
                      f"process per window, {cores} at a time; wall {wall:.2f} s",
            "upstream_finished": int(sum(ok)),
            "finished_identical_to_gpu": int(same)}
        rec["gpu_over_cpu"] = round(rec["windows_per_s"] / rec["cpu_baseline"]["value"], 2)
    log(f"seeded batch: {rec['windows_per_s']} windows/s, statuses {rec['statuses']}")
    return rec


def ingest(batch, threads: int, log) -> dict:
    """Read-set ingestion as a caller with files would do it: every window of the workload
    written as its own FASTA file (60-column lines), then all files parsed and packed by the
    library (mpoa_reads_load, minipoa's readFile rules) on `threads` host threads. The packed
    arrays must equal the workload's."""
    from poa_amd import api
    tmp = tempfile.mkdtemp(prefix="mpoa_ingest_")
    try:
        paths = []
        nbytes = 0
        for w in range(len(batch.win_off) - 1):
            r0, r1 = int(batch.win_off[w]), int(batch.win_off[w + 1])
            parts = []
            for r in range(r0, r1):
                s = batch.seq[batch.seq_off[r]:batch.seq_off[r + 1]].tobytes()
                parts.append(b">r%d window %d\n" % (r - r0, w))
                parts.extend(s[k:k + 60] + b"\n" for k in range(0, len(s), 60))
            data = b"".join(parts)
            path = os.path.join(tmp, f"w{w:06d}.fa")
            with open(path, "wb") as f:
                f.write(data)
            paths.append(path)
            nbytes += len(data)
        rs = api.ReadSet(paths, threads=threads)
        dt = rs.load_s
        same = (np.array_equal(rs.win_off, batch.win_off) and
                np.array_equal(rs.seq_off, batch.seq_off) and
                np.array_equal(rs.seq, batch.seq[:int(batch.seq_off[-1])]))
        log(f"ingest: {len(paths)} files, {nbytes / 1e6:.1f} MB parsed and packed in "
            f"{dt:.3f} s on {threads} threads, identical {same}")
        return {"files": len(paths), "bytes": nbytes, "threads": threads, "s": round(dt, 3),
                "windows_per_s": round(len(paths) / dt, 1), "MB_per_s": round(nbytes / dt / 1e6, 1),
                "arrays_identical": bool(same),
                "note": "file text to packed arrays (mpoa_reads_load); the upload is prepare_s"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def estimate_form(ctx, params, batch, steps: int, log) -> dict:
    """The same batch scheduled by the read-length estimate (no network): windows/s, device
    memory of the arenas and overflow re-runs, for the A/B against the learned schedule."""
    from poa_amd import api
    b = api.Batch(ctx, params, batch.win_off, batch.seq, batch.seq_off)
    b.execute()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    wall = time.perf_counter() - t0
    st = b.stats()
    texts = b.results().texts
    b.free()
    wps = batch.n_windows * steps / wall
    log(f"read-length estimate: {wps:.1f} windows/s, pools {st.pool_bytes / 1e9:.1f} GB, "
        f"{st.reruns} re-runs")
    return {"windows_per_s": round(wps, 1), "pool_bytes": int(st.pool_bytes),
            "reruns": int(st.reruns), "steps": steps, "texts": texts}


def wide_form(ctx, params, batch, pred, steps: int, log) -> dict:
    """The same batch and schedule with every DP call's matrices stored as int32
    (MPOA_FORM_WIDE, minipoa's width) instead of int16 where the values fit: windows/s, for
    the storage A/B."""
    from poa_amd import api
    b = api.Batch(ctx, params, batch.win_off, batch.seq, batch.seq_off)
    if pred is not None:
        b.set_predictions(pred)
    b.set_forms(api.FORM_WIDE)
    b.execute()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    wall = time.perf_counter() - t0
    st = b.stats()
    texts = b.results().texts
    b.free()
    wps = batch.n_windows * steps / wall
    log(f"int32 storage: {wps:.1f} windows/s, {st.narrow_cells} int16 cells")
    return {"windows_per_s": round(wps, 1), "reruns": int(st.reruns), "steps": steps,
            "texts": texts}


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload, from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy-kernel calibration);
    None when the summary does not cover the kernel."""
    path = os.path.join(ROOT, "profiles", "k_windows_uneven10k_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            if ln.startswith("traffic_bytes_per_launch") and kernel in ln:
                return int(ln.rsplit(None, 1)[-1])
    return None


INT32_VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # CUs x lanes x peak engine clock


def measured_int_ops(kernel: str):
    """INT32 VALU lane operations per launch of `kernel` on this workload, from the committed
    rocprofv3 --pmc summary (SQ_INSTS_VALU_INT32 x 64 lanes); None when it does not cover it."""
    path = os.path.join(ROOT, "profiles", "k_windows_uneven10k_sq_counters.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            if ln.startswith("valu_int32_lane_ops_per_launch") and kernel in ln:
                return int(ln.rsplit(None, 1)[-1])
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--windows", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None, help="threads per window workgroup")
    ap.add_argument("--cpl", type=int, default=None, help="DP columns per thread")
    ap.add_argument("--cpu-sample", type=int, default=1000)
    ap.add_argument("--cpu-cores", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule", default=None,
                    help="small_qmax,small_width,move_top for the launch split")
    ap.add_argument("--no-learned", action="store_true",
                    help="schedule by the read-length estimate instead of the trained network")
    ap.add_argument("--eager", action="store_true",
                    help="enqueue the launches one by one instead of replaying the hipGraph")
    ap.add_argument("--plain-sample", type=int, default=100,
                    help="windows run in the plain one-window-per-launch form (0: skip)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-window configurations")
    ap.add_argument("--no-ingest", action="store_true",
                    help="skip the timing of read-set ingestion from FASTA files")
    ap.add_argument("--no-seeded", action="store_true",
                    help="skip the -S run of the same windows through the batch API")
    ap.add_argument("--no-ab", action="store_true",
                    help="skip the learned-against-estimate scheduling A/B")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (torch), the known-size "
                         "kernel that scales rocprof FETCH_SIZE/WRITE_SIZE to bytes")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4,
                          "launches": 3}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(windows are independent: N GPUs would each run a 1/N share of them).",
              file=sys.stderr)
        return 2

    from poa_amd import api, sched, synth

    seed = synth.SEEDS[3] if args.seed is None else args.seed
    batch = synth.config3(seed=seed, n_windows=args.windows)
    ctx = api.Context(0)
    params = api.Params()
    tp = time.perf_counter()
    b = api.Batch(ctx, params, batch.win_off, batch.seq, batch.seq_off)
    learned = None
    pred = None
    if not args.no_learned and os.path.exists(sched.MODEL):
        # Learned scheduling: per-call features -> network on the GPU -> per-window cost,
        # arena sizes and tile choice. Host feature extraction and inference are part of the
        # preparation, outside the timed steps.
        tl = time.perf_counter()
        model = sched.Model.load()
        net = sched.Network(0)
        net.set_weights(model.weights)
        pred = sched.predict(model, net, params, batch.win_off, batch.seq, batch.seq_off)
        net.close()
        b.set_predictions(pred)
        learned = {"model": os.path.relpath(sched.MODEL, ROOT), "predict_s":
                   round(time.perf_counter() - tl, 3), "trained_on": model.meta}
    if args.schedule:
        b.set_schedule(*[int(x) for x in args.schedule.split(",")])
    b.set_graph(not args.eager)
    prepare_s = time.perf_counter() - tp
    if args.nt or args.cpl:
        b.set_tile(args.nt or 64, args.cpl or 4)
        b.tuning = {"small_tile": [args.nt or 64, args.cpl or 4], "from": "command line"}

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    log(f"{args.windows} windows, {int(batch.win_off[-1])} reads prepared")
    for i in range(args.warmup):
        log(f"warmup {i}: {b.execute():.1f} ms")
    kernel_ms = []
    launch_ms: dict[int, list[float]] = {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        b.execute()
        kernel_ms.append(b.stats().kernel_ms)
        for ls in b.launch_stats():
            launch_ms.setdefault(ls.launch, []).append(ls.ms)
        log(f"step {i}: kernel {kernel_ms[-1]:.1f} ms")
    wall = time.perf_counter() - t0
    tr = time.perf_counter()
    res = b.results()
    results_s = time.perf_counter() - tr
    st = res.stats
    ws, tick_s = b.window_stats()
    ph = np.array([[w.ticks[k] for k in range(8)] for w in ws], dtype=np.float64) * tick_s
    slow = int(np.argmax(ph[:, :6].sum(axis=1)))
    log("phase seconds summed over windows: " +
        ", ".join(f"{n}={v:.3f}" for n, v in zip(api.PHASES, ph.sum(axis=0))))
    log(f"slowest window {slow}: nodes={ws[slow].nodes} cells={ws[slow].cells} " +
        ", ".join(f"{n}={v * 1e3:.1f}ms" for n, v in zip(api.PHASES, ph[slow])))
    tb = np.array([w.t_begin for w in ws], dtype=np.float64)
    te = np.array([w.t_end for w in ws], dtype=np.float64)
    lb = np.array([w.launch for w in ws])
    t0d = tb.min()
    for L in sorted(set(lb.tolist())):
        sel = lb == L
        log(f"launch {L}: {int(sel.sum())} windows, first start {(tb[sel].min() - t0d) * tick_s * 1e3:.1f} ms, "
            f"last start {(tb[sel].max() - t0d) * tick_s * 1e3:.1f} ms, "
            f"last end {(te[sel].max() - t0d) * tick_s * 1e3:.1f} ms, "
            f"busy {((te[sel] - tb[sel]).sum()) * tick_s:.2f} window-s")

    ms_per_step = wall * 1000.0 / args.steps
    windows_per_s = args.windows * args.steps / wall
    gcups = st.cells * args.steps / wall / 1e9
    ebar = st.edge_cells / max(st.cells, 1)
    # Dominant kernel: the launch that processes the most DP cells. Algorithmic bytes per cell:
    # int32 H, D and I written once, H and D of the predecessor row read once per edge-cell.
    lst = {ls.launch: ls for ls in b.launch_stats()}
    dom = max(lst.values(), key=lambda ls: ls.cells)
    dom_ebar = dom.edge_cells / max(dom.cells, 1)
    # Score bytes per cell of the launch: 2 for the calls stored as int16, 4 for the rest.
    dom_narrow = sum(w.narrow_cells for w in ws if w.launch == dom.launch)
    score_bytes = 4 - 2 * dom_narrow / max(dom.cells, 1)
    bytes_per_cell = score_bytes * (3 + 2 * dom_ebar)
    alg_bytes = dom.cells * bytes_per_cell
    dom_ms = float(np.mean(launch_ms[dom.launch]))
    achieved = alg_bytes / (dom_ms / 1000.0) / 1e9
    kname = f"mpoa::k_windows<{dom.threads}, {dom.cols}>"
    traffic = measured_traffic(kname)
    for L, ls in sorted(lst.items()):  # 0 small, 1 medium, 2 re-run, 3 wide tile
        log(f"launch {L} ({ls.threads}x{ls.cols}, {ls.slots} slots): {ls.windows} windows, "
            f"{ls.cells / 1e9:.2f} G cells, {np.mean(launch_ms[L]):.1f} ms (HIP events)")
    line = {
        "metric": "POA windows/sec on 1 MI355X (bit-exact to minipoa), with GCUPS",
        "value": round(windows_per_s, 1),
        "unit": "windows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "gcups": round(gcups, 2),
        "config": {
            "workload": "batched uneven windows: 10,000 POA jobs, 5-60 reads each, template "
                        "lognormal(median 500 bp, sigma 0.5) clipped to 150-2,000 bp, error "
                        "U(5%,15%) split 1:1:1; minipoa defaults, consensus",
            "windows": args.windows,
            "seed": seed,
            "reads": int(batch.win_off[-1]),
            "bases": int(batch.seq_off[-1]),
            "dp_calls": int(st.dp_calls),
            "band_cells": int(st.cells),
            "edge_cells_per_cell": round(ebar, 4),
            "logical_cells": int(st.logical_cells),
            "score_storage": {"int16_cells": int(st.narrow_cells),
                              "int32_cells": int(st.cells - st.narrow_cells),
                              "note": "int32 arithmetic; H, D and I stored as int16 for the "
                                      "DP calls whose values fit, else int32"},
            "reruns": int(st.reruns),
            "launches": int(st.launches),
            "slots": int(st.slots),
            "pool_bytes": int(st.pool_bytes),
            "pcie_inclusive_windows_per_s": round(
                args.windows / (prepare_s + ms_per_step / 1000.0 + results_s), 1),
            "prepare_s": round(prepare_s, 3),
            "results_s": round(results_s, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": kname,
            "kernel_ms": round(dom_ms, 3),
            "windows": int(dom.windows),
            "band_cells": int(dom.cells),
            "bytes_per_cell": round(bytes_per_cell, 3),
            "score_bytes": round(score_bytes, 4),
            "alg_bytes": int(alg_bytes),
            "all_launches_ms": round(float(np.mean(kernel_ms)), 3),
        },
    }
    int_ops = measured_int_ops(kname)
    if int_ops is not None:
        tops = int_ops / (dom_ms / 1e3) / 1e12
        line["roofline"]["int32_valu"] = {
            "achieved": round(tops, 3), "peak": INT32_VALU_PEAK_TOPS, "unit": "Tops/s",
            "frac": round(tops / INT32_VALU_PEAK_TOPS, 4), "lane_ops_per_launch": int_ops}
    line["config"]["scheduling"] = learned or "read-length estimate"
    line["config"]["tuning"] = b.tuning or "library defaults (no autotuner output)"
    line["config"]["launch_form"] = ("eager launches on three streams" if args.eager else
                                     f"hipGraph replay ({st.graph_builds} build, "
                                     f"{st.graph_launches} launch per step)")
    if args.plain_sample:
        line["plain_one_window_per_launch"] = plain_form(ctx, params, batch, args.plain_sample,
                                                         res.texts, windows_per_s, log)
    cores = args.cpu_cores or min(16, os.cpu_count() or 1)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(batch, res.texts, args.cpu_sample, cores)
    pool_learned = int(st.pool_bytes)
    b.free()
    if learned and not args.no_ab:
        est = estimate_form(ctx, params, batch, 2, log)
        line["scheduling_ab"] = {
            "learned": {"windows_per_s": round(windows_per_s, 1), "pool_bytes": pool_learned,
                        "reruns": int(st.reruns)},
            "read_length_estimate": {k: v for k, v in est.items() if k != "texts"},
            "outputs_identical": est["texts"] == res.texts}
    if not args.no_ab:
        wide = wide_form(ctx, params, batch, pred, 2, log)
        line["storage_ab"] = {
            "int16_where_fits": {"windows_per_s": round(windows_per_s, 1),
                                 "int16_cells": int(st.narrow_cells)},
            "int32_only": {k: v for k, v in wide.items() if k != "texts"},
            "outputs_identical": wide["texts"] == res.texts}
    if not args.no_ingest:
        line["ingest"] = ingest(batch, cores, log)
    if not args.no_single:
        line["single_window"] = single_window(ctx, cores, not args.no_cpu, log)
    if not args.no_seeded:
        line["seeded_batch"] = seeded_batch(ctx, batch, not args.no_cpu, 200, cores, log)
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
