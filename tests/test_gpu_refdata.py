"""The device path on the reference's own data and failure paths, through the C ABI: the whole
data/mtDNA.fasta (393 reads) under -r 0/1/2 and -B, the README's MSA command (-S -t 12 -r 1:
upstream crashes, the library returns MPOA_E_EMPTY_WINDOW before any kernel sees the window),
upstream's backtrack errors (MPOA_E_BACKTRACK; the command-line tool exits 1 as upstream
does), and the MSA with a consensus row (result 3) in the single-window, seeded, batched and
call-level forms."""
from __future__ import annotations

import gzip
import json
import os
import subprocess

import numpy as np
import pytest

import full_util as fu
import golden_util as gu
import oracle_lib
from poa_amd import api

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(ROOT, "poa_amd", "minipoa_amd")
MTDNA = json.load(gzip.open(os.path.join(HERE, "golden", "mtdna_full.json.gz"), "rt"))
FAILURES = json.load(open(os.path.join(HERE, "golden", "failures.json")))["cases"]

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", MTDNA["cases"], ids=lambda c: " ".join(c["args"]))
def test_device_mtdna_full(ctx, case):
    seq, off, names, comments = gu.parse_fasta(MTDNA["fasta"])
    params = api.Params(**gu.parse_args(case["args"]))
    print(f"mtdna {' '.join(case['args'])}", flush=True)
    if case["returncode"] != 0:
        with pytest.raises(api.MpoaError) as e:
            ctx.align_reads(params, seq, off, names, comments)
        assert e.value.rc == -5
        return
    out, stats = ctx.align_reads(params, seq, off, names, comments, with_stats=True)
    assert [s.score for s in stats] == case["scores"]
    assert out == case["stdout"]


@pytest.mark.parametrize("case", FAILURES, ids=lambda c: c["case"])
def test_device_backtrack_error(ctx, case):
    seq, off, names, comments = gu.parse_fasta(case["fasta"])
    print(f"failure {case['case']}", flush=True)
    with pytest.raises(api.MpoaError) as e:
        ctx.align_reads(api.Params(**gu.parse_args(case["args"])), seq, off, names, comments)
    assert e.value.rc == -3


@pytest.mark.skipif(not os.path.exists(CLI), reason="minipoa_amd not built")
def test_cli_exit_statuses(tmp_path):
    for case in FAILURES:
        path = tmp_path / f"{case['case']}.fa"
        path.write_text(case["fasta"])
        p = subprocess.run([CLI, str(path), *case["args"]], capture_output=True, text=True,
                           timeout=120)
        assert p.returncode == case["returncode"] == 1
        assert "backtrack error" in p.stderr and p.stdout == case["stdout"]
    path = tmp_path / "mtDNA.fasta"
    path.write_text(MTDNA["fasta"])
    p = subprocess.run([CLI, str(path), "-S", "-t", "12", "-r", "1"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 1 and p.stdout == "" and "sub-window" in p.stderr


def _pairs():
    full = {c["case"]: c for c in fu.load_cases()}
    out = [("config2", full["config2"], full["config2_r0"]),
           ("config2_S", full["config2_S"], full["config2_S_r0"])]
    small = gu.load_cases()
    for name in ("mtdna40", "cfg2_small", "edge_one_read", "edge_identical", "edge_with_n"):
        r1 = [c for c in small if c["case"] == name and c["args"] == ["-r", "1"]]
        r0 = [c for c in small if c["case"] == name and c["args"] == ["-r", "0"]]
        if r1 and r0:
            out.append((name, r1[0], r0[0]))
    mt = {tuple(c["args"]): c for c in MTDNA["cases"]}
    out.append(("mtdna_full", mt[("-r", "1")], mt[("-r", "0")]))
    return out


def _text(name, case):
    if "fasta" in case:
        return case["fasta"]
    if name == "mtdna_full":
        return MTDNA["fasta"]
    return fu.fasta_of(fu.reads_of(case))[1]


@pytest.mark.parametrize("name,r1,r0", _pairs(), ids=lambda x: x if isinstance(x, str) else "")
def test_device_msa_consensus_row(ctx, name, r1, r0):
    seq, off, names, comments = gu.parse_fasta(_text(name, r1))
    p = gu.parse_args(r1["args"])
    p["result"] = 3
    print(f"msa+cons {name}", flush=True)
    out = ctx.align_reads(api.Params(**p), seq, off, names, comments)
    gu.check_msa_consensus_row(out, r1["stdout"], r0["stdout"])
    want, _ = oracle_lib.window(p, seq, off, names, comments)
    assert out == want


def test_device_msa_consensus_row_batched_and_calls(ctx):
    from poa_amd import synth
    b = synth.config3(seed=0x3c3c, n_windows=24)
    params = api.Params(result=3)
    bat = api.Batch(ctx, params, b.win_off, b.seq, b.seq_off)
    bat.execute()
    got = bat.results().texts
    want = oracle_lib.batch(params, b.win_off, b.seq, b.seq_off)
    assert got == want
    # the call-level graph's result-3 output equals the single-window text
    r = synth.window(0x3c3d, 700, 15, 0.06, 0.04, 0.04, 0.7)
    g = api.Graph(ctx, params)
    n = len(r.off) - 1
    for i in range(n):
        q = bytes(r.seq[r.off[i]:r.off[i + 1]])
        g.add_path(i, g.alignment(i, q))
        g.topsort(1 if i + 1 == n else 0)
    assert g.output(3) == ctx.align_reads(params, r.seq, r.off)
    g.close()


LOGICAL_CASES = [c for c in gu.load_cases()
                 if (c["case"], tuple(c["args"])) in {
                     ("mtdna40", ("-r", "0")), ("mtdna40", ("-B", "-r", "0")),
                     ("mtdna40", ("-f", "0", "-r", "0")), ("cfg2_seeded", ("-S", "-r", "1")),
                     ("cfg4_small", ("-S", "-k", "15", "-w", "10", "-r", "0"))}]


@pytest.mark.parametrize("case", LOGICAL_CASES, ids=gu.case_id)
def test_device_logical_cells(ctx, case):
    """Every DP call's logical cells (rows 1..n-1 of Me - Ms + 1) equal the oracle's, in the
    static, full-width, adaptive and seeded bands; the batch total is their sum."""
    seq, off, names, comments = gu.parse_fasta(case["fasta"])
    p = gu.parse_args(case["args"])
    _, stats = ctx.align_reads(api.Params(**p), seq, off, names, comments, with_stats=True)
    _, want = oracle_lib.window(p, seq, off, names, comments)
    assert [s.logical_cells for s in stats] == [w.logical_cells for w in want]
    assert [s.cells for s in stats] == [w.cells for w in want]
    if not p.get("seeding"):
        b = api.Batch(ctx, api.Params(**p), np.array([0, len(off) - 1], np.int32), seq, off)
        b.execute()
        assert b.stats().logical_cells == sum(w.logical_cells for w in want)
        b.free()


def test_device_progressive_batch(ctx):
    """-p in the batch: every window's reads in its own guide-tree order, MSA rows back in
    input order, equal to the oracle window by window; -p with GFA output is refused."""
    from poa_amd import synth
    b = synth.config3(seed=0x9f, n_windows=20)
    for result in (1, 3, 0):
        params = api.Params(result=result, progressive=1)
        bat = api.Batch(ctx, params, b.win_off, b.seq, b.seq_off)
        bat.execute()
        assert bat.results().texts == oracle_lib.batch(params, b.win_off, b.seq, b.seq_off)
        bat.free()
    with pytest.raises(api.MpoaError) as e:
        api.Batch(ctx, api.Params(result=2, progressive=1), b.win_off, b.seq, b.seq_off)
    assert e.value.rc == -1
