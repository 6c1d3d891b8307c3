"""Incremental mode (minipoa -i graph.gfa) through the call-level C ABI on the device:
mpoa_graph_read_gfa builds the same graph as the oracle's read_gfa (nodes, edges, weights,
ranks), then minipoa's loop over the second read set gives every read's res exactly as the
oracle's and the text upstream printed (tests/golden/gfa_input.json.gz); where upstream dies on
an empty -S sub-window the device reports MPOA_E_EMPTY_WINDOW at the same read. Inputs the
library refuses fail with MPOA_E_ARG."""
from __future__ import annotations

import gzip
import json
import os

import numpy as np
import pytest

import golden_util as gu
import oracle_lib as ol
from poa_amd import api
from test_gpu_calls import _same_graph, res_rows

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(gzip.open(os.path.join(HERE, "golden", "gfa_input.json.gz"), "rt"))


@pytest.mark.gpu
@pytest.mark.parametrize("case", FIX["cases"], ids=lambda c: f"{c['case']}[{' '.join(c['args'])}]")
def test_gfa_input_loop(ctx, case):
    print(f"gfa input {case['case']} {' '.join(case['args'])}", flush=True)
    inp = FIX["inputs"][case["case"]]
    seq, off, names, comments, n_paths = gu.gfa_read_set(inp["gfa"], inp["fasta"])
    p = gu.parse_args(case["args"])
    p.pop("progressive", None)  # -p is ignored with -i (src/main.cpp:101)
    params = api.Params(**p)
    dag, g = ol.Dag(p, seq, off), api.Graph(ctx, params)
    seeder = api.Seeder(params, seq, off)
    gfa = inp["gfa"].encode()
    assert g.read_gfa(gfa) == dag.read_gfa(gfa) == names[:n_paths]
    _same_graph(dag.export(), g.export())
    n = len(off) - 1
    failed = None
    for i in range(n_paths, n):
        q = bytes(seq[off[i]:off[i + 1]])
        try:
            want = dag.alignment(i, q)
        except RuntimeError as e:
            assert "oracle error -5" in str(e)
            failed = i
            break
        got = g.alignment(i, q, seeder)
        np.testing.assert_array_equal(res_rows(got), want, err_msg=f"read {i}")
        dag.add_path(i, want)
        g.add_path(i, got)
        op = 1 if i + 1 == n else 0
        dag.topsort(op)
        g.topsort(op)
    if case["returncode"] != 0:
        assert failed is not None and case["died"]
        with pytest.raises(api.MpoaError) as e:
            g.alignment(failed, bytes(seq[off[failed]:off[failed + 1]]), seeder)
        assert e.value.rc == -5
    else:
        assert failed is None
        assert g.output(params.result, names, comments) == case["stdout"]
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("result,gfa", [
    (0, b"S\ts1\tACGT\nS\ts2\tGG\nP\tp\ts1+,s2-\t*\n"),
    (0, b"S\ts1\tACGT\nP\tp\ts1+,s9+\t*\n"),
    (0, b"S\ts1\tACGT\nS\ts1\tGG\n"),
    (0, b"S\ts1\n"),
    (0, b"S\ts1\tACGT\nP\tp\t\t*\n"),
    (2, b"S\ts1\tACGT\nP\tp\ts1+\t*\n"),
], ids=["reverse", "unknown", "duplicated", "no-bases", "empty-path", "gfa-output"])
def test_gfa_input_refused(ctx, result, gfa):
    g = api.Graph(ctx, api.Params(result=result))
    with pytest.raises(api.MpoaError) as e:
        g.read_gfa(gfa)
    assert e.value.rc == -1
    g.close()


@pytest.mark.gpu
def test_gfa_input_after_reads_refused(ctx):
    """read_gfa fills a new graph only."""
    g = api.Graph(ctx, api.Params())
    g.add_path(0, np.array([[0, 4, -1], [-1, 0, -1]], np.int32))
    with pytest.raises(api.MpoaError) as e:
        g.read_gfa(b"S\ts1\tACGT\nP\tp\ts1+\t*\n")
    assert e.value.rc == -1
    g.close()


ROOT = os.path.dirname(HERE)
PROGRAMS = {"cli": os.path.join(ROOT, "poa_amd", "minipoa_amd"),
            "loop": os.path.join(ROOT, "poa_amd", "minipoa_loop")}


@pytest.mark.gpu
@pytest.mark.parametrize("prog", sorted(PROGRAMS))
@pytest.mark.parametrize("case", [c for c in FIX["cases"] if c["case"] in ("hand", "synth500")],
                         ids=lambda c: f"{c['case']}[{' '.join(c['args'])}]")
def test_gfa_input_programs(tmp_path, prog, case):
    """minipoa_amd -i (the command line) and minipoa_loop -i (the loop over the C++ facade)
    print what upstream printed, with its exit status."""
    import subprocess
    exe = PROGRAMS[prog]
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    inp = FIX["inputs"][case["case"]]
    gfa, fa = tmp_path / "in.gfa", tmp_path / "reads.fa"
    gfa.write_text(inp["gfa"])
    fa.write_text(inp["fasta"])
    print(f"{prog} -i {case['case']} {' '.join(case['args'])}", flush=True)
    p = subprocess.run([exe, *case["args"], "-i", str(gfa), str(fa)], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == case["returncode"], p.stderr
    assert p.stdout == case["stdout"]
