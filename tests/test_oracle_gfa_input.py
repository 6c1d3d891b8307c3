"""The oracle's incremental mode (minipoa -i graph.gfa: src/main.cpp:86-126,
src/handle_input.cpp:444-495) against upstream runs (tests/golden/gfa_input.json.gz): GFAs that
upstream wrote with -r 2 and a hand-written one, a second read set aligned to each under -r 0,
-r 1, -B, -S and -p (which -i turns off, main.cpp:101); outputs and every DP score. Where
upstream dies (a -S sub-window that ends before it begins, std::length_error) the oracle
reports MPOA_E_EMPTY_WINDOW after the same DP calls. Also the inputs the library refuses."""
from __future__ import annotations

import gzip
import json
import os

import pytest

import golden_util as gu
import oracle_lib

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(gzip.open(os.path.join(HERE, "golden", "gfa_input.json.gz"), "rt"))


def run_case(dag_of, case, **kw):
    """(text or None, the error, dag): minipoa -i's loop for one fixture run through dag_of(p,
    seq, off), a call-level object (the oracle's Dag or api.Graph with its read set)."""
    inp = FIX["inputs"][case["case"]]
    seq, off, names, comments, _ = gu.gfa_read_set(inp["gfa"], inp["fasta"])
    p = gu.parse_args(case["args"])
    p.pop("progressive", None)  # -p is ignored with -i (main.cpp:101)
    dag = dag_of(p, seq, off)
    try:
        text, _ = oracle_lib.run_loop(dag, len(off) - 1, seq, off, p.get("result", 0), names,
                                      comments, gfa=inp["gfa"].encode(), **kw)
        return text, None, dag
    except RuntimeError as e:
        return None, e, dag


@pytest.mark.parametrize("case", FIX["cases"], ids=lambda c: f"{c['case']}[{' '.join(c['args'])}]")
def test_oracle_gfa_input(case):
    text, err, dag = run_case(lambda p, s, o: oracle_lib.Dag(p, s, o), case)
    if case["returncode"] != 0:
        assert case["died"] and "-S" in case["args"]
        assert err is not None and "oracle error -5" in str(err)
    else:
        assert err is None, err
        assert text == case["stdout"]
    assert dag.scores() == case["scores"]


def test_oracle_gfa_paths_are_reads():
    """The hand-written GFA's paths come back as reads 0..3 with their names (a blank kept),
    each row of the MSA without reads aligned is its path's bases."""
    inp = FIX["inputs"]["hand"]
    dag = oracle_lib.Dag(gu.parse_args(["-r", "1"]))
    names = dag.read_gfa(inp["gfa"].encode())
    assert names == ["p1", "p2", "p3 with blank", "p4"]
    dag.topsort(1)
    text = dag.output(1, names, [""] * 4)
    rows = text.splitlines()[1::2]
    segs = {"s1": "ACGTTGCA", "s2": "GGAT", "s3": "GCAN", "s4": "TTACGGATCCA"}
    want = ["s1 s2 s4", "s1 s3 s4", "s2 s4", "s1 s2 s4"]
    for row, w in zip(rows, want):
        assert row.replace("-", "") == "".join(segs[s] for s in w.split())


@pytest.mark.parametrize("gfa,why", [
    (b"S\ts1\tACGT\nS\ts2\tGG\nP\tp\ts1+,s2-\t*\n", "reverse segment"),
    (b"S\ts1\tACGT\nP\tp\ts1+,s9+\t*\n", "unknown segment"),
    (b"S\ts1\tACGT\nS\ts1\tGG\n", "duplicated segment"),
    (b"S\ts1\n", "segment without bases"),
    (b"S\ts1\tACGT\nP\tp\t\t*\n", "path without segments"),
])
def test_oracle_gfa_refused(gfa, why):
    with pytest.raises(RuntimeError, match="oracle error -1"):
        oracle_lib.Dag(gu.parse_args(["-r", "0"])).read_gfa(gfa)


def test_oracle_gfa_refused_with_gfa_output():
    with pytest.raises(RuntimeError, match="oracle error -1"):
        oracle_lib.Dag(gu.parse_args(["-r", "2"])).read_gfa(b"S\ts1\tACGT\nP\tp\ts1+\t*\n")
