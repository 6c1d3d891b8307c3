"""The oracle on the reference's own data and failure paths: the whole data/mtDNA.fasta (393
reads) under the README's commands and -r 0/1/2 and -B (tests/golden/mtdna_full.json.gz), the
README's MSA command (-S -t 12 -r 1) on which upstream crashes, upstream's backtrack errors
(tests/golden/failures.json), and the MSA with a consensus row (result 3) against upstream's
-r 1 and -r 0 texts of the same runs."""
from __future__ import annotations

import gzip
import json
import os

import pytest

import full_util as fu
import golden_util as gu
import oracle_lib

HERE = os.path.dirname(os.path.abspath(__file__))
MTDNA = json.load(gzip.open(os.path.join(HERE, "golden", "mtdna_full.json.gz"), "rt"))
FAILURES = json.load(open(os.path.join(HERE, "golden", "failures.json")))["cases"]


def mtdna_reads():
    import hashlib
    assert hashlib.sha256(MTDNA["fasta"].encode()).hexdigest() == MTDNA["fasta_sha256"]
    return gu.parse_fasta(MTDNA["fasta"])


@pytest.mark.parametrize("case", MTDNA["cases"], ids=lambda c: " ".join(c["args"]))
def test_oracle_mtdna_full(case):
    seq, off, names, comments = mtdna_reads()
    assert len(off) - 1 == 393
    p = gu.parse_args(case["args"])
    if case["returncode"] != 0:
        # upstream dies (SIGSEGV) on a -S sub-window that ends before it begins; the oracle
        # reports it
        with pytest.raises(RuntimeError, match="oracle error -5"):
            oracle_lib.window(p, seq, off, names, comments)
        return
    out, stats = oracle_lib.window(p, seq, off, names, comments)
    assert [s.score for s in stats] == case["scores"]
    assert out == case["stdout"]


@pytest.mark.parametrize("case", FAILURES, ids=lambda c: c["case"])
def test_oracle_backtrack_error(case):
    assert case["returncode"] == 1 and "backtrack error" in case["stderr"]
    seq, off, names, comments = gu.parse_fasta(case["fasta"])
    with pytest.raises(RuntimeError, match="oracle error -3"):
        oracle_lib.window(gu.parse_args(case["args"]), seq, off, names, comments)


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


@pytest.mark.parametrize("name,r1,r0", _pairs(), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_msa_consensus_row(name, r1, r0):
    if "fasta" in r1:
        text = r1["fasta"]
    elif name == "mtdna_full":
        text = MTDNA["fasta"]
    else:
        text = fu.fasta_of(fu.reads_of(r1))[1]
    seq, off, names, comments = gu.parse_fasta(text)
    p = gu.parse_args(r1["args"])
    p["result"] = 3
    out, _ = oracle_lib.window(p, seq, off, names, comments)
    gu.check_msa_consensus_row(out, r1["stdout"], r0["stdout"])
