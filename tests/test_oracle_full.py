"""The oracle on the full-size single-window workloads against upstream minipoa's outputs
(tests/golden/full_configs.json.gz): the 16,569 bp x 20 read consensus run, the 2,000 bp x 100
read MSA run with and without -S, and the 20 kb x 50 read seeded run (a graph of ~103,000
nodes). Also pins the generator: the FASTA it writes hashes to the fixture's SHA-256."""
from __future__ import annotations

import pytest

import full_util as fu
import golden_util as gu
import oracle_lib

CASES = fu.load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["case"] for c in CASES])
def test_oracle_full_config(case):
    reads = fu.reads_of(case)
    sha, text = fu.fasta_of(reads)
    assert sha == case["fasta_sha256"]
    seq, off, names, comments = gu.parse_fasta(text)
    out, stats = oracle_lib.window(gu.parse_args(case["args"]), seq, off, names, comments)
    assert [s.score for s in stats] == case["scores"]
    assert out == case["stdout"]
