"""The CPU oracle against the golden fixtures: for every case, minipoa's exact stdout
(consensus, row-column MSA or GFA) and the score of every poa() call (the "score:" lines of
-V 1, src/align.cpp:1577), as printed by the upstream binary when the fixtures were made."""
from __future__ import annotations

import pytest

import golden_util as gu
import oracle_lib

CASES = gu.load_cases()


@pytest.mark.parametrize("case", CASES, ids=[gu.case_id(c) for c in CASES])
def test_oracle_matches_upstream(case):
    assert case["returncode"] == 0
    seq, off, names, comments = gu.parse_fasta(case["fasta"])
    p = gu.parse_args(case["args"])
    text, stats = oracle_lib.window(p, seq, off, names, comments)
    assert [s.score for s in stats] == case["scores"]
    assert text == case["stdout"]


def test_fixture_coverage():
    """The fixtures cover every output format, the adaptive band, seeding, the full band and
    non-default scores."""
    args = [" ".join(c["args"]) for c in CASES]
    assert any("-r 0" in a for a in args) and any("-r 1" in a for a in args)
    assert any("-r 2" in a for a in args)
    assert any("-B" in a for a in args) and any("-S" in a for a in args)
    assert any("-f 0" in a for a in args) and any("-M 3" in a for a in args)
