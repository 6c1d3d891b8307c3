"""The device path on the full-size single-window workloads against upstream minipoa's outputs
(tests/golden/full_configs.json.gz), through the C ABI."""
from __future__ import annotations

import pytest

import full_util as fu
import golden_util as gu
from poa_amd import api

CASES = fu.load_cases()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["case"] for c in CASES])
def test_device_full_config(ctx, case):
    reads = fu.reads_of(case)
    sha, text = fu.fasta_of(reads)
    assert sha == case["fasta_sha256"], "the generator's reads differ from the fixture's"
    seq, off, names, comments = gu.parse_fasta(text)
    print(f"full {case['case']} {' '.join(case['args'])}", flush=True)
    out, stats = ctx.align_reads(api.Params(**gu.parse_args(case["args"])), seq, off, names,
                                 comments, with_stats=True)
    assert [s.score for s in stats] == case["scores"]
    assert out == case["stdout"]
