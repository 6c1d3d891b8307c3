"""The device path against the golden fixtures (upstream minipoa's own outputs), through the
C ABI: stdout text and the score of every DP call."""
from __future__ import annotations

import pytest

import golden_util as gu

CASES = gu.load_cases()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[gu.case_id(c) for c in CASES])
def test_device_matches_upstream(ctx, case):
    from poa_amd import api
    seq, off, names, comments = gu.parse_fasta(case["fasta"])
    p = api.Params(**gu.parse_args(case["args"]))
    print(f"golden {gu.case_id(case)}", flush=True)
    text, stats = ctx.align_reads(p, seq, off, names, comments, with_stats=True)
    assert [s.score for s in stats] == case["scores"]
    assert text == case["stdout"]


@pytest.mark.gpu
def test_golden_batch(ctx):
    """All consensus fixtures without seeding as one batch of independent windows."""
    import numpy as np
    from poa_amd import api
    sel = [c for c in CASES if c["args"] == ["-r", "0"]]
    seqs, offs, win = [], [0], [0]
    base = 0
    for c in sel:
        seq, off, _, _ = gu.parse_fasta(c["fasta"])
        seqs.append(seq)
        offs.extend((off[1:] + base).tolist())
        base += int(off[-1])
        win.append(win[-1] + len(off) - 1)
    b = api.Batch(ctx, api.Params(), np.array(win, np.int32), np.concatenate(seqs),
                  np.array(offs, np.int64))
    b.execute()
    texts = b.results().texts
    b.free()
    assert texts == [c["stdout"] for c in sel]
