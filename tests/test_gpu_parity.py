"""GPU path vs the CPU oracle: byte-identical minipoa output and identical per-call scores.

Every case prints what it runs before touching the GPU, so a failing run names its case.
"""
import numpy as np
import pytest

import oracle_lib
from poa_amd import api, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _check_window(ctx, reads: synth.Reads, **flags):
    p = api.Params(**flags)
    want, calls = oracle_lib.window(p, reads.seq, reads.off)
    got, gstats = ctx.align_reads(p, reads.seq, reads.off, with_stats=True)
    assert got == want
    assert [s.score for s in gstats] == [c.score for c in calls]


@pytest.mark.parametrize("result", [0, 1])
def test_small_window(ctx, result):
    reads = synth.window(11, 300, 8, 0.04, 0.03, 0.03, 0.7)
    _say("small window", result)
    _check_window(ctx, reads, result=result)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_config3_like_windows(ctx, seed):
    b = synth.config3(seed=seed, n_windows=4)
    for w in range(b.n_windows):
        _say("config3-like", seed, w)
        _check_window(ctx, b.window(w))


def test_adaptive_band(ctx):
    reads = synth.window(5, 600, 10, 0.05, 0.03, 0.03, 0.7)
    _say("adaptive band -B")
    _check_window(ctx, reads, ab_band=1)
    _check_window(ctx, reads, ab_band=1, result=1)


def test_full_band(ctx):
    reads = synth.window(6, 200, 6, 0.05, 0.05, 0.05)
    _say("full band -f 0")
    _check_window(ctx, reads, band_f=0)


def test_scores_nondefault(ctx):
    reads = synth.window(7, 250, 6, 0.05, 0.05, 0.05)
    _say("scores -M 3 -X 5 -O 6 -E 1 -b 20 -f 10")
    _check_window(ctx, reads, match=3, mismatch=5, gap_open=6, gap_ext=1, band_b=20, band_f=10)


def test_edge_cases(ctx):
    cases = {
        "one read": [b"ACGTACGTAC"],
        "two reads": [b"ACGTACGTAC", b"ACGTTACGTAC"],
        "identical": [b"GATTACAGATTACA"] * 4,
        "with N": [b"ACGTNNACGT", b"ACGTACGT", b"ACNTACGTA"],
        "short": [b"AC", b"A", b"ACG", b"C"],
        "lowercase": [b"acgtacgt", b"ACGTACGT", b"acgTTacgt"],
    }
    for name, rs in cases.items():
        _say("edge", name)
        seq = np.frombuffer(b"".join(rs), np.uint8).copy()
        off = np.zeros(len(rs) + 1, np.int64)
        off[1:] = np.cumsum([len(r) for r in rs])
        reads = synth.Reads(seq, off)
        for result in (0, 1):
            _check_window(ctx, reads, result=result)


def test_batch_matches_oracle(ctx):
    b = synth.config3(seed=21, n_windows=64)
    p = api.Params()
    _say("batch 64 windows")
    bat = api.Batch(ctx, p, b.win_off, b.seq, b.seq_off)
    bat.execute()
    got = bat.results().texts
    want = oracle_lib.batch(p, b.win_off, b.seq, b.seq_off)
    assert got == want


@pytest.mark.parametrize("result", [1, 2])
def test_batch_msa_gfa_matches_oracle(ctx, result):
    """Row-column MSA and GFA through the batched launches (the node paths of every read)."""
    b = synth.config3(seed=23, n_windows=48)
    p = api.Params(result=result)
    _say("batch 48 windows, result", result)
    bat = api.Batch(ctx, p, b.win_off, b.seq, b.seq_off)
    bat.execute()
    got = bat.results().texts
    bat.free()
    assert got == oracle_lib.batch(p, b.win_off, b.seq, b.seq_off)


@pytest.mark.parametrize("result", [0, 1])
def test_seeded_windows(ctx, result):
    """-S: host seeding, one workgroup per sub-window, adaptive band (align.cpp:1763-1827)."""
    reads = synth.window(31, 1500, 14, 0.04, 0.03, 0.03, 0.7)
    _say("seeded config2-like", result)
    _check_window(ctx, reads, seeding=1, result=result)
    reads = synth.window(32, 5000, 6, 0.024, 0.048, 0.048, 0.7)
    _say("seeded config4-like", result)
    _check_window(ctx, reads, seeding=1, k=15, mm_w=10, result=result)
