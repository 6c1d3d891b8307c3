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


def test_held_text_follow_up_calls(ctx):
    """A call whose text does not fit keeps it for the follow-up call with the same arguments;
    a call with other arguments in between runs its own alignment."""
    import ctypes
    ra = synth.window(41, 400, 7, 0.04, 0.03, 0.03, 0.7)
    rb = synth.window(42, 350, 5, 0.04, 0.03, 0.03, 0.7)
    p = api.Params(result=1)
    want_a, calls_a = oracle_lib.window(p, ra.seq, ra.off)
    want_b, _ = oracle_lib.window(p, rb.seq, rb.off)
    L = api.lib()

    def call(reads, cap):
        seq = np.ascontiguousarray(reads.seq, np.uint8)
        off = np.ascontiguousarray(reads.off, np.int64)
        need, n_stats = ctypes.c_int64(), ctypes.c_int32()
        buf = np.zeros(max(cap, 1), np.uint8) if cap is not None else None
        stats = (api.CallStat * 64)()
        rc = L.mpoa_align_reads(ctx.h, ctypes.byref(p), len(off) - 1, api._ptr(seq), api._ptr(off),
                                None, None, None, None, None if buf is None else api._ptr(buf),
                                0 if buf is None else cap, ctypes.byref(need), stats, 64,
                                ctypes.byref(n_stats))
        text = buf[:need.value].tobytes().decode() if buf is not None and rc == 0 else None
        return rc, need.value, text, [stats[i].score for i in range(n_stats.value)]

    _say("held text: sizes of A, then B in full, then A short and in full")
    rc, need_a, _, _ = call(ra, None)
    assert rc == 0 and need_a == len(want_a)
    rc, need_b, text, _ = call(rb, 1 << 20)
    assert rc == 0 and text == want_b
    rc, need, _, _ = call(ra, need_a - 1)
    assert rc == -4 and need == need_a  # MPOA_E_OUT_CAP
    rc, need, text, scores = call(ra, need_a)
    assert rc == 0 and text == want_a and scores == [c.score for c in calls_a]
    rc, need, text, _ = call(ra, need_a)  # delivered: the next call runs again
    assert rc == 0 and text == want_a


@pytest.mark.parametrize("forms", [api.FORM_SORT_HBM, api.FORM_SORT_LANE, api.FORM_BT_LANE,
                                   api.FORM_ADD_LANE, api.FORM_BT_RANKS, 15])
def test_fallback_forms_match_oracle(ctx, forms):
    """Every read through the forms taken otherwise only for unusual graphs: the sort over a CSR
    in HBM with atomic degrees, the one-lane sort, backtrack and add_path (graph.cpp:59-123,
    align.cpp:1572-1736, graph.cpp:170-235), and the backtrack's runs along consecutive ranks."""
    b = synth.config3(seed=25, n_windows=24)
    for result in (0, 1):
        p = api.Params(result=result)
        _say("fallback forms", forms, "result", result)
        bat = api.Batch(ctx, p, b.win_off, b.seq, b.seq_off)
        bat.set_forms(forms)
        bat.execute()
        got = bat.results().texts
        bat.set_forms(0)
        bat.execute()
        again = bat.results().texts
        bat.free()
        want = oracle_lib.batch(p, b.win_off, b.seq, b.seq_off)
        assert got == want
        assert again == want


def test_fallback_forms_reject_unknown_bits(ctx):
    b = synth.config3(seed=26, n_windows=2)
    bat = api.Batch(ctx, api.Params(), b.win_off, b.seq, b.seq_off)
    with pytest.raises(api.MpoaError):
        bat.set_forms(256)
    bat.free()


@pytest.mark.parametrize("forms", [api.FORM_CONS_ROUNDS, 31])
def test_seeded_fallback_forms_match_oracle(ctx, forms):
    """-S windows with the fallback forms forced through the context: the consensus settled in
    rounds (graph.cpp:294-350), and with 31 every other fallback form as well."""
    ctx.set_forms(forms)
    try:
        reads = synth.window(33, 1500, 10, 0.04, 0.03, 0.03, 0.7)
        _say("seeded fallback forms", forms)
        _check_window(ctx, reads, seeding=1)
        reads = synth.window(34, 4000, 5, 0.024, 0.048, 0.048, 0.7)
        _check_window(ctx, reads, seeding=1, k=15, mm_w=10, result=1)
    finally:
        ctx.set_forms(0)
