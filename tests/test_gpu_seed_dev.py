"""Device seeding of -S windows: every read seeded by the workgroup that committed the read
before it (minimizers, anchors, local and global chains, sub-windows on the GPU), against the
host-seeded form (MPOA_FORM_HOST_SEED) and upstream's outputs. Both forms must split every read
into the same sub-windows: the per-call (read, rows, qlen, score) lists are compared whole."""
from __future__ import annotations

import numpy as np
import pytest

import full_util as fu
import golden_util as gu
import oracle_lib
from poa_amd import api, synth


def _run(ctx, params, seq, off, forms, names=None, comments=None):
    ctx.set_forms(forms)
    try:
        out, stats = ctx.align_reads(params, seq, off, names, comments, with_stats=True)
        how = ctx.last_seeding()
    finally:
        ctx.set_forms(0)
    return out, [(s.read, s.rows, s.qlen, s.score) for s in stats], how


def _outcome(ctx, params, seq, off, forms):
    """(text or None, calls, seeding form, status): minipoa's failures (an empty sub-window,
    a backtrack error) are statuses to compare, not test errors."""
    try:
        out, calls, how = _run(ctx, params, seq, off, forms)
        return out, calls, how, 0
    except api.MpoaError as e:
        return None, None, ctx.last_seeding(), e.rc


def _oracle(params, seq, off):
    try:
        return oracle_lib.window(params, seq, off)[0], 0
    except RuntimeError as e:  # "oracle error <status>"
        return None, int(str(e).split()[-1])


def _say(*a):
    print(*a, flush=True)


SEEDED = [c for c in fu.load_cases() if "-S" in c["args"]]
EMPTY = -5  # MPOA_E_EMPTY_WINDOW


@pytest.mark.gpu
@pytest.mark.parametrize("case", SEEDED, ids=[c["case"] for c in SEEDED])
def test_full_configs_seeded_on_device(ctx, case):
    reads = fu.reads_of(case)
    _, text = fu.fasta_of(reads)
    seq, off, names, comments = gu.parse_fasta(text)
    p = api.Params(**gu.parse_args(case["args"]))
    _say("seed_dev", case["case"])
    out, calls, how = _run(ctx, p, seq, off, 0, names, comments)
    assert how == 1, f"the window left device seeding ({how})"
    assert out == case["stdout"]
    assert [c[3] for c in calls] == case["scores"]
    out_h, calls_h, how_h = _run(ctx, p, seq, off, api.FORM_HOST_SEED, names, comments)
    assert how_h == 0
    assert out_h == out
    assert calls_h == calls


def _windows(seed: int, n: int):
    b = synth.config3(seed=seed, n_windows=n)
    for w in range(n):
        r0, r1 = int(b.win_off[w]), int(b.win_off[w + 1])
        o = b.seq_off[r0:r1 + 1]
        yield w, b.seq[o[0]:o[-1]].copy(), (o - o[0]).astype(np.int64)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [dict(seeding=1, result=0), dict(seeding=1, result=1),
                                   dict(seeding=1, k=15, mm_w=5, result=0),
                                   dict(seeding=1, k=11, mm_w=3, poa_w=40, result=1)],
                         ids=["S_r0", "S_r1", "S_k15w5", "S_k11w3W40"])
def test_random_windows_against_oracle(ctx, flags):
    p = api.Params(**flags)
    for w, seq, off in _windows(31 + flags.get("k", 19), 12):
        _say("seed_dev window", w, flags)
        want, want_rc = _oracle(p, seq, off)
        out, calls, how, rc = _outcome(ctx, p, seq, off, 0)
        assert how == 1, f"window {w} left device seeding ({how})"
        if rc == EMPTY and want_rc != 0:
            # anchors overlapping by more than one base on the read: a sub-window of query
            # length below -1, where minipoa's DP reads past its query and exits one way or
            # another; the library stops the window
            pass
        else:
            assert (rc, out) == (want_rc, want), f"window {w}"
        out_h, calls_h, how_h, rc_h = _outcome(ctx, p, seq, off, api.FORM_HOST_SEED)
        assert how_h == 0
        assert (rc_h, out_h, calls_h) == (rc, out, calls), f"window {w}"


def _repeat_reads(seed: int, n_reads: int, unit: int, copies: int):
    """Reads of a template holding a tandem repeat (identical minimizers within each read)."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    rep = np.tile(rng.integers(0, 4, unit), copies)
    tpl = np.concatenate([rng.integers(0, 4, 300), rep, rng.integers(0, 4, 300)])
    seqs = []
    for _ in range(n_reads):
        s = tpl.copy()
        hit = rng.random(len(s)) < 0.03
        s[hit] = rng.integers(0, 4, int(hit.sum()))
        seqs.append(acgt[s])
    off = np.zeros(n_reads + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return np.concatenate(seqs), off


@pytest.mark.gpu
@pytest.mark.parametrize("unit,copies", [(40, 3), (90, 12)])
def test_repeats_against_oracle(ctx, unit, copies):
    """Equal minimizer keys within reads and consensus: the device repeats minipoa's radix
    sort steps for them (the order of equal keys decides which anchors chain)."""
    seq, off = _repeat_reads(7 + unit, 8, unit, copies)
    p = api.Params(seeding=1, result=1)
    _say("seed_dev repeats", unit, copies)
    want, want_rc = _oracle(p, seq, off)
    out, calls, how, rc = _outcome(ctx, p, seq, off, 0)
    assert how == 1, f"left device seeding ({how})"
    assert (rc, out) == (want_rc, want)
    out_h, calls_h, _, rc_h = _outcome(ctx, p, seq, off, api.FORM_HOST_SEED)
    assert (rc_h, out_h, calls_h) == (rc, out, calls)


def _finishing_windows(p, seed: int, n: int):
    """A seeded synthetic batch restricted to the windows minipoa finishes, with their texts."""
    b = synth.config3(seed=seed, n_windows=n)
    keep, want = [], []
    for w in range(n):
        r0, r1 = int(b.win_off[w]), int(b.win_off[w + 1])
        o = b.seq_off[r0:r1 + 1]
        text, rc = _oracle(p, b.seq[o[0]:o[-1]].copy(), (o - o[0]).astype(np.int64))
        if rc == 0:
            keep.append(w)
            want.append(text)
    return b.subset(keep), want


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [dict(seeding=1, result=0), dict(seeding=1, result=1),
                                   dict(seeding=1, result=2),
                                   dict(seeding=1, k=15, mm_w=5, result=3),
                                   dict(seeding=1, progressive=1, result=1)],
                         ids=["S_r0", "S_r1", "S_r2", "S_k15w5_r3", "S_p_r1"])
def test_seeded_batch_against_oracle(ctx, flags):
    """-S windows in the batch API: one workgroup per window seeds every read itself and runs
    its sub-windows in turn (k_seeded); the host-seeded form gives the same texts."""
    p = api.Params(**flags)
    sub, want = _finishing_windows(p, 77 + flags.get("k", 19), 48)
    assert len(want) >= 30
    _say("seeded batch", flags, len(want), "windows")
    bat = api.Batch(ctx, p, sub.win_off, sub.seq, sub.seq_off)
    bat.execute()
    got = bat.results()
    assert got.texts == want
    # a window past its estimate re-runs in a second seeded launch; few should
    assert got.stats.reruns <= len(want) // 10, "windows left the seeded launches"
    assert 1 <= got.stats.launches <= 3
    bat.set_forms(api.FORM_HOST_SEED)
    bat.execute()
    assert bat.results().texts == want
    bat.free()


# Windows of the bench workload (synth.config3(), default seed) whose reads under -S have two
# anchors overlapping by one base: sub-windows of query length -1, which minipoa runs as DP
# calls without res entries (window 21 has 27 of them, some over 4 rows).
OVERLAP = [14, 21, 22, 61, 104, 110, 118, 170, 175, 182, 189, 209, 243, 344, 372, 392]


@pytest.mark.gpu
def test_overlapping_anchors_against_oracle(ctx):
    full = synth.config3()
    p = api.Params(seeding=1, result=0)
    n_null = 0
    for w in OVERLAP:
        win = full.window(w)
        _say("seed_dev overlap window", w)
        want, ocalls = oracle_lib.window(p, win.seq, win.off)
        want_calls = [(c.read, c.rows, c.qlen, c.score) for c in ocalls]
        n_null += sum(1 for c in want_calls if c[2] == -1)
        for forms in (0, api.FORM_HOST_SEED):
            out, calls, how = _run(ctx, p, win.seq, win.off, forms)
            assert how == (1 if forms == 0 else 0)
            assert out == want, f"window {w} forms {forms}"
            assert calls == want_calls, f"window {w} forms {forms}"
    assert n_null >= 50
    sub = full.subset(OVERLAP)
    bat = api.Batch(ctx, p, sub.win_off, sub.seq, sub.seq_off)
    bat.execute()
    assert bat.results().texts == [oracle_lib.window(p, full.window(w).seq,
                                                     full.window(w).off)[0] for w in OVERLAP]
    bat.free()
