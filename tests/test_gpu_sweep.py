"""Random parameter sets against the oracle: scores (-M -X -O -E), bands (-b -f, -B), output
modes (-r 0/1/2) and progressive order (-p) on small random windows (one read to fourteen,
30-700 bp templates, 0-25% error), through the single-window call (mpoa_align_reads) and the
batched launches (mpoa_batch_*). minipoa's own failures (a band violation, align.cpp:1584-1588;
a backtrack failure, :1730-1731) are statuses to compare: the oracle's status and the GPU's must
be the same, and where both finish, the text and every call's score.

Every case prints what it runs before touching the GPU, so a failing run names its case.
"""
import numpy as np
import pytest

import oracle_lib
from poa_amd import api, synth

pytestmark = pytest.mark.gpu

N_SETS = 40
WINDOWS_PER_SET = 6
E_ARG = -1  # MPOA_E_ARG


def _say(*a):
    print("[case]", *a, flush=True)


def _param_set(rng) -> dict:
    return dict(match=int(rng.integers(1, 7)), mismatch=int(rng.integers(1, 9)),
                gap_open=int(rng.integers(0, 9)), gap_ext=int(rng.integers(1, 5)),
                band_b=int(rng.choice([3, 10, 40, 100])), band_f=int(rng.choice([0, 5, 20, 40])),
                ab_band=int(rng.integers(0, 2)), result=int(rng.integers(0, 3)),
                progressive=int(rng.random() < 0.25))


def _windows(rng, seed: int) -> list[synth.Reads]:
    out = []
    for w in range(WINDOWS_PER_SET):
        err = float(rng.uniform(0.0, 0.25))
        split = rng.dirichlet([1.0, 1.0, 1.0]) * err
        out.append(synth.window(seed * 100 + w, int(rng.integers(30, 700)),
                                int(rng.integers(1, 15)), float(split[0]), float(split[1]),
                                float(split[2]), float(rng.choice([0.0, 0.7]))))
    return out


def _oracle(p, reads):
    try:
        text, calls = oracle_lib.window(p, reads.seq, reads.off)
        return text, [c.score for c in calls], 0
    except RuntimeError as e:  # "oracle error <status>"
        return None, None, int(str(e).split()[-1])


def _gpu(ctx, p, reads):
    try:
        text, stats = ctx.align_reads(p, reads.seq, reads.off, with_stats=True)
        return text, [s.score for s in stats], 0
    except api.MpoaError as e:
        return None, None, e.rc


def _check_batch(ctx, p, wins, want):
    """The same windows as one batch: texts of finished windows, statuses of the others."""
    seq = np.concatenate([r.seq for r in wins])
    win_off = np.zeros(len(wins) + 1, np.int32)
    seq_off, base = [np.zeros(1, np.int64)], 0
    for i, r in enumerate(wins):
        seq_off.append(r.off[1:] + base)
        base += int(r.off[-1])
        win_off[i + 1] = win_off[i] + r.n
    seq_off = np.concatenate(seq_off)
    if all(rc == E_ARG for _, _, rc in want):  # flags refused as a whole (-p with -r 2)
        with pytest.raises(api.MpoaError) as e:
            api.Batch(ctx, p, win_off, seq, seq_off)
        assert e.value.rc == E_ARG
        return
    bat = api.Batch(ctx, p, win_off, seq, seq_off)
    first = next((rc for _, _, rc in want if rc != 0), 0)
    try:
        bat.execute()
        rc_all = 0
    except api.MpoaError as e:  # the first window that did not finish
        rc_all = e.rc
    assert rc_all == first
    texts = bat.results().texts
    ws, _ = bat.window_stats()
    bat.free()
    for w, (text, _, rc) in enumerate(want):
        assert ws[w].status == rc, f"window {w}: status {ws[w].status} against {rc}"
        if rc == 0:
            assert texts[w] == text, f"window {w}"


def _case_ids():
    return [f"set{i}" for i in range(N_SETS)]


@pytest.mark.parametrize("k", range(N_SETS), ids=_case_ids())
def test_random_parameters_match_oracle(ctx, k):
    rng = np.random.default_rng(9000 + k)
    flags = _param_set(rng)
    p = api.Params(**flags)
    wins = _windows(rng, 9000 + k)
    want = []
    for w, reads in enumerate(wins):
        _say("sweep", k, flags, "window", w, "reads", reads.n, "bases", int(reads.off[-1]))
        o = _oracle(p, reads)
        g = _gpu(ctx, p, reads)
        assert g[2] == o[2], f"status {g[2]} against the oracle's {o[2]}"
        assert g[0] == o[0]
        assert g[1] == o[1]
        want.append(o)
    _say("sweep", k, "batch of", len(wins))
    _check_batch(ctx, p, wins, want)


def _seeded_set(rng) -> dict:
    return dict(match=int(rng.integers(1, 7)), mismatch=int(rng.integers(1, 9)),
                gap_open=int(rng.integers(0, 9)), gap_ext=int(rng.integers(1, 5)),
                band_b=int(rng.choice([10, 40, 100])), band_f=int(rng.choice([5, 20, 40])),
                seeding=1, k=int(rng.integers(11, 29)), mm_w=int(rng.integers(3, 21)),
                result=int(rng.integers(0, 3)), progressive=int(rng.random() < 0.25))


N_SEEDED = 16


@pytest.mark.parametrize("k", range(N_SEEDED), ids=[f"seeded{i}" for i in range(N_SEEDED)])
def test_random_seeded_parameters_match_oracle(ctx, k):
    """-S with random scores, bands, k-mer and window sizes on 0.3-3 kb windows: the device
    seeding, its sub-windows and the seeded batch against the oracle, statuses included."""
    rng = np.random.default_rng(9500 + k)
    flags = _seeded_set(rng)
    p = api.Params(**flags)
    wins = []
    for w in range(4):
        err = float(rng.uniform(0.02, 0.15))
        split = rng.dirichlet([1.0, 1.0, 1.0]) * err
        wins.append(synth.window(9500 * 100 + k * 10 + w, int(rng.integers(300, 3000)),
                                 int(rng.integers(2, 11)), float(split[0]), float(split[1]),
                                 float(split[2]), 0.7))
    want = []
    for w, reads in enumerate(wins):
        _say("seeded sweep", k, flags, "window", w, "reads", reads.n, "bases", int(reads.off[-1]))
        o = _oracle(p, reads)
        g = _gpu(ctx, p, reads)
        assert g[2] == o[2], f"status {g[2]} against the oracle's {o[2]}"
        assert g[0] == o[0]
        assert g[1] == o[1]
        want.append(o)
    _say("seeded sweep", k, "batch of", len(wins))
    _check_batch(ctx, p, wins, want)
