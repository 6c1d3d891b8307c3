"""The seeded synthetic-read generator: determinism and the workload shapes it promises."""
from __future__ import annotations

import numpy as np

from poa_amd import synth


def test_deterministic():
    a = synth.config3(seed=77, n_windows=50)
    b = synth.config3(seed=77, n_windows=50)
    assert np.array_equal(a.seq, b.seq) and np.array_equal(a.seq_off, b.seq_off)
    assert np.array_equal(a.win_off, b.win_off)
    c = synth.config3(seed=78, n_windows=50)
    assert not np.array_equal(a.seq[:1000], c.seq[:1000])


def test_config3_shape():
    b = synth.config3(seed=synth.SEEDS[3], n_windows=400)
    nr = np.diff(b.win_off)
    assert nr.min() >= 5 and nr.max() <= 60
    lens = np.diff(b.seq_off)
    assert lens.min() > 0
    # read lengths follow the template (150-2,000 bp) within the indel noise
    assert 100 < np.median(lens) < 1000 and lens.max() < 2600
    assert set(np.unique(b.seq).tolist()) <= set(b"ACGT")


def test_window_rates():
    w = synth.window(9, 20000, 20, 0.04, 0.03, 0.03)
    lens = np.diff(w.off)
    assert w.n == 20
    # insertions and deletions balance: mean length stays near the template length
    assert abs(lens.mean() - 20000) < 400
