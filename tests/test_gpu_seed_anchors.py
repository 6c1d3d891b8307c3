"""The device seeding's anchors against the oracle's (mpoa_seed_anchors_device against
orc_anchors): the minimizers, sorts, joins, local and global chains of the -S windows, run on
one workgroup for a read and a consensus, as x/y pairs with the band-mode mark. The same read
sets and consensus choices as the host seeding's check (tests/test_seed_cpu.py), plus a
repeat-rich set whose equal minimizer keys take the radix sort's own order."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle_lib
from poa_amd import api, synth
from test_seed_cpu import _anchors_oracle


def _anchors_device(ctx, p, seq, off, rid, cons: bytes):
    L = api.lib()
    L.mpoa_seed_anchors_device.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(api.Params), ctypes.c_int32, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p,
        ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    n = ctypes.c_int64()
    cap = 1 << 16
    out = np.zeros(2 * cap, np.uint64)
    rc = L.mpoa_seed_anchors_device(ctx.h, ctypes.byref(p), len(off) - 1, seq.ctypes.data,
                                    off.ctypes.data, rid, cons, len(cons), out.ctypes.data, cap,
                                    ctypes.byref(n))
    assert rc == 0, rc
    return out[:2 * n.value].reshape(-1, 2)


def _repeat_window(seed: int):
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    unit = rng.integers(0, 4, 70)
    tpl = np.concatenate([rng.integers(0, 4, 400), np.tile(unit, 14), rng.integers(0, 4, 400)])
    seqs = []
    for _ in range(6):
        s = tpl.copy()
        hit = rng.random(len(s)) < 0.02
        s[hit] = rng.integers(0, 4, int(hit.sum()))
        seqs.append(acgt[s])
    off = np.zeros(len(seqs) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    return np.concatenate(seqs), off


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(tlen=2000, n=12, err=(0.04, 0.03, 0.03), flags={}),
    dict(tlen=2000, n=12, err=(0.04, 0.03, 0.03), flags=dict(k=15, mm_w=10)),
    dict(tlen=6000, n=6, err=(0.024, 0.048, 0.048), flags=dict(k=15, mm_w=10)),
    dict(tlen=3000, n=8, err=(0.01, 0.01, 0.01), flags=dict(k=11, mm_w=5, poa_w=300)),
    dict(tlen=600, n=5, err=(0.01, 0.01, 0.01), flags={}),
    dict(tlen=0, n=0, err=None, flags={}),
])
def test_device_anchors_equal_oracle(ctx, case):
    if case["err"] is None:
        seq, off = _repeat_window(99)
        n_reads = len(off) - 1
        read = lambda i: seq[off[i]:off[i + 1]].tobytes()
    else:
        reads = synth.window(4242 + case["tlen"], case["tlen"], case["n"], *case["err"], 0.7)
        seq = np.ascontiguousarray(reads.seq, np.uint8)
        off = np.ascontiguousarray(reads.off, np.int64)
        n_reads = reads.n
        read = reads.read
    p = api.Params(seeding=1, **case["flags"])
    total = marked = 0
    for rid in range(1, n_reads):
        for cons_src in (0, rid - 1):
            cons = read(cons_src)
            print("anchors", case["tlen"], rid, cons_src, flush=True)
            a = _anchors_device(ctx, p, seq, off, rid, cons)
            b = _anchors_oracle(p, seq, off, rid, cons)
            assert np.array_equal(a, b), (rid, cons_src, len(a), len(b))
            total += len(a)
            marked += int(((b[:, 1] >> 40) & 1).sum()) if len(b) else 0
    assert total > 0
