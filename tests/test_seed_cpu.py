"""-S host seeding of the product library against the oracle's (no GPU): identical anchors
for seeded reads against consensus-like templates, across k, w and -W."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle_lib
from poa_amd import api, synth


def _anchors_product(p, seq, off, rid, cons: bytes):
    L = api.lib()
    L.mpoa_seed_anchors.argtypes = [ctypes.POINTER(api.Params), ctypes.c_int32, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p,
                                    ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.POINTER(ctypes.c_int64)]
    n = ctypes.c_int64()
    cap = 1 << 16
    out = np.zeros(2 * cap, np.uint64)
    rc = L.mpoa_seed_anchors(ctypes.byref(p), len(off) - 1, seq.ctypes.data, off.ctypes.data, rid,
                             cons, len(cons), out.ctypes.data, cap, ctypes.byref(n))
    assert rc == 0
    return out[:2 * n.value].reshape(-1, 2)


def _anchors_oracle(p, seq, off, rid, cons: bytes):
    L = oracle_lib.lib()
    L.orc_anchors.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p,
                              ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    op = oracle_lib.oparams(p)
    n = ctypes.c_int64()
    cap = 1 << 16
    out = np.zeros(2 * cap, np.uint64)
    rc = L.orc_anchors(ctypes.byref(op), len(off) - 1, seq.ctypes.data, off.ctypes.data, rid, cons,
                       len(cons), out.ctypes.data, cap, ctypes.byref(n))
    assert rc == 0
    return out[:2 * n.value].reshape(-1, 2)


@pytest.mark.parametrize("case", [
    dict(tlen=2000, n=12, err=(0.04, 0.03, 0.03), flags={}),
    dict(tlen=2000, n=12, err=(0.04, 0.03, 0.03), flags=dict(k=15, mm_w=10)),
    dict(tlen=6000, n=6, err=(0.024, 0.048, 0.048), flags=dict(k=15, mm_w=10)),
    dict(tlen=3000, n=8, err=(0.01, 0.01, 0.01), flags=dict(k=11, mm_w=5, poa_w=300)),
    dict(tlen=600, n=5, err=(0.01, 0.01, 0.01), flags={}),
])
def test_product_seeding_equals_oracle(case):
    reads = synth.window(4242 + case["tlen"], case["tlen"], case["n"], *case["err"], 0.7)
    seq = np.ascontiguousarray(reads.seq, np.uint8)
    off = np.ascontiguousarray(reads.off, np.int64)
    p = api.Params(seeding=1, **case["flags"])
    total = 0
    for rid in range(1, reads.n):
        for cons_src in (0, rid - 1):
            cons = reads.read(cons_src)
            a = _anchors_product(p, seq, off, rid, cons)
            b = _anchors_oracle(p, seq, off, rid, cons)
            assert np.array_equal(a, b), (rid, cons_src)
            total += len(a)
    assert total > 0


def test_guide_tree_order_matches_oracle():
    """The product's progressive order (-p, minimizer.cpp:195-386; host code, no GPU) equals the
    oracle's on every -p fixture and on seeded read sets of uneven length."""
    import golden_util as gu
    import oracle_lib as ol
    from poa_amd import api, synth
    sets = []
    for c in gu.load_cases():
        if "-p" in c["args"]:
            seq, off, _, _ = gu.parse_fasta(c["fasta"])
            sets.append((gu.parse_args(c["args"]), seq, off))
    b = synth.config3(seed=0x9e, n_windows=12)
    for w in range(b.n_windows):
        r = b.window(w)
        sets.append(({"progressive": 1}, r.seq, r.off))
    for p, seq, off in sets:
        want = ol.Dag(p, seq, off).order()
        got = api.Seeder(api.Params(**p), seq, off).order()
        assert list(got) == list(want)
        assert sorted(got) == list(range(len(off) - 1))


def test_seeding_with_gfa_paths_equals_oracle():
    """minipoa -i's read set starts with the GFA's paths as reads without bases (src/main.cpp:86,
    minimizers over all of them at :98): the reads after them seed as in the oracle."""
    import gzip
    import json
    import os

    import golden_util as gu
    fix = json.load(gzip.open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                           "gfa_input.json.gz"), "rt"))
    total = 0
    for inp in fix["inputs"].values():
        seq, off, _, _, n_paths = gu.gfa_read_set(inp["gfa"], inp["fasta"])
        p = api.Params(seeding=1)
        for rid in range(n_paths, len(off) - 1):
            cons = bytes(seq[off[n_paths]:off[n_paths + 1]])
            a = _anchors_product(p, seq, off, rid, cons)
            b = _anchors_oracle(p, seq, off, rid, cons)
            assert np.array_equal(a, b), rid
            total += len(a)
    assert total > 0
