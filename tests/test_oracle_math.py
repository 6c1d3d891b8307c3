"""The CPU oracle against what the mathematics fixes, independently of minipoa's code shape
(Q=8 blocks, bands, sentinels, the I/H scan):

- a linear graph (two reads) scores as textbook global affine alignment (Gotoh) of the two
  reads, each prefixed with the N sentinel that matches N (+M) (src/align.cpp:1363,1453);
- with the full band (-f 0) every finite H/D/I cell of a DAG equals a brute-force DAG Gotoh
  over all cells (values below -1e9 stand for minus infinity: their offsets differ);
- the row-column MSA of two reads, rescored with affine gaps, gives the DP score, and every
  MSA row without gaps is its read;
- graph invariants after add_path: edge weights count the reads using the edge, in- and
  out-lists agree, and rank is a topological order;
- the I/H scan equals its prefix-max form, exactly in integers (the form the device uses);
- consensus special cases: one read, identical reads, a minority substitution.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib

NEG = -(10 ** 12)
CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}

SCORES = [dict(match=2, mismatch=-4, gap_open=-4, gap_ext=-2),
          dict(match=3, mismatch=-5, gap_open=-6, gap_ext=-1),
          dict(match=1, mismatch=-2, gap_open=-3, gap_ext=-3)]


def norm(p):
    """o and e as minipoa uses them: the first gap base costs O+E (src/align.cpp:1369)."""
    return p["match"], p["mismatch"], p["gap_open"] + p["gap_ext"], p["gap_ext"]


def mutate(rng, t: str, err: float) -> str:
    out = []
    for ch in t:
        r = rng.random()
        if r < err / 3:
            out.append("ACGT"[rng.integers(4)])
        elif r < 2 * err / 3:
            out.append(ch + "ACGT"[rng.integers(4)])
        elif r < err:
            continue
        else:
            out.append(ch)
    s = "".join(out)
    return s if s else "A"


def reads_arrays(reads: list[str]):
    off = np.zeros(len(reads) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return np.frombuffer("".join(reads).encode(), np.uint8).copy(), off


def gotoh(x: list[int], y: list[int], M, X, o, e) -> int:
    """Textbook global affine alignment score of code lists x and y (gap of length L costs
    o + (L-1)e)."""
    n, m = len(x), len(y)
    H = [[NEG] * (m + 1) for _ in range(n + 1)]
    D = [[NEG] * (m + 1) for _ in range(n + 1)]
    I = [[NEG] * (m + 1) for _ in range(n + 1)]
    H[0][0] = 0
    for i in range(n + 1):
        for j in range(m + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                D[i][j] = max(D[i - 1][j] + e, H[i - 1][j] + o)
            if j > 0:
                I[i][j] = max(I[i][j - 1] + e, H[i][j - 1] + o)
            h = max(D[i][j], I[i][j])
            if i > 0 and j > 0:
                h = max(h, H[i - 1][j - 1] + (M if x[i - 1] == y[j - 1] else X))
            H[i][j] = h
    return H[n][m]


def codes(s: str) -> list[int]:
    return [CODE.get(ord(c.upper()), 4) for c in s]


@pytest.mark.parametrize("si", range(len(SCORES)))
def test_linear_graph_is_gotoh(si):
    rng = np.random.default_rng(1000 + si)
    p = SCORES[si]
    M, X, o, e = norm(p)
    for trial in range(25):
        L = int(rng.integers(5, 60))
        t = "".join("ACGT"[k] for k in rng.integers(0, 4, L))
        a, b = mutate(rng, t, 0.15), mutate(rng, t, 0.15)
        seq, off = reads_arrays([a, b])
        for f in (40, 0):
            _, stats = oracle_lib.window(dict(p, band_f=f), seq, off)
            assert len(stats) == 1
            want = gotoh([4] + codes(a), [4] + codes(b), M, X, o, e)
            assert stats[0].score == want, (a, b, f)


def brute_dag(g: dict, q: list[int], M, X, o, e):
    """Full DAG Gotoh over ranks and columns 0..m, minipoa's conventions: row r's M-source
    compares the predecessor's base (N for the source row) with seq[c-1]; D takes the
    predecessor's H/D at the same column; I runs along the row."""
    N = g["n_nodes"]
    rank_to_id = g["rank_to_id"]
    rank = np.empty(N, np.int64)
    rank[rank_to_id] = np.arange(N)
    seq = [4] + q
    m = len(seq)
    H = np.full((N, m + 1), NEG, np.int64)
    D = np.full((N, m + 1), NEG, np.int64)
    I = np.full((N, m + 1), NEG, np.int64)
    H[0, 0] = 0
    for r in range(1, N):
        u = rank_to_id[r]
        preds = [int(rank[s]) for s in g["in_src"][g["in_off"][u]:g["in_off"][u + 1]]]
        for c in range(m + 1):
            best = NEG
            dv = NEG
            for pr in preds:
                pb = 4 if pr == 0 else int(g["base"][rank_to_id[pr]])
                if c >= 1 and H[pr, c - 1] > NEG // 2:
                    best = max(best, H[pr, c - 1] + (M if pb == seq[c - 1] else X))
                dv = max(dv, D[pr, c] + e, H[pr, c] + o)
            iv = max(I[r, c - 1] + e, H[r, c - 1] + o) if c >= 1 else NEG
            D[r, c] = dv if dv > NEG // 2 else NEG
            I[r, c] = iv if iv > NEG // 2 else NEG
            H[r, c] = max(best, D[r, c], I[r, c])
    return H, D, I, m


@pytest.mark.parametrize("seed", range(6))
def test_full_band_dag_matches_brute_force(seed):
    rng = np.random.default_rng(2000 + seed)
    p = dict(SCORES[seed % len(SCORES)], band_f=0)
    M, X, o, e = norm(p)
    L = int(rng.integers(6, 16))
    t = "".join("ACGT"[k] for k in rng.integers(0, 4, L))
    reads = [mutate(rng, t, 0.25) for _ in range(int(rng.integers(3, 6)))]
    seq, off = reads_arrays(reads)
    seq_g, off_g = reads_arrays(reads[:-1])
    g = oracle_lib.graph(p, seq_g, off_g)
    dm = oracle_lib.dp_matrix(p, seq, off)
    H, D, I, m = brute_dag(g, codes(reads[-1]), M, X, o, e)
    assert dm["n_rows"] == g["n_nodes"]
    finite = 0
    for r in range(dm["n_rows"]):
        c0, c1 = 8 * dm["Bs"][r], 8 * (dm["Be"][r] - 1)
        assert c0 == 0 and c1 >= m
        base = dm["row_off"][r]
        for c in range(m + 1):
            for name, want, got in (("H", H, dm["H"]), ("D", D, dm["D"]), ("I", I, dm["I"])):
                if r == 0 and name == "I":
                    continue  # row 0's I is never written (src/align.cpp:1410-1423)
                v = int(got[base + c - c0])
                if want[r, c] > NEG // 2:
                    assert v == want[r, c], (name, r, c)
                    finite += 1
                else:
                    assert v < -10 ** 9, (name, r, c)
    assert finite > 0
    assert dm["score"] == H[dm["n_rows"] - 1, m]


def msa_rows(text: str) -> list[str]:
    lines = text.splitlines()
    return [lines[i + 1] for i in range(0, len(lines), 2)]


def rescore(a: str, b: str, M, X, o, e) -> int:
    """Affine score of two aligned rows (columns with two gaps do not occur for two reads)."""
    s, gap = 0, None
    for x, y in zip(a, b):
        if x == "-" or y == "-":
            kind = "D" if y == "-" else "I"
            s += e if gap == kind else o
            gap = kind
        else:
            s += M if x == y else X
            gap = None
    return s


@pytest.mark.parametrize("si", range(len(SCORES)))
def test_two_read_msa_rescores_to_dp_score(si):
    rng = np.random.default_rng(3000 + si)
    p = dict(SCORES[si], result=1)
    M, X, o, e = norm(p)
    for trial in range(20):
        L = int(rng.integers(8, 80))
        t = "".join("ACGT"[k] for k in rng.integers(0, 4, L))
        a, b = mutate(rng, t, 0.2), mutate(rng, t, 0.2)
        seq, off = reads_arrays([a, b])
        text, stats = oracle_lib.window(p, seq, off, ["a", "b"], ["", ""])
        ra, rb = msa_rows(text)
        assert ra.replace("-", "") == a and rb.replace("-", "") == b
        assert M + rescore(ra, rb, M, X, o, e) == stats[0].score


@pytest.mark.parametrize("seed", range(4))
def test_graph_invariants(seed):
    rng = np.random.default_rng(4000 + seed)
    t = "".join("ACGT"[k] for k in rng.integers(0, 4, 300))
    reads = [mutate(rng, t, 0.12) for _ in range(12)]
    seq, off = reads_arrays(reads)
    g = oracle_lib.graph({}, seq, off)
    N = g["n_nodes"]
    # every read path has len+1 edges (source -> first base ... last base -> sink)
    assert int(g["in_w"].sum()) == sum(len(r) + 1 for r in reads)
    assert int(g["out_w"].sum()) == int(g["in_w"].sum())
    ein = {}
    for v in range(N):
        for k in range(g["in_off"][v], g["in_off"][v + 1]):
            ein[(int(g["in_src"][k]), v)] = int(g["in_w"][k])
    eout = {}
    for u in range(N):
        for k in range(g["out_off"][u], g["out_off"][u + 1]):
            eout[(u, int(g["out_dst"][k]))] = int(g["out_w"][k])
    assert ein == eout
    assert all(1 <= w <= len(reads) for w in ein.values())
    rank = np.empty(N, np.int64)
    rank[g["rank_to_id"]] = np.arange(N)
    assert sorted(g["rank_to_id"].tolist()) == list(range(N))
    assert all(rank[u] < rank[v] for (u, v) in ein)
    assert g["rank_to_id"][0] == 0 and g["rank_to_id"][N - 1] == 1  # source first, sink last


def scan_serial(G: np.ndarray, o: int, e: int):
    I = np.empty_like(G)
    H = G.copy()
    I[0] = -1061109568
    H[0] = max(H[0], I[0])
    for j in range(1, len(G)):
        I[j] = max(I[j - 1] + e, H[j - 1] + o)
        H[j] = max(H[j], I[j])
    return H, I


def scan_prefix_max(G: np.ndarray, o: int, e: int):
    """I[j] = j*e + max(NEG_INF, max_{k<j} (G[k] + o - (k+1)*e)), valid because o <= e."""
    n = len(G)
    k = np.arange(n, dtype=np.int64)
    t = G.astype(np.int64) + o - (k + 1) * e
    pm = np.maximum.accumulate(t)
    I = np.empty(n, np.int64)
    I[0] = -1061109568
    I[1:] = np.maximum(-1061109568 + k[1:] * e, pm[:-1] + k[1:] * e)
    return np.maximum(G, I), I


def test_scan_prefix_max_identity():
    rng = np.random.default_rng(5)
    for o, e in ((-6, -2), (-7, -1), (-6, -3), (-4, -4)):
        for _ in range(200):
            n = int(rng.integers(1, 300))
            G = rng.integers(-500, 500, n).astype(np.int64)
            G[rng.random(n) < 0.3] = -1061109568 + rng.integers(-2000, 2000)
            h1, i1 = scan_serial(G, o, e)
            h2, i2 = scan_prefix_max(G, o, e)
            fin = h1 > -10 ** 9
            assert np.array_equal(h1[fin], h2[fin]) and np.array_equal(i1 > -10 ** 9, i2 > -10 ** 9)
            assert np.array_equal(fin, h2 > -10 ** 9)


def consensus(reads: list[str]) -> str:
    seq, off = reads_arrays(reads)
    text, _ = oracle_lib.window({}, seq, off)
    assert text.startswith(">consensus sequence\n")
    return text.splitlines()[1]


def test_consensus_special_cases():
    rng = np.random.default_rng(6)
    t = "".join("ACGT"[k] for k in rng.integers(0, 4, 200))
    assert consensus([t]) == t
    assert consensus([t] * 5) == t
    assert consensus([t.lower()] * 3) == t
    v = t[:100] + ("A" if t[100] != "A" else "C") + t[101:]
    assert consensus([t, t, v, t]) == t
    assert consensus([v, t, t]) == t
