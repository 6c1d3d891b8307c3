"""ctypes binding of the CPU oracle (oracle/liboracle.so), the checker the GPU path is compared
against. Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg import this module."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "minipoa")
ORACLE_CLI = os.path.join(ROOT, "oracle", "oracle_cli")


class OParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "match", "mismatch", "gap_open", "gap_ext", "band_b", "band_f", "ab_band", "seeding",
        "k", "mm_w", "poa_w", "result", "progressive")]


class OCallStat(ctypes.Structure):
    _fields_ = [("read", ctypes.c_int32), ("rows", ctypes.c_int32), ("qlen", ctypes.c_int32),
                ("score", ctypes.c_int32), ("ab_band", ctypes.c_int32),
                ("max_width", ctypes.c_int32), ("cells", ctypes.c_int64),
                ("edge_cells", ctypes.c_int64), ("logical_cells", ctypes.c_int64),
                ("graph_nodes", ctypes.c_int32), ("graph_edges", ctypes.c_int32)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} not built (make -C oracle)")
        _lib = ctypes.CDLL(ORACLE_SO)
    return _lib


def oparams(p) -> OParams:
    """From poa_amd.api.Params (or a dict) to the oracle's struct."""
    d = p.as_dict() if hasattr(p, "as_dict") else dict(p)
    base = dict(match=2, mismatch=-4, gap_open=-4, gap_ext=-2, band_b=100, band_f=40,
                ab_band=0, seeding=0, k=19, mm_w=10, poa_w=0, result=0, progressive=0)
    base.update(d)
    return OParams(**base)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _names(names):
    if not names:
        return None, None
    enc = [n.encode() for n in names]
    off = np.zeros(len(enc) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in enc])
    return np.frombuffer(b"".join(enc) + b"\0", np.uint8).copy(), off


def window(p, seq: np.ndarray, off: np.ndarray, names=None, comments=None):
    """minipoa's stdout for one read set, plus per-poa()-call statistics."""
    L = lib()
    seq = np.ascontiguousarray(seq, np.uint8)
    off = np.ascontiguousarray(off, np.int64)
    n = len(off) - 1
    op = oparams(p)
    nb, no = _names(names)
    cb, co = _names(comments)
    need = ctypes.c_int64()
    cap = 1 << 16
    stats = (OCallStat * cap)()
    ns = ctypes.c_int32()
    rc = L.orc_window(ctypes.byref(op), n, _p(seq), _p(off), _p(nb), _p(no), _p(cb), _p(co),
                      None, ctypes.c_int64(0), ctypes.byref(need), stats, cap, ctypes.byref(ns))
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    buf = np.zeros(max(need.value, 1), np.uint8)
    rc = L.orc_window(ctypes.byref(op), n, _p(seq), _p(off), _p(nb), _p(no), _p(cb), _p(co),
                      _p(buf), ctypes.c_int64(len(buf)), ctypes.byref(need), None, 0, None)
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    return buf[:need.value].tobytes().decode(), [stats[i] for i in range(ns.value)]


def batch(p, win_off, seq, seq_off, threads=0):
    """Independent windows on `threads` OpenMP threads; list of stdout texts."""
    L = lib()
    win_off = np.ascontiguousarray(win_off, np.int32)
    seq = np.ascontiguousarray(seq, np.uint8)
    seq_off = np.ascontiguousarray(seq_off, np.int64)
    nw = len(win_off) - 1
    op = oparams(p)
    out_off = np.zeros(nw + 1, np.int64)
    rc = L.orc_batch(ctypes.byref(op), nw, _p(win_off), _p(seq), _p(seq_off), threads, None,
                     ctypes.c_int64(0), _p(out_off), None, ctypes.c_int64(0), None)
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    buf = np.zeros(max(int(out_off[-1]), 1), np.uint8)
    rc = L.orc_batch(ctypes.byref(op), nw, _p(win_off), _p(seq), _p(seq_off), threads, _p(buf),
                     ctypes.c_int64(len(buf)), _p(out_off), None, ctypes.c_int64(0), None)
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    raw = buf.tobytes()
    return [raw[out_off[i]:out_off[i + 1]].decode() for i in range(nw)]


def graph(p, seq: np.ndarray, off: np.ndarray) -> dict:
    """The graph the oracle builds from all reads (orc_graph), as numpy arrays."""
    L = lib()
    seq = np.ascontiguousarray(seq, np.uint8)
    off = np.ascontiguousarray(off, np.int64)
    op = oparams(p)
    nn, ne = ctypes.c_int32(), ctypes.c_int32()
    n = len(off) - 1
    none = [None] * 13
    rc = L.orc_graph(ctypes.byref(op), n, _p(seq), _p(off), ctypes.byref(nn), ctypes.byref(ne),
                     *none)
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    N, E = nn.value, ne.value
    g = dict(base=np.zeros(N, np.uint8), in_off=np.zeros(N + 1, np.int32),
             in_src=np.zeros(E, np.int32), in_w=np.zeros(E, np.int32),
             out_off=np.zeros(N + 1, np.int32), out_dst=np.zeros(E, np.int32),
             out_w=np.zeros(E, np.int32), rank_to_id=np.zeros(N, np.int32),
             hlen=np.zeros(N, np.int32), tlen=np.zeros(N, np.int32), par_id=np.zeros(N, np.int32),
             aligned=np.zeros(5 * N, np.int32))
    order = ["base", "in_off", "in_src", "in_w", "out_off", "out_dst", "out_w", "rank_to_id",
             "hlen", "tlen", "par_id", "aligned"]
    rc = L.orc_graph(ctypes.byref(op), n, _p(seq), _p(off), ctypes.byref(nn), ctypes.byref(ne),
                     *[_p(g[k]) for k in order])
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    g["n_nodes"], g["n_edges"] = N, E
    return g


def dp_matrix(p, seq: np.ndarray, off: np.ndarray) -> dict:
    """H, D, I of the DP that aligns the last read to the graph of the reads before it."""
    L = lib()
    seq = np.ascontiguousarray(seq, np.uint8)
    off = np.ascontiguousarray(off, np.int64)
    op = oparams(p)
    n = len(off) - 1
    nr, tot, score = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
    rc = L.orc_dp_matrix(ctypes.byref(op), n, _p(seq), _p(off), ctypes.byref(nr),
                         ctypes.byref(tot), None, None, None, None, None, None,
                         ctypes.byref(score))
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    R, T = nr.value, tot.value
    m = dict(Bs=np.zeros(R, np.int32), Be=np.zeros(R, np.int32), row_off=np.zeros(R + 1, np.int64),
             H=np.zeros(T, np.int32), D=np.zeros(T, np.int32), I=np.zeros(T, np.int32))
    rc = L.orc_dp_matrix(ctypes.byref(op), n, _p(seq), _p(off), ctypes.byref(nr),
                         ctypes.byref(tot), _p(m["Bs"]), _p(m["Be"]), _p(m["row_off"]),
                         _p(m["H"]), _p(m["D"]), _p(m["I"]), ctypes.byref(score))
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")
    m["n_rows"], m["score"] = R, score.value
    return m


class Dag:
    """The oracle's call-level mirror (orc_dag_*): minipoa's poa(), alignment(),
    graph::add_path, graph::topsort and outputs on one graph. res arrays are int32 (n, 3):
    from, base code, aligned_id."""

    def __init__(self, p, seq: np.ndarray | None = None, off: np.ndarray | None = None):
        L = lib()
        L.orc_dag_new.restype = ctypes.c_void_p
        L.orc_dag_new.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                  ctypes.c_void_p]
        L.orc_dag_free.argtypes = [ctypes.c_void_p]
        L.orc_dag_free.restype = None
        self.op = oparams(p)
        self._seq = None if seq is None else np.ascontiguousarray(seq, np.uint8)
        self._off = None if off is None else np.ascontiguousarray(off, np.int64)
        n = 0 if off is None else len(off) - 1
        self.h = L.orc_dag_new(ctypes.byref(self.op), n, _p(self._seq), _p(self._off))
        if not self.h:
            raise RuntimeError("orc_dag_new failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dag_free(self.h)
            self.h = None

    def order(self) -> np.ndarray:
        """The order reads are added in: minipoa's guide tree with -p, else input order."""
        n = lib().orc_dag_order(ctypes.c_void_p(self.h), None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_dag_order(ctypes.c_void_p(self.h), _p(out), n)
        return out[:n]

    def _res(self, fn, *args, guess: int = 64) -> np.ndarray:
        """One call with room for `guess` entries; a second, sized one only when that was too
        small (the call is computed again then)."""
        n = ctypes.c_int64()
        out = np.zeros((max(guess, 1), 3), np.int32)
        rc = fn(ctypes.c_void_p(self.h), *args, _p(out), ctypes.c_int64(len(out)),
                ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")
        if n.value > len(out):
            out = np.zeros((n.value, 3), np.int32)
            rc = fn(ctypes.c_void_p(self.h), *args, _p(out), ctypes.c_int64(len(out)),
                    ctypes.byref(n))
            if rc != 0:
                raise RuntimeError(f"oracle error {rc}")
        return out[:n.value]

    def poa(self, beg_id: int, end_id: int, q: bytes, ab_band: bool):
        score = ctypes.c_int32()
        n = ctypes.c_int64()
        buf = ctypes.create_string_buffer(q, len(q) + 1)
        out = np.zeros((len(q) + 2, 3), np.int32)
        rc = lib().orc_dag_poa(ctypes.c_void_p(self.h), beg_id, end_id, buf, len(q),
                               int(ab_band), _p(out), ctypes.c_int64(len(out)), ctypes.byref(n),
                               ctypes.byref(score))
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")
        return out[:n.value], score.value

    def alignment(self, rid: int, q: bytes) -> np.ndarray:
        buf = ctypes.create_string_buffer(q, len(q) + 1)
        return self._res(lib().orc_dag_alignment, rid, buf, len(q), guess=2 * len(q) + 8)

    def add_path(self, seq_id: int, res: np.ndarray) -> None:
        res = np.ascontiguousarray(res, np.int32)
        rc = lib().orc_dag_add_path(ctypes.c_void_p(self.h), seq_id, _p(res),
                                    ctypes.c_int64(len(res)))
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")

    def topsort(self, op: int) -> None:
        if lib().orc_dag_topsort(ctypes.c_void_p(self.h), op) != 0:
            raise RuntimeError("oracle topsort failed")

    def scores(self) -> list[int]:
        """The scores of the DP calls completed on this graph, in order (-V 1's lines)."""
        n = lib().orc_dag_scores(ctypes.c_void_p(self.h), None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_dag_scores(ctypes.c_void_p(self.h), _p(out), n)
        return [int(x) for x in out[:n]]

    def read_gfa(self, text: bytes) -> list[str]:
        """read_gfa (minipoa -i) into this new graph; returns the path names (a name and its
        newline take no more room than its P line)."""
        buf = ctypes.create_string_buffer(text, len(text) + 1)
        out = ctypes.create_string_buffer(len(text) + 1)
        n, k = ctypes.c_int64(), ctypes.c_int32()
        rc = lib().orc_dag_read_gfa(ctypes.c_void_p(self.h), buf, ctypes.c_int64(len(text)), out,
                                    ctypes.c_int64(len(out)), ctypes.byref(n), ctypes.byref(k))
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")
        return out.raw[:n.value].decode().split("\n")[:k.value]

    def output(self, result: int, names=None, comments=None) -> str:
        nb, no = _names(names)
        cb, co = _names(comments)
        n = ctypes.c_int64()
        L = lib()
        rc = L.orc_dag_output(ctypes.c_void_p(self.h), result, _p(nb), _p(no), _p(cb), _p(co),
                              None, ctypes.c_int64(0), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")
        buf = np.zeros(max(n.value, 1), np.uint8)
        rc = L.orc_dag_output(ctypes.c_void_p(self.h), result, _p(nb), _p(no), _p(cb), _p(co),
                              _p(buf), ctypes.c_int64(len(buf)), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")
        return buf[:n.value].tobytes().decode()

    def export(self) -> dict:
        L = lib()
        nn, ne = ctypes.c_int32(), ctypes.c_int32()
        rc = L.orc_dag_export(ctypes.c_void_p(self.h), ctypes.byref(nn), ctypes.byref(ne),
                              *([None] * 12))
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")
        N, E = nn.value, ne.value
        g = dict(base=np.zeros(N, np.uint8), in_off=np.zeros(N + 1, np.int32),
                 in_src=np.zeros(E, np.int32), in_w=np.zeros(E, np.int32),
                 out_off=np.zeros(N + 1, np.int32), out_dst=np.zeros(E, np.int32),
                 out_w=np.zeros(E, np.int32), rank_to_id=np.zeros(N, np.int32),
                 hlen=np.zeros(N, np.int32), tlen=np.zeros(N, np.int32),
                 par_id=np.zeros(N, np.int32), aligned=np.zeros(5 * N, np.int32))
        order = ["base", "in_off", "in_src", "in_w", "out_off", "out_dst", "out_w",
                 "rank_to_id", "hlen", "tlen", "par_id", "aligned"]
        rc = L.orc_dag_export(ctypes.c_void_p(self.h), ctypes.byref(nn), ctypes.byref(ne),
                              *[_p(g[k]) for k in order])
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}")
        return g


def run_loop(dag, n_reads: int, seq: np.ndarray, off: np.ndarray, result: int,
             names=None, comments=None, seeder=None, order=None, gfa: bytes | None = None):
    """minipoa's main loop (src/main.cpp:86-126) driven through a call-level object (the
    oracle's Dag or the product's api.Graph): with `gfa`, read_gfa first, whose paths are the
    first reads of the set (seq/off hold them as empty reads, names their names); then per read,
    in `order` (the guide tree with -p), alignment -> add_path -> topsort (op 1 after the last
    read), then the output, whose MSA rows map back through the order. Returns (text, [res of
    every read aligned])."""
    ord_ = list(range(n_reads)) if order is None else [int(x) for x in order]
    rid_to_ord = np.zeros(n_reads, np.int32)
    for i, r in enumerate(ord_):
        rid_to_ord[r] = i
    first = 0
    if gfa is not None:
        first = len(dag.read_gfa(gfa))
    all_res = []
    for i, rid in enumerate(ord_):
        if i < first:
            continue
        q = bytes(seq[off[rid]:off[rid + 1]])
        res = dag.alignment(rid, q) if seeder is None else dag.alignment(rid, q, seeder)
        all_res.append(res)
        dag.add_path(i, res)
        dag.topsort(1 if i + 1 == n_reads else 0)
    if seeder is None:
        return dag.output(result, names, comments), all_res
    return dag.output(result, names, comments, rid_to_ord), all_res
