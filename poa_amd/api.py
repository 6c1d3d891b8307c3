"""Python binding of the C ABI in include/mpoa.h (ctypes over poa_amd/libmpoa.so).

Mirrors minipoa's interface: ``Params`` carries minipoa's flags (src/main.cpp:14-35);
``align_reads`` is minipoa's main loop (src/main.cpp:84-126) and returns exactly the text
minipoa prints; ``Batch`` runs many independent windows in one device launch.

There is no CPU path: every call fails with ``MpoaError`` when the library or the GPU is
missing.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import PRODUCT_SO

MPOA_OK = 0
MPOA_E_OUT_CAP = -4
FORM_SORT_HBM, FORM_SORT_LANE, FORM_BT_LANE, FORM_ADD_LANE, FORM_CONS_ROUNDS = 1, 2, 4, 8, 16
FORM_WIDE = 32  # int32 storage of the DP matrices for every call (no int16 form)
FORM_HOST_SEED = 64  # -S reads seeded on the host between launches (no device seeding)
FORM_BT_RANKS = 128  # backtrack runs along consecutive ranks (not the heaviest-predecessor chain)
ERRORS = {
    -1: "E_ARG", -2: "E_BAND", -3: "E_BACKTRACK", -4: "E_OUT_CAP", -5: "E_EMPTY_WINDOW",
    -6: "E_OOM", -7: "E_DEVICE", -8: "E_OVERFLOW", -9: "E_SCRATCH",
}

EXPORTS = [
    "mpoa_create", "mpoa_destroy", "mpoa_strerror", "mpoa_align_reads", "mpoa_batch_prepare",
    "mpoa_batch_execute", "mpoa_batch_results", "mpoa_batch_free", "mpoa_batch_set_tile",
    "mpoa_batch_window_stats", "mpoa_batch_launch_stats", "mpoa_batch_set_traces",
    "mpoa_batch_traces", "mpoa_batch_set_mode", "mpoa_batch_set_graph", "mpoa_batch_set_forms",
    "mpoa_set_forms", "mpoa_last_seeding", "mpoa_align_reads_noalloc", "mpoa_graph_create", "mpoa_graph_destroy",
    "mpoa_poa", "mpoa_seeder_create", "mpoa_seeder_destroy", "mpoa_seeder_order",
    "mpoa_alignment",
    "mpoa_graph_add_path", "mpoa_graph_topsort", "mpoa_graph_output", "mpoa_graph_export",
    "mpoa_graph_read_gfa", "mpoa_reads_load", "mpoa_reads_parse", "mpoa_reads_free",
    "mpoa_reads_view", "mpoa_batch_set_predictions",
    "mpoa_batch_set_schedule", "mpoa_seed_anchors", "mpoa_seed_anchors_device", "mpoa_sched_features", "mpoa_mlp_create", "mpoa_mlp_destroy", "mpoa_mlp_n_params",
    "mpoa_mlp_get_weights", "mpoa_mlp_set_weights", "mpoa_mlp_forward", "mpoa_mlp_gradients",
    "mpoa_mlp_train", "mpoa_mlp_error",
]

MODE_BATCHED, MODE_PLAIN = 0, 1

PHASES = ("plan", "dp", "backtrack", "add_path", "topsort", "output", "topsort.order",
          "topsort.depth")


class MpoaError(RuntimeError):
    def __init__(self, rc: int, msg: str):
        super().__init__(f"mpoa error {rc} ({ERRORS.get(rc, '?')}): {msg}")
        self.rc = rc


class Params(ctypes.Structure):
    """minipoa flags; defaults are minipoa's (src/main.cpp:18-33)."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "match", "mismatch", "gap_open", "gap_ext", "band_b", "band_f", "ab_band", "seeding",
        "k", "mm_w", "poa_w", "result", "progressive")]

    def __init__(self, match=2, mismatch=-4, gap_open=-4, gap_ext=-2, band_b=100, band_f=40,
                 ab_band=0, seeding=0, k=19, mm_w=10, poa_w=0, result=0, progressive=0):
        super().__init__(match, mismatch, gap_open, gap_ext, band_b, band_f, ab_band, seeding,
                         k, mm_w, poa_w, result, progressive)

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class CallStat(ctypes.Structure):
    _fields_ = [("read", ctypes.c_int32), ("rows", ctypes.c_int32), ("qlen", ctypes.c_int32),
                ("score", ctypes.c_int32), ("cells", ctypes.c_int64),
                ("edge_cells", ctypes.c_int64), ("width", ctypes.c_int32),
                ("nodes", ctypes.c_int32), ("edges", ctypes.c_int32), ("window", ctypes.c_int32),
                ("logical_cells", ctypes.c_int64)]


class WindowStat(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("nodes", ctypes.c_int32), ("edges", ctypes.c_int32),
                ("launch", ctypes.c_int32), ("cells", ctypes.c_int64),
                ("edge_cells", ctypes.c_int64), ("ticks", ctypes.c_int64 * 8),
                ("t_begin", ctypes.c_int64), ("t_end", ctypes.c_int64),
                ("logical_cells", ctypes.c_int64), ("narrow_cells", ctypes.c_int64)]


class LaunchStat(ctypes.Structure):
    _fields_ = [("launch", ctypes.c_int32), ("threads", ctypes.c_int32), ("cols", ctypes.c_int32),
                ("slots", ctypes.c_int32), ("windows", ctypes.c_int64), ("cells", ctypes.c_int64),
                ("edge_cells", ctypes.c_int64), ("ms", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("windows", ctypes.c_int64), ("dp_calls", ctypes.c_int64),
                ("cells", ctypes.c_int64), ("edge_cells", ctypes.c_int64),
                ("reruns", ctypes.c_int64), ("device_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("launches", ctypes.c_int32),
                ("slots", ctypes.c_int32), ("pool_bytes", ctypes.c_int64),
                ("graph_launches", ctypes.c_int32), ("graph_builds", ctypes.c_int32),
                ("logical_cells", ctypes.c_int64),
                ("narrow_cells", ctypes.c_int64)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(PRODUCT_SO):
            raise MpoaError(-7, f"{PRODUCT_SO} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(PRODUCT_SO)
        P, I32, I64, VP = ctypes.POINTER, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
        L.mpoa_create.argtypes = [ctypes.c_int, P(VP)]
        L.mpoa_destroy.argtypes = [VP]
        L.mpoa_destroy.restype = None
        L.mpoa_strerror.argtypes = [ctypes.c_int]
        L.mpoa_strerror.restype = ctypes.c_char_p
        L.mpoa_align_reads.argtypes = [VP, P(Params), I32, VP, VP, VP, VP, VP, VP, VP, I64,
                                       P(I64), VP, I32, P(I32)]
        L.mpoa_batch_prepare.argtypes = [VP, P(Params), I32, VP, VP, VP, VP, P(VP)]
        L.mpoa_batch_execute.argtypes = [VP]
        L.mpoa_batch_results.argtypes = [VP, VP, I64, VP, P(BatchStats)]
        L.mpoa_batch_free.argtypes = [VP]
        L.mpoa_batch_free.restype = None
        L.mpoa_batch_set_tile.argtypes = [VP, I32, I32]
        L.mpoa_batch_window_stats.argtypes = [VP, VP, I32, P(ctypes.c_double)]
        L.mpoa_batch_launch_stats.argtypes = [VP, VP, I32, P(I32)]
        L.mpoa_batch_set_traces.argtypes = [VP, I32]
        L.mpoa_batch_set_mode.argtypes = [VP, I32]
        L.mpoa_batch_set_graph.argtypes = [VP, I32]
        L.mpoa_batch_set_forms.argtypes = [VP, I32]
        L.mpoa_set_forms.argtypes = [VP, I32]
        L.mpoa_last_seeding.argtypes = [VP]
        L.mpoa_batch_set_schedule.argtypes = [VP, I32, I32, I32, I32]
        L.mpoa_batch_set_predictions.argtypes = [VP, VP, I32]
        L.mpoa_batch_traces.argtypes = [VP, VP, I64, P(I64)]
        L.mpoa_graph_create.argtypes = [VP, P(Params), P(VP)]
        L.mpoa_graph_destroy.argtypes = [VP]
        L.mpoa_graph_destroy.restype = None
        L.mpoa_poa.argtypes = [VP, I32, I32, VP, I32, I32, VP, I64, P(I64), P(I32)]
        L.mpoa_seeder_create.argtypes = [P(Params), I32, VP, VP, P(VP)]
        L.mpoa_seeder_destroy.argtypes = [VP]
        L.mpoa_seeder_destroy.restype = None
        L.mpoa_alignment.argtypes = [VP, VP, I32, VP, I32, VP, I64, P(I64)]
        L.mpoa_graph_add_path.argtypes = [VP, I32, VP, I64]
        L.mpoa_graph_topsort.argtypes = [VP, I32]
        L.mpoa_graph_read_gfa.argtypes = [VP, VP, I64, VP, I64, P(I64), P(I32)]
        L.mpoa_reads_load.argtypes = [VP, I32, I32, P(VP)]
        L.mpoa_reads_parse.argtypes = [VP, I64, P(VP)]
        L.mpoa_reads_free.argtypes = [VP]
        L.mpoa_reads_free.restype = None
        L.mpoa_reads_view.argtypes = [VP, P(I32), P(VP), P(I32), P(VP), P(VP), P(VP), P(VP),
                                      P(VP), P(VP)]
        L.mpoa_graph_output.argtypes = [VP, I32, VP, VP, VP, VP, VP, VP, I64, P(I64)]
        L.mpoa_seeder_order.argtypes = [VP, VP, I32]
        L.mpoa_seeder_order.restype = I32
        L.mpoa_graph_export.argtypes = [VP, P(I32), P(I32)] + [VP] * 12
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != MPOA_OK:
        raise MpoaError(rc, lib().mpoa_strerror(rc).decode())


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _names(names: list[str] | None):
    if not names:
        return None, None
    enc = [n.encode() for n in names]
    off = np.zeros(len(enc) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in enc])
    buf = np.frombuffer(b"".join(enc) + b"\0", np.uint8).copy()
    return buf, off


class Context:
    """Device context (mpoa_create). One per process and device is enough."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().mpoa_create(device, ctypes.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().mpoa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_forms(self, forms: int) -> None:
        """Test hook: the FORM_* fallback forms of every batch created afterwards, including
        the one behind align_reads; 0 restores the normal choice. Results do not change."""
        _check(lib().mpoa_set_forms(self.h, int(forms)))

    def last_seeding(self) -> int:
        """How the last seeded align_reads call was seeded: 1 on the device, 0 on the host by
        choice, above 1 on the host after the device form stopped (bits naming why, as
        mpoa_last_seeding documents), -1 before any seeded call."""
        return int(lib().mpoa_last_seeding(self.h))

    def align_reads(self, params: Params, seq: np.ndarray, off: np.ndarray,
                    names: list[str] | None = None, comments: list[str] | None = None,
                    with_stats: bool = False):
        """minipoa's main loop on the GPU. Returns the stdout text (and DP call stats)."""
        n = len(off) - 1
        seq = np.ascontiguousarray(seq, np.uint8)
        off = np.ascontiguousarray(off, np.int64)
        nb, no = _names(names)
        cb, co = _names(comments)
        need = ctypes.c_int64()
        n_stats = ctypes.c_int32()
        rc = lib().mpoa_align_reads(self.h, ctypes.byref(params), n, _ptr(seq), _ptr(off),
                                    _ptr(nb), _ptr(no), _ptr(cb), _ptr(co), None, 0,
                                    ctypes.byref(need), None, 0, ctypes.byref(n_stats))
        _check(rc)
        buf = np.zeros(max(need.value, 1), np.uint8)
        cap = max(n_stats.value, 1)
        stats = (CallStat * cap)()
        rc = lib().mpoa_align_reads(self.h, ctypes.byref(params), n, _ptr(seq), _ptr(off),
                                    _ptr(nb), _ptr(no), _ptr(cb), _ptr(co), _ptr(buf),
                                    len(buf), ctypes.byref(need), stats, cap,
                                    ctypes.byref(n_stats))
        _check(rc)
        text = buf[:need.value].tobytes().decode()
        if with_stats:
            return text, [stats[i] for i in range(min(n_stats.value, cap))]
        return text


@dataclass
class BatchResult:
    texts: list[str]
    stats: BatchStats
    device_ms: float = field(default=0.0)


class Batch:
    """Independent windows prepared on the device (mpoa_batch_prepare)."""

    def __init__(self, ctx: Context, params: Params, win_off: np.ndarray, seq: np.ndarray,
                 seq_off: np.ndarray, size_hint: np.ndarray | None = None, tuned: bool = True):
        self.ctx = ctx
        self.win_off = np.ascontiguousarray(win_off, np.int32)
        seq = np.ascontiguousarray(seq, np.uint8)
        seq_off = np.ascontiguousarray(seq_off, np.int64)
        hint = None if size_hint is None else np.ascontiguousarray(size_hint, np.float64)
        h = ctypes.c_void_p()
        _check(lib().mpoa_batch_prepare(ctx.h, ctypes.byref(params), len(self.win_off) - 1,
                                        _ptr(self.win_off), _ptr(seq), _ptr(seq_off), _ptr(hint),
                                        ctypes.byref(h)))
        self.h = h
        self.n = len(self.win_off) - 1
        self.tuning = None
        if tuned:
            # the autotuner's choice for this device (poa_amd/autotune.py), when present
            from . import autotune
            t = autotune.load()
            if t:
                self.set_tile(*t["small_tile"])
                self.set_schedule(move_top=t["move_top"])
                self.tuning = {"small_tile": t["small_tile"], "move_top": t["move_top"]}

    def set_tile(self, threads: int, cols: int) -> None:
        _check(lib().mpoa_batch_set_tile(self.h, threads, cols))

    def execute(self) -> float:
        """Runs every window on the device; returns device milliseconds."""
        _check(lib().mpoa_batch_execute(self.h))
        return self.stats().device_ms

    def window_stats(self):
        """(records, seconds per tick) of the last execute."""
        arr = (WindowStat * max(self.n, 1))()
        hz = ctypes.c_double()
        _check(lib().mpoa_batch_window_stats(self.h, arr, self.n, ctypes.byref(hz)))
        return [arr[i] for i in range(self.n)], 1.0 / hz.value

    def set_schedule(self, small_qmax: int = 1500, small_width: int = 344,
                     move_top: int = -1, small_nodes: int = 1 << 30) -> None:
        """Launch split between the small tile and the large launch."""
        _check(lib().mpoa_batch_set_schedule(self.h, small_qmax, small_width, move_top,
                                             small_nodes))

    def set_mode(self, mode: int) -> None:
        """MODE_BATCHED (persistent launches, work queue) or MODE_PLAIN (one launch per
        window, in input order)."""
        _check(lib().mpoa_batch_set_mode(self.h, mode))

    def set_graph(self, enable: bool) -> None:
        """Replay the batched pass as a captured hipGraph (default) or enqueue its launches
        one by one."""
        _check(lib().mpoa_batch_set_graph(self.h, int(bool(enable))))

    def set_forms(self, forms: int) -> None:
        """Test hook: run every read through the fallback forms given by FORM_* bits (the sort
        over a CSR in HBM, the one-lane sort, backtrack or add_path); 0 restores the normal
        choice. Results do not change."""
        _check(lib().mpoa_batch_set_forms(self.h, int(forms)))

    def set_predictions(self, pred: np.ndarray) -> None:
        """Per-window (cost, max call cells, max band width, final nodes) predictions that
        order the queue and size the arenas; results never depend on them."""
        pred = np.ascontiguousarray(pred, np.float32)
        assert pred.shape == (self.n, 4)
        _check(lib().mpoa_batch_set_predictions(self.h, _ptr(pred), 4))

    def set_traces(self, enable: bool = True) -> None:
        """Record every DP call's geometry on the next execute (see traces())."""
        _check(lib().mpoa_batch_set_traces(self.h, 1 if enable else 0))

    def traces(self) -> np.ndarray:
        """Per-DP-call records of the last execute as a structured array (window order)."""
        n = ctypes.c_int64()
        _check(lib().mpoa_batch_traces(self.h, None, 0, ctypes.byref(n)))
        arr = (CallStat * max(n.value, 1))()
        _check(lib().mpoa_batch_traces(self.h, arr, n.value, ctypes.byref(n)))
        return np.ctypeslib.as_array(arr)[:n.value].copy()

    def launch_stats(self) -> list[LaunchStat]:
        """Per-launch records (tile, windows, cells, HIP-event milliseconds) of the last execute."""
        arr = (LaunchStat * 8)()
        n = ctypes.c_int32()
        _check(lib().mpoa_batch_launch_stats(self.h, arr, 8, ctypes.byref(n)))
        return [arr[i] for i in range(min(n.value, 8))]

    def stats(self) -> BatchStats:
        st = BatchStats()
        _check(lib().mpoa_batch_results(self.h, None, 0, None, ctypes.byref(st)))
        return st

    def results(self) -> BatchResult:
        off = np.zeros(self.n + 1, np.int64)
        st = BatchStats()
        _check(lib().mpoa_batch_results(self.h, None, 0, _ptr(off), ctypes.byref(st)))
        buf = np.zeros(max(int(off[-1]), 1), np.uint8)
        _check(lib().mpoa_batch_results(self.h, _ptr(buf), len(buf), _ptr(off),
                                        ctypes.byref(st)))
        raw = buf.tobytes()
        texts = [raw[off[i]:off[i + 1]].decode() for i in range(self.n)]
        return BatchResult(texts, st, st.device_ms)

    def free(self) -> None:
        if self.h:
            lib().mpoa_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# Call-level API: minipoa's poa() / alignment() / graph::add_path / graph::topsort / outputs
# on a graph kept in device memory (include/mpoa.h, "Call-level API").
# ---------------------------------------------------------------------------------------
RES_DTYPE = np.dtype([("from", np.int32), ("base", np.int32), ("aligned_id", np.int32)])


class ReadSet:
    """minipoa's readFile for one or many FASTA/FASTQ(.gz) files (one window each), parsed by
    the library on host threads: packed arrays ready for Context.align_reads and Batch."""

    def __init__(self, paths: list[str] | None = None, threads: int = 0,
                 text: bytes | None = None):
        import time
        h = ctypes.c_void_p()
        if text is not None:
            buf = ctypes.create_string_buffer(text, len(text) + 1)
            t0 = time.perf_counter()
            _check(lib().mpoa_reads_parse(buf, len(text), ctypes.byref(h)))
        else:
            enc = [os.fsencode(p) for p in paths]
            arr = (ctypes.c_char_p * max(len(enc), 1))(*enc)
            t0 = time.perf_counter()
            _check(lib().mpoa_reads_load(arr, len(enc), threads, ctypes.byref(h)))
        self.load_s = time.perf_counter() - t0  # the library call alone
        try:
            nw, nr = ctypes.c_int32(), ctypes.c_int32()
            ptr = [ctypes.c_void_p() for _ in range(7)]
            _check(lib().mpoa_reads_view(h, ctypes.byref(nw), ctypes.byref(ptr[0]),
                                         ctypes.byref(nr), *[ctypes.byref(x) for x in ptr[1:]]))

            def arr_of(p, n, dt):
                if n == 0 or not p.value:
                    return np.zeros(n, dt)
                return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(
                    np.ctypeslib.as_ctypes_type(dt))), (n,)).copy()

            self.win_off = arr_of(ptr[0], nw.value + 1, np.int32)
            self.seq_off = arr_of(ptr[2], nr.value + 1, np.int64)
            self.name_off = arr_of(ptr[4], nr.value + 1, np.int64)
            self.comment_off = arr_of(ptr[6], nr.value + 1, np.int64)
            self.seq = arr_of(ptr[1], int(self.seq_off[-1]), np.uint8)
            self.name = arr_of(ptr[3], int(self.name_off[-1]), np.uint8).tobytes()
            self.comment = arr_of(ptr[5], int(self.comment_off[-1]), np.uint8).tobytes()
        finally:
            lib().mpoa_reads_free(h)

    @staticmethod
    def _split(text: bytes, off: np.ndarray) -> list[str]:
        return [text[off[i]:off[i + 1]].decode("latin-1") for i in range(len(off) - 1)]

    @property
    def names(self) -> list[str]:
        return self._split(self.name, self.name_off)

    @property
    def comments(self) -> list[str]:
        return self._split(self.comment, self.comment_off)

    @property
    def n_windows(self) -> int:
        return len(self.win_off) - 1

    @property
    def n_reads(self) -> int:
        return len(self.seq_off) - 1


class Seeder:
    """minipoa's minimizer_t over a read set (needed by Graph.alignment with -S)."""

    def __init__(self, params: Params, seq: np.ndarray, off: np.ndarray):
        seq = np.ascontiguousarray(seq, np.uint8)
        off = np.ascontiguousarray(off, np.int64)
        h = ctypes.c_void_p()
        _check(lib().mpoa_seeder_create(ctypes.byref(params), len(off) - 1, _ptr(seq), _ptr(off),
                                        ctypes.byref(h)))
        self.h = h

    def order(self) -> np.ndarray:
        """The order reads are added in: minipoa's guide tree with -p, else input order."""
        n = lib().mpoa_seeder_order(self.h, None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().mpoa_seeder_order(self.h, _ptr(out), n)
        return out[:n]

    def close(self) -> None:
        if self.h:
            lib().mpoa_seeder_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A POA graph in device memory driven call by call, as minipoa's main loop drives its
    graph (src/main.cpp:110-126). res arrays are structured arrays of RES_DTYPE."""

    def __init__(self, ctx: Context, params: Params):
        self.ctx = ctx
        self.params = params
        h = ctypes.c_void_p()
        _check(lib().mpoa_graph_create(ctx.h, ctypes.byref(params), ctypes.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().mpoa_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _res_call(fn, *args, guess: int = 64) -> np.ndarray:
        """One call with room for `guess` entries; a second, sized one only on MPOA_E_OUT_CAP."""
        n = ctypes.c_int64()
        out = np.zeros(max(guess, 1), RES_DTYPE)
        rc = fn(*args, _ptr(out), len(out), ctypes.byref(n))
        if rc == MPOA_E_OUT_CAP:
            out = np.zeros(max(n.value, 1), RES_DTYPE)
            rc = fn(*args, _ptr(out), len(out), ctypes.byref(n))
        _check(rc)
        return out[:n.value]

    def poa(self, beg_id: int, end_id: int, q: bytes, ab_band: bool):
        """poa() on the rank window [rank(beg_id), rank(end_id)]: (res, score)."""
        score = ctypes.c_int32()
        n = ctypes.c_int64()
        buf = ctypes.create_string_buffer(q, len(q) + 1)
        out = np.zeros(len(q) + 2, RES_DTYPE)
        _check(lib().mpoa_poa(self.h, beg_id, end_id, buf, len(q), int(ab_band), _ptr(out),
                              len(out), ctypes.byref(n), ctypes.byref(score)))
        return out[:n.value], score.value

    def alignment(self, rid: int, q: bytes, seeder: Seeder | None = None) -> np.ndarray:
        """alignment(): the res of read rid against the graph."""
        buf = ctypes.create_string_buffer(q, len(q) + 1)
        return self._res_call(lib().mpoa_alignment, self.h, seeder.h if seeder else None, rid,
                              buf, len(q), guess=2 * len(q) + 8)

    def add_path(self, seq_id: int, res: np.ndarray) -> None:
        """graph::add_path of a res given as RES_DTYPE records or as (n, 3) int32 rows."""
        if res.dtype != RES_DTYPE:
            res = np.ascontiguousarray(res, np.int32).reshape(-1, 3).view(RES_DTYPE).reshape(-1)
        res = np.ascontiguousarray(res, RES_DTYPE)
        _check(lib().mpoa_graph_add_path(self.h, seq_id, _ptr(res), len(res)))

    def topsort(self, op: int) -> None:
        _check(lib().mpoa_graph_topsort(self.h, op))

    def read_gfa(self, text: bytes) -> list[str]:
        """read_gfa (minipoa -i): the GFA's segments and '+' paths into this new graph, the
        paths as reads 0, 1, ...; returns the path names (a name and its newline take no more
        room than its P line)."""
        buf = ctypes.create_string_buffer(text, len(text) + 1)
        out = ctypes.create_string_buffer(len(text) + 1)
        n, k = ctypes.c_int64(), ctypes.c_int32()
        _check(lib().mpoa_graph_read_gfa(self.h, buf, len(text), out, len(out), ctypes.byref(n),
                                         ctypes.byref(k)))
        return out.raw[:n.value].decode().split("\n")[:k.value]

    def output(self, result: int, names: list[str] | None = None,
               comments: list[str] | None = None, rid_to_ord: np.ndarray | None = None) -> str:
        nb, no = _names(names)
        cb, co = _names(comments)
        ro = None if rid_to_ord is None else np.ascontiguousarray(rid_to_ord, np.int32)
        n = ctypes.c_int64()
        _check(lib().mpoa_graph_output(self.h, result, _ptr(nb), _ptr(no), _ptr(cb), _ptr(co),
                                       _ptr(ro), None, 0, ctypes.byref(n)))
        buf = np.zeros(max(n.value, 1), np.uint8)
        _check(lib().mpoa_graph_output(self.h, result, _ptr(nb), _ptr(no), _ptr(cb), _ptr(co),
                                       _ptr(ro), _ptr(buf), len(buf), ctypes.byref(n)))
        return buf[:n.value].tobytes().decode()

    def export(self) -> dict:
        """Node bases, par_id, aligned_node, in/out edge CSR with weights, rank -> id, hlen,
        tlen (by rank of the last op 0 sort)."""
        nn, ne = ctypes.c_int32(), ctypes.c_int32()
        L = lib()
        _check(L.mpoa_graph_export(self.h, ctypes.byref(nn), ctypes.byref(ne), *([None] * 12)))
        N, E = nn.value, ne.value
        d = {"base": np.zeros(N, np.uint8), "in_off": np.zeros(N + 1, np.int32),
             "in_src": np.zeros(max(E, 1), np.int32), "in_w": np.zeros(max(E, 1), np.int32),
             "out_off": np.zeros(N + 1, np.int32), "out_dst": np.zeros(max(E, 1), np.int32),
             "out_w": np.zeros(max(E, 1), np.int32), "rank_to_id": np.zeros(N, np.int32),
             "hlen": np.zeros(N, np.int32), "tlen": np.zeros(N, np.int32),
             "par_id": np.zeros(N, np.int32), "aligned": np.zeros(5 * N, np.int32)}
        keys = ("base", "in_off", "in_src", "in_w", "out_off", "out_dst", "out_w", "rank_to_id",
                "hlen", "tlen", "par_id", "aligned")
        _check(L.mpoa_graph_export(self.h, ctypes.byref(nn), ctypes.byref(ne),
                                   *[_ptr(d[k]) for k in keys]))
        d["in_src"], d["in_w"] = d["in_src"][:E], d["in_w"][:E]
        d["out_dst"], d["out_w"] = d["out_dst"][:E], d["out_w"][:E]
        return d

