"""Seeded synthetic workloads for the five measured configurations.

The generator itself is C++ (``csrc/synth.cpp``, exported through ``libmpoa_synth.so``) so the
same seed gives the same reads on every machine. Parameters that the workload descriptions
leave open are fixed here and recorded:

* config 3/5: lognormal sigma = 0.5, error split sub:ins:del = 1:1:1, indel length 1;
* config 4: indel length geometric with P(L=1) = 0.7, as config 2.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


class BatchSpec(ctypes.Structure):
    _fields_ = [
        ("min_reads", ctypes.c_int32), ("max_reads", ctypes.c_int32),
        ("median_len", ctypes.c_double), ("sigma", ctypes.c_double),
        ("min_len", ctypes.c_int32), ("max_len", ctypes.c_int32),
        ("err_lo", ctypes.c_double), ("err_hi", ctypes.c_double),
        ("split_sub", ctypes.c_double), ("split_ins", ctypes.c_double),
        ("split_del", ctypes.c_double), ("split_sum", ctypes.c_double),
        ("geo_p", ctypes.c_double),
    ]


@dataclass
class Reads:
    """Concatenated reads: ``seq`` (uint8 ASCII) and ``off`` (n+1 int64 offsets)."""
    seq: np.ndarray
    off: np.ndarray

    @property
    def n(self) -> int:
        return len(self.off) - 1

    def read(self, i: int) -> bytes:
        return self.seq[self.off[i]:self.off[i + 1]].tobytes()

    def to_fasta(self, path: str, names: list[str] | None = None) -> None:
        with open(path, "w") as f:
            for i in range(self.n):
                name = names[i] if names else f"read{i + 1}"
                f.write(f">{name}\n{self.read(i).decode()}\n")


@dataclass
class Batch:
    """Independent windows: window w owns reads [win_off[w], win_off[w+1])."""
    seq: np.ndarray
    seq_off: np.ndarray
    win_off: np.ndarray

    @property
    def n_windows(self) -> int:
        return len(self.win_off) - 1

    def window(self, w: int) -> Reads:
        r0, r1 = int(self.win_off[w]), int(self.win_off[w + 1])
        b0, b1 = int(self.seq_off[r0]), int(self.seq_off[r1])
        return Reads(self.seq[b0:b1].copy(), self.seq_off[r0:r1 + 1] - b0)

    def subset(self, windows) -> "Batch":
        parts = [self.window(int(w)) for w in windows]
        seq = np.concatenate([p.seq for p in parts]) if parts else np.zeros(0, np.uint8)
        win_off = np.zeros(len(parts) + 1, np.int32)
        offs, base, nr = [np.zeros(1, np.int64)], 0, 0
        for i, p in enumerate(parts):
            offs.append(p.off[1:] + base)
            base += int(p.off[-1])
            nr += p.n
            win_off[i + 1] = nr
        return Batch(seq, np.concatenate(offs), win_off)


def window(seed: int, tlen: int, n_reads: int, sub: float, ins: float, dele: float,
           geo_p: float = 0.0) -> Reads:
    lib = _lib.synth()
    off = np.zeros(n_reads + 1, np.int64)
    rc = lib.mpoa_synth_window(seed, tlen, n_reads, sub, ins, dele, geo_p, None, 0,
                               off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc != 0:
        raise RuntimeError(f"mpoa_synth_window failed ({rc})")
    seq = np.zeros(max(int(off[-1]), 1), np.uint8)
    rc = lib.mpoa_synth_window(seed, tlen, n_reads, sub, ins, dele, geo_p,
                               seq.ctypes.data_as(ctypes.c_char_p), len(seq),
                               off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc != 0:
        raise RuntimeError(f"mpoa_synth_window failed ({rc})")
    return Reads(seq[:int(off[-1])], off)


def batch(seed: int, n_windows: int, *, min_reads=5, max_reads=60, median_len=500.0, sigma=0.5,
          min_len=150, max_len=2000, err_lo=0.05, err_hi=0.15, split=(1.0, 1.0, 1.0),
          geo_p=0.0) -> Batch:
    lib = _lib.synth()
    spec = BatchSpec(min_reads, max_reads, median_len, sigma, min_len, max_len, err_lo, err_hi,
                     split[0], split[1], split[2], sum(split), geo_p)
    nr, nb = ctypes.c_int64(), ctypes.c_int64()
    rc = lib.mpoa_synth_batch(seed, ctypes.byref(spec), n_windows, None, None, 0, None,
                              ctypes.byref(nr), ctypes.byref(nb))
    if rc != 0:
        raise RuntimeError(f"mpoa_synth_batch failed ({rc})")
    win_off = np.zeros(n_windows + 1, np.int32)
    seq = np.zeros(max(nb.value, 1), np.uint8)
    seq_off = np.zeros(nr.value + 1, np.int64)
    rc = lib.mpoa_synth_batch(seed, ctypes.byref(spec), n_windows,
                              win_off.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                              seq.ctypes.data_as(ctypes.c_char_p), len(seq),
                              seq_off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                              ctypes.byref(nr), ctypes.byref(nb))
    if rc != 0:
        raise RuntimeError(f"mpoa_synth_batch failed ({rc})")
    return Batch(seq[:nb.value], seq_off, win_off)


# Seeds of the five measured workloads.
SEEDS = {1: 0x5151, 2: 0x5252, 3: 0x5353, 4: 0x5454, 5: 0x5555, "train": 0x7a7a}


def config1(seed: int = SEEDS[1], tlen: int = 16569, n_reads: int = 20) -> Reads:
    """Consensus, plain method: 16,569 bp template, 20 reads, 1% sub / 0.5% ins / 0.5% del."""
    return window(seed, tlen, n_reads, 0.01, 0.005, 0.005)


def config2(seed: int = SEEDS[2], tlen: int = 2000, n_reads: int = 100) -> Reads:
    """MSA: 2,000 bp template, 100 reads, 10% error split 4:3:3, geometric(0.7) indels."""
    return window(seed, tlen, n_reads, 0.04, 0.03, 0.03, 0.7)


def config3(seed: int = SEEDS[3], n_windows: int = 10000) -> Batch:
    """Batched uneven windows: 5-60 reads, lognormal(ln 500, 0.5) length in [150, 2000],
    per-window error U(5%, 15%) split 1:1:1."""
    return batch(seed, n_windows)


def config4(seed: int = SEEDS[4], tlen: int = 20000, n_reads: int = 50,
            geo_p: float = 0.35) -> Reads:
    """Long-read graph growth: 20 kb template, 50 reads, 12% error split 2:4:4, indel lengths
    geometric(0.35) (mean 2.9 bases): with -S -k 15 -w 10 the graph grows to ~103,000 nodes
    (geometric(0.7) indels stop at ~75,000)."""
    return window(seed, tlen, n_reads, 0.024, 0.048, 0.048, geo_p)


def config5(seed: int = SEEDS[5], n_windows: int = 100000) -> Batch:
    """Learned scheduling: 100,000 windows drawn as in config 3."""
    return batch(seed, n_windows)
