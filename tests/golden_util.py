"""Helpers for the golden fixtures (tests/golden/golden.json): FASTA text to read arrays, and
minipoa's command-line flags (src/main.cpp:14-35) to a parameter dict."""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "golden.json")

# flag -> (field, takes a value)
_FLAGS = {
    "-M": ("match", True), "-X": ("mismatch", True), "-O": ("gap_open", True),
    "-E": ("gap_ext", True), "-b": ("band_b", True), "-f": ("band_f", True),
    "-B": ("ab_band", False), "-S": ("seeding", False), "-k": ("k", True), "-w": ("mm_w", True),
    "-W": ("poa_w", True), "-r": ("result", True),
}


def load_cases() -> list[dict]:
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def case_id(c: dict) -> str:
    return f"{c['case']}[{' '.join(c['args'])}]"


def parse_args(args: list[str]) -> dict:
    """minipoa flags to parameter fields. Scores are taken as typed: initPara negates positive
    penalties itself (src/handle_input.cpp:45-77), and so do both libraries."""
    p: dict = {}
    i = 0
    while i < len(args):
        field, val = _FLAGS[args[i]]
        if val:
            p[field] = int(args[i + 1])
            i += 2
        else:
            p[field] = 1
            i += 1
    return p


def parse_fasta(text: str):
    """(seq uint8 concatenation, offsets int64, names, comments) of a FASTA text."""
    names, comments, seqs = [], [], []
    cur = None
    for line in text.splitlines():
        if line.startswith(">"):
            head = line[1:]
            sp = head.find(" ")
            names.append(head if sp < 0 else head[:sp])
            comments.append("" if sp < 0 else head[sp + 1:])
            cur = []
            seqs.append(cur)
        elif line:
            cur.append(line.strip())
    joined = ["".join(s) for s in seqs]
    off = np.zeros(len(joined) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in joined])
    seq = np.frombuffer("".join(joined).encode(), np.uint8).copy()
    return seq, off, names, comments
