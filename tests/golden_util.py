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
    "-W": ("poa_w", True), "-r": ("result", True), "-p": ("progressive", False),
}


def load_cases() -> list[dict]:
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def case_id(c: dict) -> str:
    return f"{c['case']}[{' '.join(c['args'])}]"


def parse_args(args: list[str]) -> dict:
    """minipoa flags to parameter fields. Scores are taken as typed: initPara negates positive
    penalties itself (src/handle_input.cpp:45-77), and so do both libraries. -t (threads) is
    accepted and dropped: it changes no output (src/align.cpp:1796-1823 gives the same
    segments), and the device path has no host threads to set."""
    p: dict = {}
    i = 0
    while i < len(args):
        if args[i] == "-t":
            i += 2
            continue
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


def check_msa_consensus_row(text3: str, text_r1: str, text_r0: str) -> None:
    """The -r 3 output against upstream's -r 1 and -r 0 texts of the same run: the MSA rows are
    exactly -r 1's; one more row follows, ">consensus sequence", whose bases without gaps are
    -r 0's consensus and each of whose bases stands in a column where some read has the same
    base (a consensus node lies on a read's path, and both sit in the node's column)."""
    assert text3.startswith(text_r1)
    tail = text3[len(text_r1):].split("\n")
    assert len(tail) == 3 and tail[0] == ">consensus sequence" and tail[2] == ""
    row = tail[1]
    r0 = text_r0.split("\n")
    assert r0[0] == ">consensus sequence"
    assert row.replace("-", "") == r0[1]
    rows = [ln for ln in text_r1.splitlines() if not ln.startswith(">")]
    assert rows and all(len(r) == len(row) for r in rows)
    for c, ch in enumerate(row):
        if ch != "-":
            assert any(r[c] == ch for r in rows), f"column {c}"



def gfa_read_set(gfa_text: str, fasta_text: str):
    """minipoa -i's read set (src/main.cpp:86-89): the GFA's P lines first, as reads with a name,
    no comment and no bases, then the FASTA's reads. Returns (seq, off, names, comments,
    number of paths)."""
    paths = [ln.split("\t")[1] for ln in gfa_text.splitlines() if ln.startswith("P\t")]
    seq, off, names, comments = parse_fasta(fasta_text)
    off = np.concatenate([np.zeros(len(paths), np.int64), off])
    return seq, off, paths + names, [""] * len(paths) + comments, len(paths)
