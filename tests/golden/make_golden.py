#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/golden.json.

Each case is a small read set (seeded synthetic reads, the first reads of the reference's own
data/mtDNA.fasta, or hand-written edge cases) and a minipoa flag set. The expected output is
what the upstream minipoa binary prints: stdout (consensus / row-column MSA / GFA) and the
"score:" line of every poa() call under -V 1 (src/align.cpp:1577).

The upstream binary is oracle/_ref/minipoa, built from the upstream sources by
`make -C oracle ref`. It runs only here, when the fixtures are (re)made; the fixtures hold
inputs and outputs only.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from poa_amd import synth  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "minipoa")
MTDNA = os.environ.get("MPOA_MTDNA", "/root/reference/data/mtDNA.fasta")


def fasta_of(reads: list[tuple[str, str, str]]) -> str:
    out = []
    for name, comment, seq in reads:
        out.append(f">{name}{(' ' + comment) if comment else ''}\n{seq}\n")
    return "".join(out)


def synth_reads(r: synth.Reads, prefix: str) -> list[tuple[str, str, str]]:
    return [(f"{prefix}{i + 1}", "", r.read(i).decode()) for i in range(r.n)]


def mtdna_reads(n: int) -> list[tuple[str, str, str]]:
    reads, name, comment, seq = [], None, "", []
    with open(MTDNA) as f:
        for line in f:
            line = line.rstrip("\n")
            if line.startswith(">"):
                if name is not None:
                    reads.append((name, comment, "".join(seq)))
                head = line[1:]
                name = head.split()[0] if head.split() else ""
                sp = head.find(" ")
                comment = head[sp + 1:] if sp >= 0 else ""
                seq = []
            elif line:
                seq.append(line.strip())
        if name is not None:
            reads.append((name, comment, "".join(seq)))
    return reads[:n]


def cases():
    c = []
    mt = mtdna_reads(40)
    for args in (["-r", "0"], ["-r", "1"], ["-r", "2"], ["-B", "-r", "0"], ["-f", "0", "-r", "0"]):
        c.append(("mtdna40", mt, args))
    c1 = synth_reads(synth.window(101, 1500, 5, 0.01, 0.005, 0.005), "c1_")
    for args in (["-r", "0"], ["-r", "2"]):
        c.append(("cfg1_small", c1, args))
    c2 = synth_reads(synth.window(102, 600, 20, 0.04, 0.03, 0.03, 0.7), "c2_")
    for args in (["-r", "0"], ["-r", "1"], ["-r", "2"], ["-B", "-r", "1"],
                 ["-M", "3", "-X", "5", "-O", "6", "-E", "1", "-b", "20", "-f", "10", "-r", "1"]):
        c.append(("cfg2_small", c2, args))
    c2s = synth_reads(synth.window(103, 2000, 12, 0.04, 0.03, 0.03, 0.7), "s_")
    for args in (["-S", "-r", "1"], ["-S", "-k", "15", "-w", "10", "-r", "0"]):
        c.append(("cfg2_seeded", c2s, args))
    b = synth.config3(seed=104, n_windows=6)
    for w in range(b.n_windows):
        c.append((f"cfg3_w{w}", synth_reads(b.window(w), f"w{w}_"), ["-r", "0"]))
    c4 = synth_reads(synth.window(105, 3000, 5, 0.024, 0.048, 0.048, 0.7), "c4_")
    for args in (["-S", "-k", "15", "-w", "10", "-r", "0"], ["-S", "-k", "15", "-w", "10", "-r", "2"]):
        c.append(("cfg4_small", c4, args))
    edges = {
        "one_read": ["ACGTACGTAC"],
        "two_reads": ["ACGTACGTAC", "ACGTTACGTAC"],
        "identical": ["GATTACAGATTACA"] * 4,
        "with_n": ["ACGTNNACGT", "ACGTACGT", "ACNTACGTA"],
        "short": ["AC", "A", "ACG", "C"],
        "lowercase": ["acgtacgt", "ACGTACGT", "acgTTacgt"],
    }
    for name, rs in edges.items():
        reads = [(f"e{i + 1}", "x y" if i == 0 else "", s) for i, s in enumerate(rs)]
        for args in (["-r", "0"], ["-r", "1"], ["-r", "2"]):
            c.append((f"edge_{name}", reads, args))
    # progressive order (-p, minimizer.cpp:195-386); its GFA output reads past minipoa's own
    # path arrays when read lengths differ (graph.cpp:369,423), so -r 2 is left out
    for args in (["-p", "-r", "0"], ["-p", "-r", "1"], ["-p", "-B", "-r", "1"]):
        c.append(("mtdna40", mt, args))
    for args in (["-p", "-r", "0"], ["-p", "-r", "1"]):
        c.append(("cfg2_small", c2, args))
    c.append(("cfg2_seeded", c2s, ["-p", "-S", "-r", "1"]))
    for w in range(2):
        c.append((f"cfg3_w{w}", synth_reads(b.window(w), f"w{w}_"), ["-p", "-r", "0"]))
    for name in ("two_reads", "short", "identical"):
        reads = [(f"e{i + 1}", "x y" if i == 0 else "", s) for i, s in enumerate(edges[name])]
        c.append((f"edge_{name}", reads, ["-p", "-r", "1"]))
    return c


def run_ref(fasta: str, args: list[str]) -> tuple[str, list[int], int]:
    with tempfile.NamedTemporaryFile("w", suffix=".fa", delete=False) as f:
        f.write(fasta)
        path = f.name
    try:
        p = subprocess.run([REF_BIN, path, *args, "-V", "1"], capture_output=True, text=True)
    finally:
        os.unlink(path)
    scores = [int(l.split(":", 1)[1]) for l in p.stderr.splitlines() if l.startswith("score:")]
    return p.stdout, scores, p.returncode


def main() -> None:
    if not os.path.exists(REF_BIN):
        sys.exit(f"{REF_BIN} missing: run make -C oracle ref")
    out = []
    for name, reads, args in cases():
        fasta = fasta_of(reads)
        stdout, scores, rc = run_ref(fasta, args)
        out.append({"case": name, "args": args, "fasta": fasta, "stdout": stdout,
                    "scores": scores, "returncode": rc})
        print(f"{name:14s} {' '.join(args):40s} rc={rc} calls={len(scores)} "
              f"stdout={len(stdout)}B", flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "upstream": "minipoa (NCl3-lhd/POA)",
                   "cases": out}, f, indent=0)


if __name__ == "__main__":
    main()
