#!/usr/bin/env python3
"""Writes tests/golden/full_configs.json.gz: the full-size single-window workloads (the
16,569 bp x 20 read consensus run, the 2,000 bp x 100 read MSA run with and without -S, and the
20 kb x 50 read seeded run), each generated by poa_amd.synth from its seed and run through the
upstream minipoa binary (oracle/_ref/minipoa, built by `make -C oracle ref`). The fixture
stores the generator call, a SHA-256 of the generated FASTA (so a changed generator is caught)
and the upstream outputs: stdout and every DP call's score (-V 1). Also writes
tests/golden/mtdna_full.json.gz: the reference's data/mtDNA.fasta (393 reads) with the
upstream outputs of the README's commands and of -r 0/1/2 and -B."""
from __future__ import annotations

import gzip
import hashlib
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

RUNS = [
    ("config1", "config1", ["-r", "0"]),
    ("config2_S", "config2", ["-S", "-r", "1"]),
    ("config2", "config2", ["-r", "1"]),
    ("config4_S", "config4", ["-S", "-k", "15", "-w", "10", "-r", "0"]),
    # the consensus of the two config-2 runs: the ungapped form of the MSA consensus row
    ("config2_r0", "config2", ["-r", "0"]),
    ("config2_S_r0", "config2", ["-S", "-r", "0"]),
]
# the reference's own data file, whole (393 reads), with the README's commands
MTDNA = os.environ.get("MPOA_MTDNA", "/root/reference/data/mtDNA.fasta")
MTDNA_RUNS = [
    ("mtdna", ["-r", "0"]),
    ("mtdna", ["-r", "1"]),
    ("mtdna", ["-r", "2"]),
    ("mtdna", ["-B", "-r", "0"]),
    ("mtdna", ["-p", "-r", "1"]),
    # README.md's MSA command: upstream dies on a sub-window that ends before it begins
    ("mtdna", ["-S", "-t", "12", "-r", "1"]),
]


def fasta_text(reads: synth.Reads) -> str:
    with tempfile.NamedTemporaryFile("r", suffix=".fa", delete=False) as f:
        path = f.name
    try:
        reads.to_fasta(path)
        with open(path) as f:
            return f.read()
    finally:
        os.unlink(path)


def main() -> None:
    out = []
    for name, gen, args in RUNS:
        reads = getattr(synth, gen)()
        fa = fasta_text(reads)
        with tempfile.NamedTemporaryFile("w", suffix=".fa", delete=False) as f:
            f.write(fa)
            path = f.name
        try:
            p = subprocess.run([REF_BIN, path, *args, "-V", "1"], capture_output=True, text=True)
        finally:
            os.unlink(path)
        scores = [int(l.split(":", 1)[1]) for l in p.stderr.splitlines() if l.startswith("score:")]
        out.append({"case": name, "generator": f"synth.{gen}()", "args": args,
                    "fasta_sha256": hashlib.sha256(fa.encode()).hexdigest(),
                    "stdout": p.stdout, "scores": scores, "returncode": p.returncode})
        print(f"{name:10s} {' '.join(args):30s} rc={p.returncode} calls={len(scores)} "
              f"stdout={len(p.stdout)}B", flush=True)
    with gzip.open(os.path.join(HERE, "full_configs.json.gz"), "wt") as f:
        json.dump({"generator": "tests/golden/make_full_configs.py",
                   "upstream": "minipoa (NCl3-lhd/POA)", "cases": out}, f)
    with open(MTDNA) as f:
        fa = f.read()
    mt = []
    for name, args in MTDNA_RUNS:
        p = subprocess.run([REF_BIN, MTDNA, *args, "-V", "1"], capture_output=True, text=True)
        scores = [int(l.split(":", 1)[1]) for l in p.stderr.splitlines()
                  if l.startswith("score:")]
        mt.append({"case": name, "args": args, "stdout": p.stdout, "scores": scores,
                   "returncode": p.returncode})
        print(f"{name:10s} {' '.join(args):30s} rc={p.returncode} calls={len(scores)} "
              f"stdout={len(p.stdout)}B", flush=True)
    with gzip.open(os.path.join(HERE, "mtdna_full.json.gz"), "wt") as f:
        json.dump({"generator": "tests/golden/make_full_configs.py",
                   "upstream": "minipoa (NCl3-lhd/POA)", "source": "data/mtDNA.fasta",
                   "fasta": fa, "fasta_sha256": hashlib.sha256(fa.encode()).hexdigest(),
                   "cases": mt}, f)


if __name__ == "__main__":
    main()
