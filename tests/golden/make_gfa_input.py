#!/usr/bin/env python3
"""Writes tests/golden/gfa_input.json.gz: minipoa's incremental mode (-i graph.gfa, src/main.cpp:86
and src/handle_input.cpp:444-495) run by the upstream binary (oracle/_ref/minipoa, built by
`make -C oracle ref`). Each case makes a GFA with upstream's own -r 2 from one read set (or
takes a hand-written GFA), then aligns a second read set to it under the case's flags. The
fixture holds the GFA and the second read set per input, and per run the flags, upstream's
stdout, its -V 1 scores and its exit status.

Read sets: the first 5 and the next 4 reads of the reference's data/mtDNA.fasta; two windows
drawn by poa_amd.synth (500 bp, 12 + 8 reads at 10% error; 1,200 bp, 16 + 16 reads at 5%);
and a hand-written GFA with multi-base segments, a bubble, lowercase and N bases, H/L/comment
lines that read_gfa skips, a path that starts past the first segment, and a path name with a
blank in it."""
from __future__ import annotations

import gzip
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

HAND_GFA = (
    "H\tVN:Z:1.0\n"
    "# a comment line\n"
    "S\ts1\tACGTTGCA\n"
    "S\ts2\tGGAT\n"
    "S\ts3\tgcaN\n"
    "L\ts1\t+\ts2\t+\t0M\n"
    "L\ts1\t+\ts3\t+\t0M\n"
    "S\ts4\tTTACGGATCCA\n"
    "L\ts2\t+\ts4\t+\t0M\n"
    "L\ts3\t+\ts4\t+\t0M\n"
    "P\tp1\ts1+,s2+,s4+\t*\n"
    "P\tp2\ts1+,s3+,s4+\t*\n"
    "P\tp3 with blank\ts2+,s4+\n"
    "P\tp4\ts1+,s2+,s4+\t*\n"
)
HAND_FASTA = (">q1 first read\nACGTTGCAGGATTTACGGATCCA\n>q2\nACGTTGCAGCANTTACGGTCCA\n"
              ">q3\nACGTGCAGGATTTACGGATCCAT\n")

FLAGS = [["-r", "0"], ["-r", "1"], ["-B", "-r", "1"], ["-S", "-r", "1"], ["-S", "-r", "0"],
         ["-p", "-r", "1"]]


def fasta_of(reads, prefix: str) -> str:
    return "".join(f">{prefix}{i} {prefix} read {i}\n{reads.read(i).decode()}\n"
                   for i in range(reads.n))


def mtdna_sets() -> tuple[str, str]:
    recs = open(MTDNA).read().split(">")[1:]
    text = [">" + r if r.endswith("\n") else ">" + r + "\n" for r in recs]
    return "".join(text[:5]), "".join(text[5:9])


def run(args, fa_path, gfa_path=None):
    cmd = [REF_BIN, fa_path, *args, "-V", "1"]
    if gfa_path:
        cmd += ["-i", gfa_path]
    return subprocess.run(cmd, capture_output=True, text=True)


def main() -> None:
    sets = []
    a, b = mtdna_sets()
    sets.append(("mtdna5_4", a, b, None))
    w = synth.window(11, 500, 20, 0.04, 0.03, 0.03)
    fa = fasta_of(w, "w")
    lines = fa.splitlines(keepends=True)
    sets.append(("synth500", "".join(lines[:24]), "".join(lines[24:]), None))
    w = synth.window(12, 1200, 32, 0.02, 0.015, 0.015)
    fa = fasta_of(w, "v")
    lines = fa.splitlines(keepends=True)
    sets.append(("synth1200", "".join(lines[:32]), "".join(lines[32:]), None))
    sets.append(("hand", None, HAND_FASTA, HAND_GFA))
    cases, inputs = [], {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, base_fa, add_fa, gfa in sets:
            if gfa is None:
                with open(os.path.join(tmp, "base.fa"), "w") as f:
                    f.write(base_fa)
                p = run(["-r", "2"], os.path.join(tmp, "base.fa"))
                assert p.returncode == 0, p.stderr
                gfa = p.stdout
            gpath = os.path.join(tmp, "in.gfa")
            with open(gpath, "w") as f:
                f.write(gfa)
            apath = os.path.join(tmp, "add.fa")
            with open(apath, "w") as f:
                f.write(add_fa)
            for args in FLAGS:
                p = run(args, apath, gpath)
                scores = [int(ln.split(":", 1)[1]) for ln in p.stderr.splitlines()
                          if ln.startswith("score:")]
                cases.append({"case": name, "args": args, "returncode": p.returncode,
                              "stdout": p.stdout, "scores": scores,
                              "died": "terminate called" in p.stderr})
                print(f"{name:10s} {' '.join(args):12s} rc={p.returncode} calls={len(scores)} "
                      f"stdout={len(p.stdout)}B", flush=True)
            inputs[name] = {"gfa": gfa, "fasta": add_fa}
    with gzip.open(os.path.join(HERE, "gfa_input.json.gz"), "wt") as f:
        json.dump({"generator": "tests/golden/make_gfa_input.py",
                   "upstream": "minipoa (NCl3-lhd/POA)", "inputs": inputs, "cases": cases}, f)


if __name__ == "__main__":
    main()
