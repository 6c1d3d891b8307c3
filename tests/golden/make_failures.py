#!/usr/bin/env python3
"""Writes tests/golden/failures.json: read sets on which upstream minipoa stops in its DP.

A seeded search draws small read sets (a random template and reads carrying large insertions
and deletions) with narrow bands and a gap-open penalty of 0, runs the upstream binary
(oracle/_ref/minipoa, built by `make -C oracle ref`) on each, and keeps the first cases where
it prints " backtrack error" and exits 1 (src/align.cpp:1730-1731). The search also counts
the cases that left the band (the "l:"/"r:" lines and exit 0 of src/align.cpp:1584-1588;
band_exits_seen): each backtrack move's own test keeps the next cell inside the predecessor's
band, so none is expected. The fixture holds inputs and upstream's exit status and stderr
only."""
from __future__ import annotations

import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "minipoa")
SEEDS = (2, 5, 6)
TRIALS = 3000


def draw(rng: random.Random):
    def rs(n):
        return "".join(rng.choice("ACGT") for _ in range(n))
    t = rs(rng.randint(20, 300))
    reads = [t]
    for _ in range(rng.randint(1, 3)):
        s = list(t)
        for _ in range(rng.randint(1, 3)):
            pos = rng.randint(0, len(s))
            if rng.random() < 0.5:
                s[pos:pos] = rs(rng.randint(1, len(t)))
            else:
                del s[pos:pos + rng.randint(1, max(1, len(s) // 2))]
        reads.append("".join(s) or "A")
    rng.shuffle(reads)
    args = ["-b", str(rng.choice([0, 0, 1, 2])), "-f", str(rng.choice([100000, 1000, 50, 20]))]
    if rng.random() < 0.5:
        args.append("-B")
    if rng.random() < 0.3:
        args += ["-M", str(rng.randint(1, 5)), "-X", str(rng.randint(0, 6)), "-O",
                 str(rng.randint(0, 8)), "-E", str(rng.randint(1, 4))]
    return "".join(f">r{i}\n{r}\n" for i, r in enumerate(reads)), args


def main() -> None:
    cases, band_exits, tried = [], 0, 0
    for seed in SEEDS:
        rng = random.Random(seed)
        for _ in range(TRIALS):
            fa, args = draw(rng)
            tried += 1
            with tempfile.NamedTemporaryFile("w", suffix=".fa", delete=False) as f:
                f.write(fa)
                path = f.name
            try:
                p = subprocess.run([REF_BIN, path, *args], capture_output=True, text=True)
            finally:
                os.unlink(path)
            band_exits += "l:" in p.stderr
            if "backtrack error" in p.stderr:
                cases.append({"case": f"backtrack_s{seed}", "args": args, "fasta": fa,
                              "returncode": p.returncode, "stderr": p.stderr,
                              "stdout": p.stdout})
                print(f"seed {seed}: {' '.join(args)} rc={p.returncode}", flush=True)
                break
    with open(os.path.join(HERE, "failures.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_failures.py",
                   "upstream": "minipoa (NCl3-lhd/POA)", "searched": tried,
                   "band_exits_seen": band_exits, "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
