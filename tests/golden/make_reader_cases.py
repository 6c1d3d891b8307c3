#!/usr/bin/env python3
"""Writes tests/golden/reader_cases.json: input files that exercise readFile's parsing
(src/handle_input.cpp:13-43, kseq_read of include/kseq.h:181-223) and the upstream binary's
-r 1 output on each (oracle/_ref/minipoa, built by `make -C oracle ref`). The MSA rows carry
every read's name, comment and bases (as codes), so they pin the parser: CRLF line ends, tabs
and a lone carriage return in headers, blanks inside sequence lines, empty lines, text before
the first header, an empty name, an empty sequence, lowercase and IUPAC letters, FASTQ with
wrapped sequence and quality lines (a quality line that starts with '@', a '+' line with a name),
a FASTQ whose last quality string is cut short, and gzip input."""
from __future__ import annotations

import gzip
import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "minipoa")


def variants(seed: int, n: int, length: int = 70) -> list[str]:
    rng = random.Random(seed)
    t = "".join(rng.choice("ACGT") for _ in range(length))
    out = []
    for _ in range(n):
        s = list(t)
        for _ in range(3):
            p = rng.randrange(len(s))
            s[p] = rng.choice("ACGT")
        out.append("".join(s))
    return out


def cases() -> list[tuple[str, str, bool]]:
    """(name, text, gzip the file)."""
    v = variants(1, 6)
    crlf = ("junk before the first header\r\n"
            f">r1\tcomment\twith tabs\r\n{v[0][:35]}\r\n{v[0][35:]}\r\n"
            f">r2 \r\n{v[1]}\r\n"
            f">r3\r\n\r\n{v[2][:20]}\r\n\r\n{v[2][20:]}\r\n"
            f">r4 plain comment\r\n{v[3]}")
    w = variants(2, 5)
    blanks = (f"; a comment line\n>b1 first\n{w[0][:30]} {w[0][30:]}\n\n"
              f"> only a comment\n{w[1].lower()}\n"
              f">b3\n{w[2][:10]}RYKM{w[2][14:]}\n"
              f">b4 empty read next\n{w[3]}\n>b5\n>b6  two blanks\n{w[4]}\n")
    q = variants(3, 4)
    fq = (f"@q1 fastq read\n{q[0][:40]}\n{q[0][40:]}\n+q1\n{'I' * 40}\n{'#' * 30}\n"
          f"@q2\n{q[1]}\n+\n@{'I' * 69}\n"
          f"@q3 three\n{q[2]}\n+\n{'5' * 70}\n"
          f"@q4\n{q[3]}\n+\n{'5' * 70}\n")
    cut = fq + f"@q5 cut short\n{q[0]}\n+\n{'I' * 50}\n"
    return [("crlf", crlf, False), ("blanks", blanks, False), ("fastq", fq, False),
            ("fastq_gz", fq, True), ("fastq_cut", cut, False)]


def main() -> None:
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        for name, text, gz in cases():
            path = os.path.join(tmp, name + (".fq.gz" if gz else ".txt"))
            data = text.encode()
            with open(path, "wb") as f:
                f.write(gzip.compress(data, mtime=0) if gz else data)
            p = subprocess.run([REF_BIN, path, "-r", "1"], capture_output=True)
            out.append({"case": name, "text": text, "gzip": gz, "returncode": p.returncode,
                        "stdout": p.stdout.decode("latin-1")})
            print(f"{name:10s} rc={p.returncode} rows={p.stdout.count(b'>')}", flush=True)
    with open(os.path.join(HERE, "reader_cases.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_reader_cases.py",
                   "upstream": "minipoa (NCl3-lhd/POA)", "cases": out}, f, indent=1)


if __name__ == "__main__":
    main()
