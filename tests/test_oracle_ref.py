"""The CPU oracle against the upstream minipoa binary built from its own sources
(oracle/_ref/minipoa, `make -C oracle ref`), on fresh seeded inputs beyond the fixtures:
identical stdout and identical per-call scores. Skipped where the binary is not built."""
from __future__ import annotations

import os
import subprocess
import tempfile

import pytest

import golden_util as gu
import oracle_lib
from poa_amd import synth

pytestmark = pytest.mark.skipif(not os.path.exists(oracle_lib.REF_BIN),
                                reason="oracle/_ref/minipoa not built")


def run_ref(reads: synth.Reads, args: list[str]):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "r.fa")
        reads.to_fasta(path)
        p = subprocess.run([oracle_lib.REF_BIN, path, *args, "-V", "1"], capture_output=True,
                           text=True, timeout=300)
        with open(path) as f:
            fasta = f.read()
    scores = [int(l.split(":", 1)[1]) for l in p.stderr.splitlines() if l.startswith("score:")]
    return p.returncode, p.stdout, scores, fasta


CASES = [
    ("cfg3", 0, ["-r", "0"]), ("cfg3", 1, ["-r", "1"]), ("cfg3", 2, ["-B", "-r", "0"]),
    ("cfg3", 3, ["-r", "2"]), ("cfg2", 4, ["-S", "-r", "1"]), ("cfg2", 5, ["-r", "1"]),
    ("cfg4", 6, ["-S", "-k", "15", "-w", "10", "-r", "0"]),
    ("cfg2", 7, ["-M", "4", "-X", "-6", "-O", "-3", "-E", "-1", "-r", "0"]),
]


def reads_for(kind: str, seed: int) -> synth.Reads:
    if kind == "cfg3":
        return synth.config3(seed=9000 + seed, n_windows=1).window(0)
    if kind == "cfg2":
        return synth.window(9000 + seed, 1200, 16, 0.04, 0.03, 0.03, 0.7)
    return synth.window(9000 + seed, 4000, 6, 0.024, 0.048, 0.048, 0.7)


@pytest.mark.parametrize("kind,seed,args", CASES, ids=[f"{k}-{s}-{' '.join(a)}" for k, s, a in CASES])
def test_oracle_equals_reference_binary(kind, seed, args):
    reads = reads_for(kind, seed)
    rc, out, scores, fasta = run_ref(reads, args)
    assert rc == 0
    seq, off, names, comments = gu.parse_fasta(fasta)
    text, stats = oracle_lib.window(gu.parse_args(args), seq, off, names, comments)
    assert [s.score for s in stats] == scores
    assert text == out
