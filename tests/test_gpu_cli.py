"""The command-line tool (poa_amd/minipoa_amd, minipoa's flags) on the golden fixtures: same
stdout and the same -V 1 score lines as upstream minipoa; several files as one batch; files that
exercise the reader."""
from __future__ import annotations

import gzip
import json
import os
import subprocess

import pytest

import golden_util as gu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "poa_amd", "minipoa_amd")
CASES = [c for c in gu.load_cases() if c["case"] in ("mtdna40", "cfg2_seeded", "edge_with_n",
                                                       "cfg1_small")]

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(CLI), reason="minipoa_amd not built")]


@pytest.mark.parametrize("case", CASES, ids=[gu.case_id(c) for c in CASES])
def test_cli_matches_upstream(case, tmp_path):
    path = tmp_path / "r.fa"
    path.write_text(case["fasta"])
    print(f"cli {gu.case_id(case)}", flush=True)
    p = subprocess.run([CLI, str(path), *case["args"], "-V", "1"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    scores = [int(l.split(":", 1)[1]) for l in p.stderr.splitlines() if l.startswith("score:")]
    assert scores == case["scores"]
    assert p.stdout == case["stdout"]


def test_cli_batch_of_files(tmp_path):
    sel = [c for c in gu.load_cases() if c["args"] == ["-r", "0"]][:8]
    paths = []
    for i, c in enumerate(sel):
        pth = tmp_path / f"w{i}.fa"
        pth.write_text(c["fasta"])
        paths.append(str(pth))
    print("cli batch", flush=True)
    p = subprocess.run([CLI, *paths, "-r", "0"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout == "".join(c["stdout"] for c in sel)


READER = json.load(open(os.path.join(ROOT, "tests", "golden", "reader_cases.json")))["cases"]


@pytest.mark.parametrize("case", READER, ids=lambda c: c["case"])
def test_cli_reader_cases(case, tmp_path):
    """Files that exercise readFile's parsing (CRLF, blanks, FASTQ, gzip, a cut-short record):
    the same MSA text as upstream."""
    path = tmp_path / ("in.fq.gz" if case["gzip"] else "in.txt")
    data = case["text"].encode("latin-1")
    path.write_bytes(gzip.compress(data) if case["gzip"] else data)
    print(f"cli reader {case['case']}", flush=True)
    p = subprocess.run([CLI, str(path), "-r", "1"], capture_output=True, timeout=120)
    assert p.returncode == case["returncode"], p.stderr
    assert p.stdout.decode("latin-1") == case["stdout"]
