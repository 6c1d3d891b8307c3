"""The per-position pieces of device seeding on the host (no GPU): tests/sketch_check.cpp runs
poa_amd/csrc/seed_core.h's per-position minimizer emission over 20,000 random sequences (with
ambiguous bases, tandem repeats and copied stretches) against the oracle's sequential mm_sketch,
and the chaining log2 and pair score against the oracle's, bit for bit."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is needed for the host check")
def test_sketch_and_chain_pieces_match_oracle(tmp_path):
    exe = tmp_path / "sketch_check"
    subprocess.run(["g++", "-std=c++20", "-O2", "-fopenmp", "-Wno-unknown-pragmas",
                    f"-I{ROOT}/oracle", f"-I{ROOT}/poa_amd/csrc",
                    os.path.join(ROOT, "tests", "sketch_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe), "20000"], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
