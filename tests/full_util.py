"""The full-size single-window fixtures (tests/golden/full_configs.json.gz)."""
from __future__ import annotations

import gzip
import hashlib
import json
import os

from poa_amd import synth

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "full_configs.json.gz")


def load_cases() -> list[dict]:
    with gzip.open(PATH, "rt") as f:
        return json.load(f)["cases"]


def reads_of(case: dict) -> synth.Reads:
    gen = case["generator"]
    assert gen.startswith("synth.") and gen.endswith("()")
    return getattr(synth, gen[len("synth."):-2])()


def fasta_of(reads: synth.Reads):
    """(SHA-256, text) of the FASTA the fixture generator wrote for these reads."""
    import tempfile
    with tempfile.NamedTemporaryFile("r", suffix=".fa", delete=False) as f:
        path = f.name
    try:
        reads.to_fasta(path)
        with open(path) as f:
            text = f.read()
    finally:
        os.unlink(path)
    return hashlib.sha256(text.encode()).hexdigest(), text
