"""minipoa's main loop written against the C++ facade (include/minipoa_amd.hpp, the program
poa_amd/minipoa_loop): on the golden fixtures it prints what upstream minipoa printed, in
consensus, MSA and GFA output, with and without -S and -B."""
from __future__ import annotations

import os
import subprocess

import pytest

import golden_util as gu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LOOP = os.path.join(ROOT, "poa_amd", "minipoa_loop")
CASES = [c for c in gu.load_cases()
         if c["case"] in ("mtdna40", "cfg2_seeded", "cfg2_small", "cfg4_small", "edge_one_read",
                          "edge_with_n")]

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(LOOP), reason="minipoa_loop not built")]


@pytest.mark.parametrize("case", CASES, ids=[gu.case_id(c) for c in CASES])
def test_facade_loop_matches_upstream(case, tmp_path):
    path = tmp_path / "r.fa"
    path.write_text(case["fasta"])
    print(f"facade {gu.case_id(case)}", flush=True)
    p = subprocess.run([LOOP, *case["args"], str(path)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout == case["stdout"]
