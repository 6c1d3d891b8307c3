"""The per-slot arena layout on the host: tests/carve_check.cpp carves slots out of buffers
sized as the library sizes its pools and checks bounds and alignment for many capacities."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None and not os.path.exists(HIPCC),
                    reason="hipcc is needed to compile the layout check")
def test_slot_layout_fits(tmp_path):
    exe = tmp_path / "carve_check"
    subprocess.run([HIPCC, "-std=c++17", "-O1", f"-I{ROOT}/include", f"-I{ROOT}/poa_amd/csrc",
                    os.path.join(ROOT, "tests", "carve_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
