"""In-tree build of the native libraries.

* ``poa_amd/libmpoa.so``        the product: HIP kernels for gfx950 + the C ABI (include/mpoa.h)
* ``poa_amd/libmpoa_synth.so``  the seeded workload generator (include/mpoa_synth.h)
* ``poa_amd/minipoa_amd``       the command-line tool with minipoa's flags
* ``poa_amd/minipoa_loop``      minipoa's main loop written against the C++ facade
* ``oracle/``                   the CPU checker (oracle/Makefile), and, where the upstream
                                sources are present, the upstream binary in oracle/_ref/
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "poa_amd")
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPOA_ARCH", "gfx950")
REF_DIR = os.environ.get("MPOA_REF_DIR", "/root/reference")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions stay off (the DP is integer code; the scheduling network's
    # float code runs unpacked).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

PRODUCT_SOURCES = ["pipeline.hip", "mlp.hip", "mpoa.cpp", "graph_api.cpp", "sched.cpp", "seed.cpp",
                   "reads.cpp"]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_synth() -> str:
    out = os.path.join(PKG, "libmpoa_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{INCLUDE}", "-o", out,
          os.path.join(CSRC, "synth.cpp")])
    return out


def build_product() -> str:
    out = os.path.join(PKG, "libmpoa.so")
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in PRODUCT_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"])
    os.replace(tmp, out)
    return out


def build_cli() -> str | None:
    src = os.path.join(CSRC, "cli.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(PKG, "minipoa_amd")
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", "-o", out, src,
          f"-L{PKG}", "-lmpoa", f"-Wl,-rpath,$ORIGIN", "-lz"])
    return out


def build_loop() -> str:
    """minipoa's main loop over the C++ facade (include/minipoa_amd.hpp)."""
    out = os.path.join(PKG, "minipoa_loop")
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", "-o", out,
          os.path.join(CSRC, "minipoa_loop.cpp"), f"-L{PKG}", "-lmpoa", "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    if os.path.isdir(os.path.join(REF_DIR, "src")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF_DIR={REF_DIR}"])


def build_all() -> None:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_synth()
    build_product()
    build_cli()
    build_loop()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "synth": build_synth, "product": build_product, "cli": build_cli, "loop": build_loop,
         "oracle": build_oracle}[w]()
