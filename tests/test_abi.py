"""The C-ABI libraries load without a GPU and export every function include/*.h declares.
No compute calls are made here."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^[A-Za-z_][\w \*]*?\b(mpoa_\w+)\s*\(", re.M)

LIBS = {
    "mpoa.h": os.path.join(ROOT, "poa_amd", "libmpoa.so"),
    "mpoa_synth.h": os.path.join(ROOT, "poa_amd", "libmpoa_synth.so"),
}


def declared(header: str) -> list[str]:
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


# header-only C++ facades and the C header whose library they call
FACADES = {"minipoa_amd.hpp": "mpoa.h"}


def test_every_header_has_a_library():
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == sorted(list(LIBS) + list(FACADES))


@pytest.mark.parametrize("facade", sorted(FACADES))
def test_facade_compiles(facade, tmp_path):
    """The facade compiles on its own and calls only functions its C header declares."""
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text(f'#include "{facade}"\nint main() {{ return 0; }}\n')
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{os.path.join(ROOT, 'include')}", str(src)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    with open(os.path.join(ROOT, "include", facade)) as f:
        called = set(re.findall(r"\b(mpoa_\w+)\s*\(", f.read()))
    assert called and called <= set(declared(FACADES[facade]))


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_declared_symbols(header):
    so = LIBS[header]
    assert os.path.exists(so), f"{so} not built: run __graft_entry__.build()"
    lib = ctypes.CDLL(so)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_every_symbol():
    from poa_amd import api
    assert sorted(api.EXPORTS) == declared("mpoa.h")


def test_strerror_without_device():
    from poa_amd import api
    lib = api.lib()
    for rc in list(api.ERRORS) + [0]:
        assert lib.mpoa_strerror(rc)
