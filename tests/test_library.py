"""CPU checks of the product library: it is built, loads, and exports exactly
the C ABI declared in include/mdgpu.h (no compute calls without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from mdacp_amd import mdgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mdgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(mdgpu_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_surface():
    assert declared_symbols() == sorted(mdgpu.EXPORTS + ["mdgpu_last_error"]) or \
        set(mdgpu.EXPORTS) <= set(declared_symbols())


def test_library_loads_and_exports_every_declared_symbol():
    assert os.path.exists(mdgpu.LIB_PATH), "libmdgpu.so not built"
    lib = ctypes.CDLL(mdgpu.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", mdgpu.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mdgpu_\w+)", out))
    assert set(declared_symbols()) == exported


def test_library_is_gfx950_code_object():
    # the offload bundle of this library (built here from source) names its target
    data = open(mdgpu.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_no_cpu_fallback_in_product():
    # the product package never imports the oracle
    for dirpath, _, files in os.walk(os.path.join(ROOT, "mdacp_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".cc", ".h")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "liboracle" not in text and "from oracle" not in text and "import oracle" not in text, f
