"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: nothing here computes)."""

import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "libpt_amd", "libpt_amd.so")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(libpt_\w+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["pt.h", "libpt_amd.h"])
def test_exports(header):
    assert os.path.exists(LIB), "build first: python -m libpt_amd.build"
    lib = ctypes.CDLL(LIB)
    names = _declared(header)
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_pt_h_has_the_reference_symbols():
    # the 11 functions of the reference header src/pt.h:27-178
    want = {"libpt_print_banner", "libpt_set_malloc", "libpt_set_free"}
    for k in ("rpt", "upt", "rft", "uft"):
        want |= {f"libpt_{k}", f"libpt_{k}_mp"}
    assert set(_declared("pt.h")) == want


def test_headers_compile_as_c():
    src = '#include "pt.h"\n#include "libpt_amd.h"\nint main(void){return 0;}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c",
                        "-I", os.path.join(ROOT, "include"), "-fsyntax-only",
                        "-"], input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr


def test_banner_and_hooks():
    lib = ctypes.CDLL(LIB)
    lib.libpt_set_malloc(None)
    lib.libpt_set_free(None)
