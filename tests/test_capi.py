"""The C-ABI library loads and exports every entry point include/oagpu.h
declares, and the Python binding declares exactly those (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oagpu.h")
LIB = os.path.join(ROOT, "openaccel_amd", "liboagpu.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oagpu_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "oagpu_assemble_momentum" in names and "oagpu_solve" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("liboagpu.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(oagpu_[a-z_0-9]+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_binding_matches_header():
    import openaccel_amd as oa
    if not os.path.exists(LIB):
        pytest.skip("liboagpu.so not built")
    L = oa.load_library()
    for n in declared():
        assert n in oa._SIGS, n
        getattr(L, n)  # resolves


def test_missing_library_fails_loudly(tmp_path):
    import openaccel_amd as oa
    saved = oa._lib
    try:
        oa._lib = None
        with pytest.raises(oa.OagpuError):
            oa.load_library(str(tmp_path / "nope.so"))
    finally:
        oa._lib = saved
