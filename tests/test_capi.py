"""The C-ABI library: it loads, and it exports every entry point include/mjg.h declares.
No compute call is made (no GPU needed)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mjg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mjg_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "mjg_engine_create" in syms and "mjg_step" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from mjolnir_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert missing == []
    assert sorted(capi.EXPORTED) == declared_symbols()


def test_version_string():
    from mjolnir_amd import capi
    assert b"gfx950" in capi.lib().mjg_version()
