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


def test_handles_are_not_destroyed_by_a_forked_child(monkeypatch):
    """A forked child that garbage-collects an engine or batch of its parent must not call into the
    HIP runtime: close() calls the destroy entry point only in the creating process."""
    import os
    from mjolnir_amd import capi

    calls = []

    class Stub:
        def mjg_engine_destroy(self, h):
            calls.append(("engine", h))

        def mjg_batch_destroy(self, h):
            calls.append(("batch", h))

    monkeypatch.setattr(capi, "lib", lambda: Stub())
    for cls, kind in ((capi.Engine, "engine"), (capi.Batch, "batch")):
        obj = cls.__new__(cls)
        obj.h, obj._pid = 1234, os.getpid() + 1          # created by another process
        obj.close()
        assert obj.h is None and calls == []
        obj.h, obj._pid = 1234, os.getpid()              # created here
        obj.close()
        assert calls == [(kind, 1234)]
        calls.clear()
