"""The C-ABI library loads and exports every entry point include/biblbm.h
declares (no GPU calls here)."""
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "biblbm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(bl_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "bl_collide_stream" in syms and "bl_ibm_step" in syms and "bl_last_error" in syms
    assert len(syms) >= 40


def test_library_exports_every_declared_symbol():
    from biblbm_amd import capi
    missing = [s for s in declared_symbols() if not hasattr(capi.lib, s)]
    assert not missing, "libbiblbm.so lacks: %s" % missing


def test_binding_covers_every_declared_symbol():
    from biblbm_amd import capi
    bound = set(capi.SIGNATURES) | {"bl_last_error"}
    assert set(declared_symbols()) <= bound


def test_library_is_gfx950_code_object():
    from biblbm_amd import capi
    data = open(capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data
