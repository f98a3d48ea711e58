"""The C ABI library loads without a GPU and exports every entry point
include/mdc.h declares; no compute calls are made here."""
import ctypes
import os
import re

from mdcraft_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "mdc.h")).read()
    decl = r"^\s*(?:mdc_status|int|const char\s*\*)\s*(mdc_[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_header_and_binding_agree():
    assert declared() == sorted(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_loads_and_reports_record_size():
    lib = engine.load_library()
    assert lib.mdc_atom_record_size() == 208


def test_errors_without_device_are_reported():
    # with no GPU the context call fails cleanly with a message
    lib = engine.load_library()
    h = ctypes.c_void_p()
    st = lib.mdc_ctx_create(-1, ctypes.byref(h))
    assert st != 0
    assert lib.mdc_last_error()
