"""CPU checks of the C ABI boundary: the HIP library loads without a GPU and
exports every entry point include/plamatrix_amd.h declares (no compute calls)."""
import os
import re

from plamatrix_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "plamatrix_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_exported_list():
    assert declared_symbols() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.pm_version() >= 100


def test_errors_without_device_work():
    lib = capi.lib()
    code = lib.pm_workspace_create(None)
    assert code == 1
    assert b"null" in lib.pm_last_error()
