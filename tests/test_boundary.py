"""The C-ABI boundary (include/hitomi_amd.h) without a GPU: the library loads
and exports every declared entry point, and the Python view binds them all."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hitomi_amd.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import hitomi_amd

    if not os.path.exists(hitomi_amd.LIB_PATH):
        pytest.fail("libhitomi_amd.so is not built: run __graft_entry__.build()")
    lib = ct.CDLL(hitomi_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import hitomi_amd

    assert sorted(hitomi_amd.exported_symbols()) == declared_symbols()


def test_import_does_not_touch_gpu():
    import hitomi_amd

    lib = hitomi_amd.load_library()
    assert lib.hm_last_error(None) == b"no context"
