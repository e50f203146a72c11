"""liblisa.so loads without a GPU and exports every function include/lisa.h
declares; the host-only entry points work; device entry points fail with a
status (never a crash) when no device is present or arguments are bad."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "lisa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lisa_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree(capi):
    names = declared()
    assert len(names) >= 20
    assert sorted(capi.EXPORTS) == names


def test_every_symbol_exported(capi):
    L = C.CDLL(capi.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name


def test_errors_are_status_codes(capi):
    L = capi.lib()
    with pytest.raises(capi.LisaError):
        capi.save_ppm("/nonexistent-dir/x.ppm", [0] * 16, 2, 2)
    assert L.lisa_render(None, None, None, None) == capi.ERR_INVALID_ARG
    assert b"ctx is NULL" in L.lisa_last_error(None)
    assert L.lisa_set_scene(None, None, None, None, 0, None, 0) == capi.ERR_INVALID_ARG


def test_oracle_is_not_linked_by_the_product(capi):
    import subprocess
    out = subprocess.run(["ldd", capi.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    src = os.path.join(ROOT, "lisa_amd")
    for dirpath, _, files in os.walk(src):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                body = open(os.path.join(dirpath, f)).read()
                assert "orc_" not in body and "import oracle" not in body, f
