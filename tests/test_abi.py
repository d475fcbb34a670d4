"""The C-ABI library loads and exports every symbol include/gapx.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

import gapx

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gapx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gapx_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(gapx.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gapx.LIB_PATH):
        pytest.fail("libgapx.so is not built: run python calypso-gap_amd/build.py")
    lib = ctypes.CDLL(gapx.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_string():
    L = gapx.lib()
    assert L.gapx_version() >= 100
    assert isinstance(L.gapx_last_error(), bytes)


def test_missing_model_file_is_an_error(tmp_path):
    with pytest.raises(gapx.GapxError) as e:
        gapx.Model.read(str(tmp_path / "gap_parameters"))
    assert e.value.code == 2
