"""The C ABI library loads and exports every function include/tpt.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import tpt
from tpt import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "tpt.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tpt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "tpt_trace_closest" in names and "tpt_trace_any" in names and "tpt_scene_create" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    bound = {name for name, _, _ in _lib._SIGNATURES}
    assert set(declared_functions()) <= bound


def test_last_error_is_a_string_and_null_arguments_are_rejected():
    L = _lib.lib()
    assert isinstance(L.tpt_last_error(), bytes)
    assert L.tpt_trace_closest(None, None, 1, None, 0) == _lib.TPT_ERROR_BAD_ARGUMENT
    assert b"null" in L.tpt_last_error()
    assert L.tpt_scene_create(None, None) is None
    assert L.tpt_qbvh_build(None, None, 0, 1) is None  # build_qbvh(&[]) is None
    assert tpt.build_qbvh([], []) is None
