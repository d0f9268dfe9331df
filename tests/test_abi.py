"""The C-ABI library loads and exports every entry point include/wedg.h declares
(no GPU needed; no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "wedg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wedg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from edg_westervelt_amd import _lib
    names = declared_functions()
    assert len(names) >= 18
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert _lib.lib.wedg_version().startswith(b"wedg")


def test_struct_layout_matches_header():
    """The ctypes mirror of wedg_pre has one pointer per table field of the header."""
    from edg_westervelt_amd import _lib
    text = open(os.path.join(ROOT, "include", "wedg.h")).read()
    body = text[text.index("typedef struct wedg_pre"):text.index("} wedg_pre;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    ptrs = re.findall(r"\*(\w+)", body)
    assert tuple(ptrs) == _lib._PTRS
