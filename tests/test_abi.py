"""The C-ABI library loads and exports every entry point include/wedg.h declares
(no GPU needed; no compute calls)."""
import ctypes
import os
import re

import pytest

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


def test_entry_points_reject_null_handles():
    """Every entry point that takes a handle checks its arguments before any
    HIP call: null handles and pointers give WEDG_ERR_ARGUMENT (the destroy
    functions accept null, as free does). No device work is done."""
    import ctypes as C
    from edg_westervelt_amd import _lib
    skip = {"wedg_probe_copy", "wedg_device_info", "wedg_version"}   # these touch the device
    checked = 0
    for name in declared_functions():
        if name in skip:
            continue
        f = getattr(_lib.lib, name)
        args = []
        for t in f.argtypes:
            if t in (C.c_int, C.c_long, C.c_size_t):
                args.append(0)
            elif t is C.c_double:
                args.append(0.0)
            else:
                args.append(None)
        st = f(*args)
        expect = _lib.OK if name.endswith("_destroy") else _lib.ERR_ARGUMENT
        assert st == expect, (name, st)
        checked += 1
    assert checked >= 30


def test_workspace_rejects_unknown_lane_counts():
    """The lane count is checked before anything reaches the device."""
    import edg_westervelt_amd as w
    from edg_westervelt_amd import precompute
    pre, _ = precompute(2, 2, mixed=False, mode="initial_value")
    for bad in (0, 3, "two"):
        with pytest.raises(ValueError, match="lanes"):
            w.Workspace(pre, 2, lanes=bad)


@pytest.mark.parametrize("given,keep,expect", [(None, None, "16"), ("4", None, "16"), ("32", None, "32"),
                                               ("4", "1", "4")])
def test_hardware_queue_setting(given, keep, expect):
    """Importing the binding raises GPU_MAX_HW_QUEUES to 16 when it is unset or
    lower (the batched mode and the gradient run concurrent streams, which
    share hardware queues otherwise), keeps a higher value, and keeps any value
    with WEDG_KEEP_HW_QUEUES=1. A fresh interpreter each time: the setting is
    made at import."""
    import subprocess
    import sys
    env = {k: v for k, v in os.environ.items() if k not in ("GPU_MAX_HW_QUEUES", "WEDG_KEEP_HW_QUEUES")}
    if given is not None:
        env["GPU_MAX_HW_QUEUES"] = given
    if keep is not None:
        env["WEDG_KEEP_HW_QUEUES"] = keep
    out = subprocess.run([sys.executable, "-c", "import os, edg_westervelt_amd._lib; "
                          "print(os.environ['GPU_MAX_HW_QUEUES'])"],
                         cwd=ROOT, env=env, capture_output=True, text=True, check=True)
    assert out.stdout.strip() == expect
