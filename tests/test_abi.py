"""The C ABI library loads, exports every entry point include/stg.h declares,
and the Python binding covers exactly those entry points. No GPU calls."""
import ctypes
import re
from pathlib import Path

from steps_amd import _native

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "stg.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:int|const char \*)\s*(stg_\w+)\s*\(", text, flags=re.M))


def test_header_declares_entry_points():
    names = declared()
    assert {"stg_create", "stg_create_batch", "stg_run", "stg_destroy", "stg_last_error",
            "stg_get_tet_counts", "stg_set_tet_counts"} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in declared():
        assert hasattr(lib, name), name


def test_binding_matches_header():
    assert set(_native.SIGNATURES) == declared()


def test_library_has_gfx950_code_object():
    data = _native.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_create_error_without_device_or_with_bad_args():
    """A create with invalid arguments fails with ArgErr before touching the device."""
    lib = _native.load()
    h = ctypes.c_void_p()
    rc = lib.stg_create_batch(None, 0, None, None, 0, ctypes.byref(h))
    assert rc == _native.STG_ERR_ARG
    assert b"realization" in lib.stg_last_error(None)
