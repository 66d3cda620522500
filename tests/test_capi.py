"""The C-ABI library loads and exports every function include/isoext_amd.h declares."""

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include/isoext_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(isoext_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "isoext_amd/libisoext_amd.so"))
    names = declared()
    assert len(names) >= 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_without_gpu():
    lib = ctypes.CDLL(str(ROOT / "isoext_amd/libisoext_amd.so"))
    lib.isoext_version.restype = ctypes.c_int
    assert lib.isoext_version() == 1
    lib.isoext_last_error.restype = ctypes.c_char_p
    assert lib.isoext_last_error() == b""
