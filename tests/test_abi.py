"""The C-ABI library loads and exports every entry point include/vfs_wind.h
declares (no GPU call is made here)."""
import ctypes
import re
from pathlib import Path

from helpers import vfs

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "vfs_wind.h").read_text()
    return sorted(set(re.findall(r"\b(vfs_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding():
    assert set(declared_symbols()) == set(vfs.exported_symbols())


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(str(vfs.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_binding_loads():
    lib = vfs.load_library()
    assert lib.vfs_last_error() is not None
