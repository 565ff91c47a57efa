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


STRUCTS = {"vfs_params": "Params", "vfs_solve_opts": "SolveOpts", "vfs_solve_stats": "SolveStats",
           "vfs_step_stats": "StepStats", "vfs_newton_opts": "NewtonOpts", "vfs_newton_stats": "NewtonStats",
           "vfs_turbine": "Turbine", "vfs_turbine_control": "TurbineControl", "vfs_foil": "Foil"}


def test_struct_layouts_match_header(tmp_path):
    """Every struct the binding mirrors has the header's size and field
    offsets, as the host C compiler lays them out."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        import pytest
        pytest.skip("no host C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vfs_wind.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        st = getattr(vfs, pyname)
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in st._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, pyname in STRUCTS.items():
        st = getattr(vfs, pyname)
        assert got[(cname, "size")] == ctypes.sizeof(st), cname
        for f in st._fields_:
            assert got[(cname, f[0])] == getattr(st, f[0]).offset, (cname, f[0])
