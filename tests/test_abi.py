"""The C-ABI library loads and exports every entry point include/e3d.h declares (no GPU needed)."""
from __future__ import annotations

import ctypes
import re
import subprocess

from conftest import ROOT

from euler3d_amd import native


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "e3d.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(e3d_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_entry_points():
    decl = declared_functions()
    assert len(decl) >= 20
    assert set(native.EXPORTED) <= set(decl)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(native.LIB_PATH))
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(native.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(declared_functions()) <= exported


def test_version_and_error_channel():
    assert "gfx950" in native.version()
    import pytest

    with pytest.raises(native.E3DError):
        native.check(native.lib().e3d_mesh_read(b"/nonexistent.par", 1, ctypes.byref(ctypes.c_void_p())))
    assert "opening" in native.lib().e3d_last_error().decode()


def test_solver_executable_usage():
    exe = ROOT / "euler3d_amd" / "bin" / "E3D_Solver"
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode != 0 and "Usage" in r.stderr
