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


def test_batch_case_bytes_and_argument_checks():
    """Footprint bound and option validation of the batch engine (host only, no device call)."""
    import pytest

    from euler3d_amd import Batch, Mesh

    m = Mesh.naca(ni=33, nj=17, nk=2).preprocess(1.0)
    b1 = Batch.case_bytes(m, max_iter=100)
    b2 = Batch.case_bytes(m, max_iter=10000)
    assert b1 > 0 and b2 - b1 >= 8 * 9900 - 256  # the residual history is part of the footprint
    lib = native.lib()
    o = native.BatchOpts()
    assert lib.e3d_batch_opts_default(ctypes.byref(o)) == 0
    assert o.max_cases > 0 and o.wg_per_cu == 0 and o.use_graph == 1  # 0: one workgroup per (case, block)
    o.max_cases = 0
    h = ctypes.c_void_p()
    assert lib.e3d_batch_create(0, ctypes.byref(o), ctypes.byref(h)) == 1
    assert "max_cases" in lib.e3d_last_error().decode()
    raw = Mesh.box(2, 2, 2)  # not preprocessed; kept alive across the call that reads it
    with pytest.raises(native.E3DError):
        native.check(lib.e3d_batch_case_bytes(raw._h, None, ctypes.byref(ctypes.c_double())))
