"""The Tecplot post-processor, `E3D_POST <case.e3d>` and e3d_post_tecplot(_files), against the
reference E3D_POST's output for the same inputs (tests/golden/post/, made by
tests/golden/make_post_golden.py): byte for byte, on the reference's own 4-partition post test
data, on a mixed-element partition with malformed .sol lines, and on a 1152-tet mesh with a
solver-written .sol file. Host only: no GPU is involved."""
from __future__ import annotations

import ctypes
import gzip
import re
import shutil
import subprocess
import sys
from pathlib import Path

import pytest
from conftest import GOLDEN, ROOT

sys.path.insert(0, str(GOLDEN))
from make_golden import CONFIG  # noqa: E402

POST = ROOT / "euler3d_amd" / "bin" / "E3D_POST"
CASES = {"block8": ("StructuredBlock_8_p#.par", 4), "mixed": ("mixed.par", 1), "bump_small": ("bump_small.par", 1)}


def stage(case: str, d: Path) -> Path:
    """Copy the case's inputs into d (gunzipped; bump_small's mesh from the solver fixtures) and
    write its .e3d; returns the configuration path."""
    src = GOLDEN / "post" / case
    for f in src.iterdir():
        if f.name.startswith("expected") or f.name == "reference_fixture.dat":
            continue
        if f.suffix == ".gz":
            (d / f.stem).write_bytes(gzip.decompress(f.read_bytes()))
        else:
            shutil.copyfile(f, d / f.name)
    if case == "bump_small":
        (d / "bump_small.par").write_bytes(gzip.decompress((GOLDEN / "bump_small.par.gz").read_bytes()))
    pattern, n = CASES[case]
    s = CONFIG.format(min_residual=0, min_aero=0, mesh="unused.su2", par=str(d / pattern), out=str(d), ocl=2, ocd=0, rx=0.0, ry=0.0, rz=0.0,
                      sref=1.0, cref=1.0, sampling=1, mach=0.5, aoa=0.0, cfl=1.0, iters=1, ti=2, smooth=0)
    cfg = d / "case.e3d"
    cfg.write_text(s.replace("PARTITION_FILES= 1", f"PARTITION_FILES= {n}"))
    return cfg


def expected(case: str) -> bytes:
    src = GOLDEN / "post" / case
    if (src / "expected.dat.gz").exists():
        return gzip.decompress((src / "expected.dat.gz").read_bytes())
    return (src / "expected.dat").read_bytes()


@pytest.mark.parametrize("case", sorted(CASES))
def test_post_matches_reference_bytes(case, tmp_path):
    cfg = stage(case, tmp_path)
    r = subprocess.run([str(POST), str(cfg)], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "TECPLOT" in r.stdout
    assert (tmp_path / "out.dat").read_bytes() == expected(case)


def test_post_matches_reference_checked_in_fixture(tmp_path):
    """The reference's own test/post/Output.dat was written by an older post-processor (lower-case
    variable names, blank-separated connectivity); every number in it must still match."""
    cfg = stage("block8", tmp_path)
    subprocess.run([str(POST), str(cfg)], cwd=str(tmp_path), check=True, capture_output=True, timeout=120)
    ours = (tmp_path / "out.dat").read_text().splitlines()
    ref = (GOLDEN / "post" / "block8" / "reference_fixture.dat").read_text().splitlines()
    assert len(ours) == len(ref) == 213
    assert ours[0].lower() == ref[0].lower()
    for a, b in zip(ours[1:], ref[1:]):
        assert re.split(r"\s+", a.strip()) == re.split(r"\s+", b.strip())


def test_post_library_entry_points(tmp_path):
    lib = ctypes.CDLL(str(ROOT / "euler3d_amd" / "lib" / "libe3d_hip.so"))
    lib.e3d_last_error.restype = ctypes.c_char_p
    cfg = stage("block8", tmp_path)
    assert lib.e3d_post_tecplot(str(cfg).encode()) == 0
    assert (tmp_path / "out.dat").read_bytes() == expected("block8")
    # explicit paths, partitions in another order: zones follow the given order
    order = [2, 0]
    par = (ctypes.c_char_p * 2)(*[str(tmp_path / f"StructuredBlock_8_p{i}.par").encode() for i in order])
    sol = (ctypes.c_char_p * 2)(*[str(tmp_path / f"StructuredBlock_8_p{i}.par.sol").encode() for i in order])
    out = tmp_path / "two.dat"
    assert lib.e3d_post_tecplot_files(str(out).encode(), 2, par, sol) == 0
    full = expected("block8").decode().splitlines(keepends=True)
    zone = 53  # lines per 2-hex partition zone: 3 header, 36 coordinates, 12 values, 2 elements
    want = full[0] + "".join("".join(full[1 + i * zone:1 + (i + 1) * zone]) for i in order)
    assert out.read_text() == want
    # a missing .sol file is an I/O error, not a crash
    (tmp_path / "StructuredBlock_8_p3.par.sol").unlink()
    assert lib.e3d_post_tecplot(str(cfg).encode()) == 2
    assert b"solution file" in lib.e3d_last_error()
    assert lib.e3d_post_tecplot(b"/nonexistent/case.e3d") == 1


def test_post_cli_errors(tmp_path):
    r = subprocess.run([str(POST)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Usage" in r.stderr
    cfg = stage("mixed", tmp_path)
    (tmp_path / "mixed.par").unlink()
    r = subprocess.run([str(POST), str(cfg)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "mesh file" in r.stderr
