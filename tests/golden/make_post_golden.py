"""Writes the Tecplot post-processor fixtures of tests/golden/post/ from the reference E3D_POST.

Runs only where the reference tree is present (it is never read at test time): builds
oracle/_ref/E3D_POST from the reference sources (oracle/ref.mk), lays out each case's partition
meshes and .sol files with an .e3d configuration, runs the reference post-processor and stores
its Tecplot file as <case>/expected.dat(.gz) next to the inputs.

Cases:
  block8     the reference's own post test data (test/post/StructuredBlock_8_p0..3.par and their
             .sol files): four 2-hex partitions. The reference post reads <partition>.par.sol, so
             the .sol files are stored under that name. The reference's checked-in Tecplot file
             (test/post/Output.dat, written by an older post with lower-case variable names and
             blank-separated connectivity) is kept as reference_fixture.dat.
  mixed      one partition of this script's own: a tet, a pyramid, a prism, a hex, a voxel and an
             element of an unknown VTK id (skipped, but counted in NELEM), with a .sol file that
             has a short line, "nan"/"inf" and hex tokens and ends before its last cell
  bump_small the bump_small test mesh (1152 tets) with the reference solver's final state of the
             bump_small fixture written as the solver writes .sol files

Usage: python tests/golden/make_post_golden.py
"""
from __future__ import annotations

import gzip
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = Path("/root/reference")
POST = ROOT / "oracle" / "_ref" / "E3D_POST"
OUT = HERE / "post"
sys.path.insert(0, str(HERE))
from make_golden import CONFIG  # noqa: E402

MIXED_PAR = """%
% Problem dimension
%
NDIME= 3
%
% Inner element connectivity
%
NELEM= 6
10 0 1 2 4 0
14 0 1 3 2 5 1
13 1 3 6 5 7 8 2
12 3 6 9 10 7 8 11 12 3
42 1 2 3 4
11 0 1 2 3 4 5 6 7 5
%
% Node coordinates
%
NPOIN= 13
0.0 0.0 0.0 0
1.0 0.0 0.0 1
0.0 1.0 0.0 2
1.0 1.0 0.0 3
0.25 0.25 -1.0 4
0.5 0.5 1.0 5
2.0 1.0 0.0 6
1.0 1.0 1.0 7
2.0 1.0 1.0 8
2.0 2.0 0.0 9
1.0 2.0 0.0 10
2.0 2.0 1.0 11
1.0 2.0 1.0 12
%
% Boundary elements
%
NMARK= 1
MARKER_TAG= wall
MARKER_ELEMS= 1
5 0 1 2
%
% Internal Boundary elements
%
NINTERNAL_MARKER= 0
"""

MIXED_SOL = ("             density          velocity_x          velocity_y          velocity_z            pressure"
             "              energy\n"
             "        1.023140e+00        2.352612e-01       -2.122000e-02        3.215000e-03        1.255320e+00"
             "        3.285112e+00\n"
             "  1.5 -2.25e-1\n"
             "  0.75 nan 0.5 0.25 inf 2.0\n"
             "  9.5e-1 0x1p3 4.0 5.0 6.0 7.0\n"
             "  1.25 -0.5 +.5 1e-300 3.0 4.5\n")


def config(d: Path, pattern: str, nparts: int) -> Path:
    s = CONFIG.format(min_residual=0, min_aero=0, mesh="unused.su2", par=pattern, out=str(d), ocl=2, ocd=0, rx=0.0, ry=0.0, rz=0.0, sref=1.0,
                      cref=1.0, sampling=1, mach=0.5, aoa=0.0, cfl=1.0, iters=1, ti=2, smooth=0)
    s = s.replace("PARTITION_FILES= 1", f"PARTITION_FILES= {nparts}")
    p = d / "case.e3d"
    p.write_text(s)
    return p


def run_post(inputs: dict[str, bytes], pattern: str, nparts: int) -> bytes:
    with tempfile.TemporaryDirectory() as t:
        d = Path(t)
        for name, data in inputs.items():
            (d / name).write_bytes(data)
        cfg = config(d, str(d / pattern), nparts)
        subprocess.run([str(POST), str(cfg)], check=True, cwd=t, capture_output=True)
        return (d / "out.dat").read_bytes()


def solver_sol(g) -> bytes:
    """The reference solver's .sol layout (src/solver/WriteSolution.cpp:30-55)."""
    n = int(g["mesh_counts"][0])
    head = "".join(f"{h:>20}" for h in ("density", "velocity_x", "velocity_y", "velocity_z", "pressure", "energy"))
    rows = [head]
    for i in range(n):
        rows.append("".join(f"{g['run_' + k][i]:20.6e}" for k in ("rho", "u", "v", "w", "p", "E")))
    return ("\n".join(rows) + "\n").encode()


def main() -> None:
    subprocess.run(["make", "-s", "-f", "oracle/ref.mk", "-j8", "oracle/_ref/E3D_POST"], check=True, cwd=ROOT)
    OUT.mkdir(exist_ok=True)

    d = OUT / "block8"
    d.mkdir(exist_ok=True)
    inputs = {}
    for i in range(4):
        inputs[f"StructuredBlock_8_p{i}.par"] = (REF / f"test/post/StructuredBlock_8_p{i}.par").read_bytes()
        inputs[f"StructuredBlock_8_p{i}.par.sol"] = (REF / f"test/post/StructuredBlock_8_p{i}.sol").read_bytes()
    for k, v in inputs.items():
        (d / k).write_bytes(v)
    shutil.copyfile(REF / "test/post/Output.dat", d / "reference_fixture.dat")
    (d / "expected.dat").write_bytes(run_post(inputs, "StructuredBlock_8_p#.par", 4))

    d = OUT / "mixed"
    d.mkdir(exist_ok=True)
    inputs = {"mixed.par": MIXED_PAR.encode(), "mixed.par.sol": MIXED_SOL.encode()}
    for k, v in inputs.items():
        (d / k).write_bytes(v)
    (d / "expected.dat").write_bytes(run_post(inputs, "mixed.par", 1))

    d = OUT / "bump_small"
    d.mkdir(exist_ok=True)
    with np.load(HERE / "bump_small.npz") as z:
        g = {k: z[k] for k in z.files}
    inputs = {"bump_small.par": gzip.decompress((HERE / "bump_small.par.gz").read_bytes()),
              "bump_small.par.sol": solver_sol(g)}
    (d / "bump_small.par.sol.gz").write_bytes(gzip.compress(inputs["bump_small.par.sol"], mtime=0))
    (d / "expected.dat.gz").write_bytes(gzip.compress(run_post(inputs, "bump_small.par", 1), mtime=0))
    print("wrote", sorted(str(p.relative_to(HERE)) for p in OUT.rglob("*") if p.is_file()))


if __name__ == "__main__":
    main()
