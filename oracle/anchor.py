"""Anchor of the CPU baseline to the reference itself. TEST INFRASTRUCTURE ONLY.

The GPU box cannot run the reference solver (its sources stay here), so the benchmark's CPU leg
times the oracle, the C restatement of the reference's arithmetic. This script, run where the
reference tree and its build (oracle/ref.mk -> oracle/_ref/E3D_Solver) exist, times both on the
same case on one core of this container and writes the ratio to oracle/cpu_anchor.json, which the
benchmark reports beside its own oracle timing.

Case: the 129 x 129 x 2 NACA0012 O-grid (16,384 hexes) of the plain-path workload, M 0.8, AoA
1.25 deg, Roe, RK5 at CFL 1.5, from freestream, N iterations with no coefficient sample. The
reference's figure is SolverLoopTime_ms of its summary.txt (EulerSolver::Run only, src/Solver.cpp
:58-66); the oracle's is the same loop (or_run_any) timed around the call. Both single-threaded;
the reference is the MPICH build of oracle/ref.mk (g++ -O2 -ffp-contract=off) under mpiexec -np 1.

Usage: python oracle/anchor.py [iterations]
"""
from __future__ import annotations

import json
import os
import platform
import subprocess
import sys
import tempfile
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(ROOT / "tests" / "golden"))

REF_EXE = HERE / "_ref" / "E3D_Solver"
MPIEXEC = Path("/opt/conda/bin/mpiexec")
NAMES = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit", "faceSurfaces",
         "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost", "farCell", "farFace",
         "symGhost", "symCell", "symFace")


def cpu_model() -> str:
    for line in open("/proc/cpuinfo"):
        if line.startswith("model name"):
            return line.split(":", 1)[1].strip()
    return platform.processor() or "unknown"


def main(iters: int) -> dict:
    from make_golden import CONFIG  # the .e3d template of the golden fixtures

    from euler3d_amd import Mesh
    from oracle_ctypes import Oracle

    if not REF_EXE.exists():
        raise SystemExit(f"{REF_EXE} missing: build it with make -f oracle/ref.mk")
    mesh = Mesh.naca(ni=129, nj=129, nk=2, thickness=0.12, wall_spacing=1e-3, farfield_radius=20.0, dz=0.1)
    mesh.preprocess(1.0)
    n = mesh.info()["n_elem"]
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "out").mkdir()
        par = td / "mesh.par"
        mesh.write(str(par), precision=16, par=True)
        cfg = td / "case.e3d"
        cfg.write_text(CONFIG.format(mesh=par.name, par=str(par), out=str(td / "out"), ocl=2, ocd=0, rx=0.25, ry=0.0,
                                     rz=0.0, sref=0.1, cref=1.0, sampling=iters + 1, mach=0.8, aoa=1.25, cfl=1.5,
                                     iters=iters, ti=2, smooth=0, min_residual=0, min_aero=0))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        t0 = time.perf_counter()
        r = subprocess.run([str(MPIEXEC), "-np", "1", str(REF_EXE), str(cfg)], cwd=str(td), capture_output=True,
                           text=True, env=env)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(r.stdout[-2000:] + r.stderr[-2000:])
        summary = (td / "out" / "summary.txt").read_text()
    ref_loop_ms = float(next(l.split("=")[1] for l in summary.splitlines() if l.strip().startswith("SolverLoopTime_ms")))
    o = Oracle({k: mesh.get(k) for k in NAMES}, dict(mach=0.8, aoa_deg=1.25, cfl=1.5, sref=0.1, max_iter=iters,
                                                     sampling=iters + 1))
    t0 = time.perf_counter()
    res = o.run()
    port_s = time.perf_counter() - t0
    assert res["iterations"] == iters
    ref_s = ref_loop_ms / 1e3
    return {"case": f"NACA0012 O-grid 129x129x2 nodes ({n} hexes), M 0.8, AoA 1.25 deg, Roe, RK5 CFL 1.5, "
                    f"{iters} iterations from freestream, no coefficient sample",
            "cells": n, "iterations": iters, "host": cpu_model(), "cores": 1,
            "reference": {"s_per_iteration": ref_s / iters, "cell_stage_per_s": n * 5 * iters / ref_s,
                          "loop_s": ref_s, "process_wall_s": wall,
                          "build": "oracle/ref.mk: g++ -O2 -ffp-contract=off, MPICH, mpiexec -np 1"},
            "port": {"s_per_iteration": port_s / iters, "cell_stage_per_s": n * 5 * iters / port_s, "loop_s": port_s,
                     "build": "gcc -O2 -ffp-contract=off, single thread"},
            "port_over_reference_time": port_s / ref_s}


if __name__ == "__main__":
    out = main(int(sys.argv[1]) if len(sys.argv) > 1 else 200)
    (HERE / "cpu_anchor.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))
