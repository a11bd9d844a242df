"""Benchmark of the explicit-solver hot path on one MI355X.

Workload: 3L x L x L bump channel, 256 x 64 x 64 hexes split into 6 tets each (6,291,456 cells,
12,656,640 faces), seeded node jitter and random cell/node numbering (seed 0), M = 0.5, Roe flux,
5-stage RK with local time stepping at CFL 1.5, FP64. A step is one RK5 iteration (5 stages, each:
boundary ghosts, face fluxes, residual assembly, local time step, state update; plus the RMS of
the last stage). Mesh generation, preprocessing and upload are outside the timed region; the
state is resident in HBM when timing starts.

Prints one JSON line. value = cell-stage updates per second = cells * 5 * steps / wall time.
roofline: algorithmic bytes of one stage (152 + 32 * faces/cells B per cell-stage, FP64) divided
by the mean duration of the stage kernel(s), measured with HIP events on the solver's stream
during the timed steps, against 8 TB/s. cpu_baseline: the CPU oracle (single-thread C
restatement of the reference's arithmetic) timed on one RK5 iteration of the same mesh.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_cell_stage(n_cells: int, n_faces: int) -> float:
    # W_{k-1} in (40) + W0 in (40) + W_k out (40) + V, sum|Sx|, sum|Sy|, sum|Sz| (32)
    # + per face: area vector 3 x f64 + owner/neighbour 2 x int32 (32)
    return 152.0 + 32.0 * n_faces / n_cells


def measured_traffic(kernel: str, n_cells: int, scheme: int):
    """HBM bytes per launch of the stage kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs (profiles/traffic.json), each counter scaled by what it reads for a plain copy of known
    size; None when no profile of this kernel and workload is recorded."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    rec = json.loads(path.read_text()).get(f"{kernel}/{n_cells}/{'ausm' if scheme else 'roe'}")
    if not rec:
        return None
    return rec["fetch_kib"] * rec["fetch_scale"] * 1024 + rec["write_kib"] * rec["write_scale"] * 1024


def cpu_baseline(mesh, params: dict, n_cells: int) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle_ctypes import Oracle

    names = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit",
             "faceSurfaces", "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost",
             "farCell", "farFace", "symGhost", "symCell", "symFace")
    o = Oracle({k: mesh.get(k) for k in names}, params)
    t0 = time.perf_counter()
    o.iterate(1)
    dt = time.perf_counter() - t0
    return {"value": n_cells * 5 / dt, "unit": "cell-stage/s", "cores": 1, "kind": "port",
            "sample": f"1 RK5 iteration (5 stages) of the full {n_cells}-cell mesh, {dt:.1f} s, oracle C "
                      f"restatement compiled -O2, single thread"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--scheme", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run one plain device-to-device copy of this many MiB (a known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs one resident mesh per GPU and measures one GPU only; "
              f"--gpus {a.gpus} is not supported (multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    from euler3d_amd import EulerSolver, Mesh, default_params

    if a.copy_calib_mib:
        import torch

        src = torch.ones(a.copy_calib_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {a.copy_calib_mib} MiB read + {a.copy_calib_mib} MiB written")
        del src, dst

    t0 = time.perf_counter()
    mesh = Mesh.bump(nx=a.nx, ny=a.ny, nz=a.nz, jitter=0.1, seed=0, permute=True)
    mesh.preprocess(1.0)
    info = mesh.info()
    n_cells, n_faces = info["n_elem"], info["n_face"]
    log(f"mesh: {n_cells} cells, {n_faces} faces, setup {time.perf_counter() - t0:.1f} s")

    pd = dict(mach=0.5, aoa_deg=0.0, gamma=1.4, cfl=1.5, sref=3.0, cref=1.0, scheme=a.scheme)
    prm = default_params(**pd, orient_cl=(1, 1), orient_cd=(0, 1), orient_cm=(2, 1))
    solver = EulerSolver(mesh, prm, device=0)
    solver.iterate(a.warmup)
    solver.set_kernel_timing(True)
    t1 = time.perf_counter()
    sums = solver.iterate(a.steps)  # synchronises its stream before returning
    wall = time.perf_counter() - t1
    kern_ms, launches, kname = solver.kernel_timing()
    solver.set_kernel_timing(False)
    rms = (sums[-1] / n_cells) ** 0.5
    log(f"timed {a.steps} steps in {wall * 1e3:.2f} ms; stage kernel {kname} mean {kern_ms * 1e3:.1f} us "
        f"over {launches} launches; last rms {rms:.4e}")

    bpc = algorithmic_bytes_per_cell_stage(n_cells, n_faces)
    achieved = bpc * n_cells / (kern_ms * 1e-3) / 1e9
    value = n_cells * 5 * a.steps / wall
    out = {
        "metric": "cell-stage updates/s",
        "value": value,
        "unit": "cell-stage/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"bump channel {a.nx}x{a.ny}x{a.nz} hexes x 6 tets ({n_cells} cells, {n_faces} "
                               f"faces), jitter 0.1 + permuted numbering seed 0, M 0.5, "
                               f"{'AUSM' if a.scheme else 'Roe'}, RK5 CFL 1.5, FP64",
                   "cells": n_cells, "faces": n_faces, "stages_per_step": 5},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kname, n_cells, a.scheme),
                     "kernel": kname, "bytes_per_cell_stage": bpc, "kernel_ms": kern_ms},
        "cpu_baseline": None,
    }
    solver.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, pd, n_cells)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} cell-stage/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
