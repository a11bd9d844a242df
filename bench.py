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
restatement of the reference's arithmetic) timed on one RK5 iteration of the same mesh, with its
recorded speed ratio to the compiled reference (oracle/cpu_anchor.json) and its timings on the
NACA and wing meshes. other_scheme: the same FP64 workload with AUSM (the bump case runs both
schemes), with its own roofline. mixed_precision: the same workload on the FP32-flux / FP64-update
engine.
served: a bounded Poisson stream of six requests, served by continuous batching against the plain
FP64 path one request at a time.
"""
from __future__ import annotations

import argparse
import json
from concurrent.futures import ThreadPoolExecutor
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


def mixed_bytes_per_cell_stage(n_cells: int, n_faces: int) -> float:
    # FP32 state copy in (20) + W0 FP64 in (40) + W FP64 out (40) + FP32 copy out (20) + FP32 cell
    # metrics (16) + per face: FP32 area vector + two int32 ids (20)
    return 136.0 + 20.0 * n_faces / n_cells


def mixed_precision_leg(mesh, prm, n_cells: int, n_faces: int, steps: int, warmup: int) -> dict:
    """The same workload through the mixed-precision engine (FP32 flux/BC, FP64 update): wall-clock
    throughput with graph replay, stage-kernel mean from HIP events in a second, ungraphed pass."""
    from euler3d_amd import Batch

    out = {}
    with Batch() as b:
        cid = b.add(mesh, prm, drop=0.0, max_iter=1 << 30, hist_cap=4 * (steps + warmup))
        b.run(warmup)
        b.wait()
        t0 = time.perf_counter()
        b.run(steps)
        b.wait()
        wall = time.perf_counter() - t0
        out["value"] = n_cells * 5 * steps / wall
        out["ms_per_step"] = wall * 1e3 / steps
        it = b.status(cid)["iterations"]
    with Batch(use_graph=False) as b:
        b.add(mesh, prm, drop=0.0, max_iter=1 << 30, hist_cap=4 * (steps + warmup))
        b.run(warmup)
        b.set_timing(True)
        b.run(steps)
        kern_ms, launches, _ = b.timing()
    bpc = mixed_bytes_per_cell_stage(n_cells, n_faces)
    achieved = bpc * n_cells / (kern_ms * 1e-3) / 1e9
    out.update({"unit": "cell-stage/s", "dtype": "f32 flux / f64 update", "iterations": it,
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "kernel": "kMixStage", "kernel_ms": kern_ms,
                             "traffic": measured_traffic("kMixStage", n_cells, prm.scheme),
                             "bytes_per_cell_stage": bpc, "launches": launches}})
    return out


def other_scheme_leg(mesh, pd: dict, scheme: int, n_cells: int, n_faces: int, steps: int, warmup: int) -> dict:
    """The same FP64 workload with the other flux scheme (the 6.3M-tet bump runs both Roe and AUSM):
    wall-clock throughput and the stage kernel's roofline, timed as the main leg."""
    from euler3d_amd import EulerSolver, default_params

    prm = default_params(**{**pd, "scheme": scheme}, orient_cl=(1, 1), orient_cd=(0, 1), orient_cm=(2, 1))
    with EulerSolver(mesh, prm, device=0) as s:
        s.iterate(warmup)
        s.set_kernel_timing(True)
        t1 = time.perf_counter()
        s.iterate(steps)
        wall = time.perf_counter() - t1
        kern_ms, launches, kname = s.kernel_timing()
    bpc = algorithmic_bytes_per_cell_stage(n_cells, n_faces)
    achieved = bpc * n_cells / (kern_ms * 1e-3) / 1e9
    return {"scheme": "AUSM" if scheme else "Roe", "value": n_cells * 5 * steps / wall, "unit": "cell-stage/s",
            "ms_per_step": wall * 1e3 / steps, "dtype": "f64",
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kname, n_cells, scheme),
                         "kernel": kname, "bytes_per_cell_stage": bpc, "kernel_ms": kern_ms, "launches": launches}}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, params: dict, n_cells: int) -> dict:
    """The CPU oracle (C restatement of the reference's arithmetic, -O2) on the same mesh: one RK5
    iteration single-threaded (the reference runs one thread per MPI rank), then on all the host
    cores this process may use (parallel face fluxes and per-cell gathers, same bits)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_ctypes
    from oracle_ctypes import Oracle

    names = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit",
             "faceSurfaces", "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost",
             "farCell", "farFace", "symGhost", "symCell", "symFace")
    arrays = {k: mesh.get(k) for k in names}
    o = Oracle(arrays, params)
    t0 = time.perf_counter()
    o.iterate(1)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    oracle_ctypes.set_threads(cores)
    try:
        om = Oracle(arrays, params)
        t0 = time.perf_counter()
        om.iterate(2)
        dtm = (time.perf_counter() - t0) / 2
    finally:
        oracle_ctypes.set_threads(1)
    out = {"value": n_cells * 5 / dt, "unit": "cell-stage/s", "cores": 1, "kind": "port",
           "sample": f"1 RK5 iteration (5 stages) of the full {n_cells}-cell mesh, {dt:.1f} s, oracle C "
                     f"restatement compiled -O2, single thread, {cpu_model()}",
           "all_cores": {"value": n_cells * 5 / dtm, "unit": "cell-stage/s", "cores": cores, "kind": "port",
                         "sample": f"2 RK5 iterations, {dtm:.2f} s each, OpenMP over {cores} threads"}}
    anchor = ROOT / "oracle" / "cpu_anchor.json"
    if anchor.exists():
        # the port against the compiled reference, both single-core on one host (recorded where the
        # reference builds): the reference's speed on this host is estimated by scaling the port's
        rec = json.loads(anchor.read_text())
        ratio = rec["port_over_reference_time"]
        out["anchor_vs_reference"] = {
            "port_over_reference_time": ratio, "reference_equivalent_value": out["value"] * ratio,
            "unit": "cell-stage/s", "case": rec["case"], "host": rec["host"],
            "reference_s_per_iteration": rec["reference"]["s_per_iteration"],
            "port_s_per_iteration": rec["port"]["s_per_iteration"], "reference_build": rec["reference"]["build"]}
    out["other_workloads"] = [oracle_timing(*w) for w in other_cpu_workloads()]
    return out


def other_cpu_workloads():
    """The plain path's two other meshes for the single-thread oracle timing: the 129 x 129 x 2 NACA
    O-grid (16,384 hexes, 20 iterations) and the 193 x 97 x 65 wing (1,179,648 hexes, 1 iteration)."""
    from euler3d_amd import Mesh

    naca = lambda: Mesh.naca(ni=129, nj=129, nk=2, thickness=0.12, wall_spacing=1e-3,  # noqa: E731
                             farfield_radius=20.0, dz=0.1)
    wing = lambda: Mesh.wing(ni=193, nj=97, nk=65, thickness=0.10, sweep_deg=30.0, taper=0.56,  # noqa: E731
                             semispan=1.5, farfield=15.0)
    return [("NACA0012 O-grid 129x129x2, M 0.8, AoA 1.25 deg, Roe, RK5", naca, dict(mach=0.8, aoa_deg=1.25), 20),
            ("swept wing O-H 193x97x65, M 0.84, AoA 3 deg, Roe, RK5", wing, dict(mach=0.84, aoa_deg=3.0), 1)]


def oracle_timing(name, make, params, iters) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle_ctypes import Oracle

    m = make().preprocess(1.0)
    n = m.info()["n_elem"]
    names = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit",
             "faceSurfaces", "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost",
             "farCell", "farFace", "symGhost", "symCell", "symFace")
    o = Oracle({k: m.get(k) for k in names}, dict(params, cfl=1.5))
    t0 = time.perf_counter()
    o.iterate(iters)
    dt = time.perf_counter() - t0
    m.close()
    return {"workload": name, "cells": n, "iterations": iters, "value": n * 5 * iters / dt, "unit": "cell-stage/s",
            "cores": 1, "kind": "port", "s_per_iteration": dt / iters}


def build_meshes(cases, workers: int = 8):
    """Generate and preprocess the meshes of `cases` in parallel (the native calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda c: c.build(), cases))


def bench_batch(a) -> dict:
    """64 uneven airfoil cases: all resident in the mixed-precision engine (persistent launch,
    graph-replayed iterations, autotuned launch shape) versus the plain FP64 path one case at a
    time. Each case runs to a 1e-6 residual drop or the iteration cap."""
    from euler3d_amd import Batch
    from euler3d_amd.autotune import autotune
    from euler3d_amd.serve import plain_solve
    from euler3d_amd.workloads import airfoil_batch

    cases = airfoil_batch(a.cases, seed=0)
    t = time.perf_counter()
    built = build_meshes(cases)
    cells = [m.info()["n_elem"] for m, _ in built]
    log(f"batch: {len(cases)} cases, {sum(cells)} cells, meshes in {time.perf_counter() - t:.1f} s")
    tune = autotune(built, iters=3, warmup=1, use_cache=False)
    log(f"autotune: best {tune['best']} ({len(tune['table'])} shapes)")
    with Batch(max_cases=len(cases), wg_per_cu=tune["best"][0], xcds=tune["best"][1],
               nontemporal=tune["best"][2]) as b:
        t0 = time.perf_counter()
        # admission (plan and upload) runs on worker threads, as in the server
        with ThreadPoolExecutor(a.admit_workers) as pool:
            ids = list(pool.map(lambda mp: b.add(mp[0], mp[1], drop=a.drop, max_iter=a.max_iter), built))
        t_admit = time.perf_counter() - t0
        while True:
            b.run(100)
            st = b.poll()
            if all(s["done"] for s in st):
                break
        wall = time.perf_counter() - t0
        outcome = {s["id"]: s["outcome"] for s in b.poll()}
        res = [b.result(i) for i in ids]
        iters = [len(r["rms"]) for r in res]
    cell_stages = 5 * sum(c * i for c, i in zip(cells, iters))
    log(f"batched mixed: {wall:.2f} s ({t_admit:.2f} s admission), {cell_stages / wall:.3e} cell-stage/s")
    t0 = time.perf_counter()
    plain = [plain_solve(m, p, drop=a.drop, max_iter=a.max_iter) for m, p in built]
    wall_plain = time.perf_counter() - t0
    log(f"plain fp64 one at a time: {wall_plain:.2f} s")
    dcl = max(abs(r["CL"] - q["CL"]) for r, q in zip(res, plain))
    dcd = max(abs(r["CD"] - q["CD"]) for r, q in zip(res, plain))
    return {
        "metric": "batched solves/h", "value": 3600.0 * len(cases) / wall, "unit": "solves/h", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32 flux / f64 update", "data": "synthetic",
        "config": {"workload": f"{len(cases)} NACA 4-digit O-grid cases (n~U{{65..257}} nodes a side, t U[0.06,0.15], "
                               f"camber U[0,0.04], M U[0.3,0.85], AoA U[-2,4] deg, seed 0), Roe, RK5 CFL 1.5, "
                               f"to a {a.drop:g} residual drop or {a.max_iter} iterations",
                   "cells_total": sum(cells), "launch_shape": list(tune["best"])},
        "wall_s": wall, "admission_s": t_admit, "cell_stage_per_s": cell_stages / wall,
        "iterations": {"min": min(iters), "max": max(iters), "mean": sum(iters) / len(iters)},
        "outcomes": {k: sum(1 for i in ids if outcome[i] == k) for k in ("converged", "capped", "failed")},
        "plain_fp64": {"wall_s": wall_plain, "solves_per_h": 3600.0 * len(cases) / wall_plain,
                       "iterations_mean": sum(q["iterations"] for q in plain) / len(plain),
                       "outcomes": {k: sum(1 for q in plain if q["outcome"] == k)
                                    for k in ("converged", "capped", "failed")}},
        "speedup_vs_plain": wall_plain / wall,
        "max_abs_dCL_vs_fp64": dcl, "max_abs_dCD_vs_fp64": dcd,
        "autotune": tune["table"],
    }


def bench_serve(a) -> list[dict]:
    """Poisson request streams over a fixed seeded request set (three mesh families, 10k-2M
    cells): the continuous-batching server on the mixed-precision engine against the plain FP64
    path serving one request at a time in arrival order (its latencies follow from its measured
    service times). Meshes are generated and preprocessed before the clock starts."""
    from euler3d_amd.serve import Server, fifo_latencies, percentile, plain_solve, summarize
    from euler3d_amd.workloads import poisson_arrivals, request_mix

    cases = request_mix(a.requests, seed=0, max_cells=a.max_cells)
    t = time.perf_counter()
    built = build_meshes(cases)
    cells = [m.info()["n_elem"] for m, _ in built]
    log(f"serve: {len(cases)} requests, {sum(cells)} cells ({min(cells)}..{max(cells)}), meshes in "
        f"{time.perf_counter() - t:.1f} s")
    plain = []
    for i, (m, p) in enumerate(built):
        plain.append(plain_solve(m, p, drop=a.drop, max_iter=a.max_iter))
        log(f"plain {i}: {cases[i].kind} {cells[i]} cells, {plain[i]['service']:.2f} s, {plain[i]['iterations']} it")
    services = [q["service"] for q in plain]
    lines = []
    srv = Server(drop=a.drop, max_iter=a.max_iter, hbm_budget=200e9, admit_workers=a.admit_workers,
                 **({"stall": None} if a.no_stall else {}))
    try:
        for lam in a.lambdas:
            arr = poisson_arrivals(len(cases), lam, seed=0)
            rec = srv.serve([(m, p, cases[i].kind) for i, (m, p) in enumerate(built)], arr)
            sv = summarize(rec)
            lat, span = fifo_latencies(arr, services)
            pl = {"p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99), "makespan_s": span,
                  "solves_per_h": 3600.0 * len(cases) / span,
                  **{k: sum(1 for q in plain if q["outcome"] == k) for k in ("converged", "capped", "failed")}}
            dcl = max(abs(r["CL"] - q["CL"]) for r, q in zip(rec, plain))
            dcd = max(abs(r["CD"] - q["CD"]) for r, q in zip(rec, plain))
            line = {"metric": "served solves/h", "value": sv["solves_per_h"], "unit": "solves/h", "n_gpus": 1,
                    "higher_is_better": True, "dtype": "f32 flux / f64 update", "data": "synthetic",
                    "config": {"workload": f"{len(cases)} requests (naca/bump/wing, 10k-{a.max_cells:.0f} cells, "
                                           f"seed 0), Poisson rate {lam} req/s, {a.drop:g} drop or {a.max_iter} it, "
                                           f"200 GB HBM budget", "rate": lam, "cells_total": sum(cells),
                               "stall_test": None if a.no_stall else list(srv.stall)},
                    "served": sv, "plain_fp64": pl, "vs_plain_solves_per_h": sv["solves_per_h"] / pl["solves_per_h"],
                    "max_abs_dCL_vs_fp64": dcl, "max_abs_dCD_vs_fp64": dcd,
                    "requests": [{"kind": cases[i].kind, "cells": cells[i], "arrival": round(r["arrival"], 4),
                                  "admitted": round(r.get("admitted", float("nan")), 4),
                                  "finished": round(r["finished"], 4), "iterations": r.get("iterations"),
                                  "outcome": r.get("outcome"), "stalled_at": r.get("stalled_at"),
                                  "fp64_iterations": r.get("fp64_iterations"),
                                  "plain_iterations": plain[i]["iterations"],
                                  "plain_outcome": plain[i]["outcome"],
                                  "plain_service": round(plain[i]["service"], 4)} for i, r in enumerate(rec)]}
            log(f"rate {lam}: served p50 {sv['p50_s']:.2f} s p99 {sv['p99_s']:.2f} s {sv['solves_per_h']:.0f}/h; "
                f"plain p50 {pl['p50_s']:.2f} s p99 {pl['p99_s']:.2f} s {pl['solves_per_h']:.0f}/h")
            print(json.dumps(line), flush=True)
            lines.append(line)
    finally:
        srv.close()
    return lines


def served_leg(a) -> dict:
    """A bounded served stream under the same clock as the default run: six requests of the three
    mesh families (10k-133k cells, seed 0), Poisson arrivals at 20 req/s, each to a 1e-6 residual
    drop or 5000 iterations. The continuous-batching server on the mixed-precision engine against
    the plain FP64 path serving one request at a time in arrival order (its latencies follow from
    its measured service times); meshes are built before either clock starts."""
    from euler3d_amd.serve import Server, fifo_latencies, percentile, plain_solve, summarize
    from euler3d_amd.workloads import poisson_arrivals, request_mix

    cases = request_mix(6, seed=0, min_cells=1e4, max_cells=2e5)
    built = build_meshes(cases)
    cells = [m.info()["n_elem"] for m, _ in built]
    plain = [plain_solve(m, p, drop=1e-6, max_iter=5000) for m, p in built]
    arr = poisson_arrivals(len(cases), 20.0, seed=0)
    srv = Server(drop=1e-6, max_iter=5000, hbm_budget=200e9, admit_workers=4)
    try:
        rec = srv.serve([(m, p, c.kind) for (m, p), c in zip(built, cases)], arr)
    finally:
        srv.close()
    for m, _ in built:
        m.close()
    sv = summarize(rec)
    lat, span = fifo_latencies(arr, [q["service"] for q in plain])
    outcomes = lambda rs: {k: sum(1 for r in rs if r["outcome"] == k) for k in ("converged", "capped", "failed")}  # noqa: E731
    return {"metric": "served solves/h", "value": sv["solves_per_h"], "unit": "solves/h",
            "workload": f"6 requests (naca/bump/wing, {min(cells)}-{max(cells)} cells, seed 0), Poisson 20 req/s, "
                        "1e-6 drop or 5000 iterations, 200 GB HBM budget",
            "served": {"p50_s": sv["p50_s"], "p99_s": sv["p99_s"], "makespan_s": sv["makespan_s"],
                       "solves_per_h": sv["solves_per_h"], **outcomes(rec), "finished_fp64": sv["finished_fp64"]},
            "plain_fp64": {"p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99), "makespan_s": span,
                           "solves_per_h": 3600.0 * len(cases) / span, **outcomes(plain)},
            "vs_plain_solves_per_h": sv["solves_per_h"] / (3600.0 * len(cases) / span),
            "max_abs_dCL_vs_fp64": max(abs(r["CL"] - q["CL"]) for r, q in zip(rec, plain)),
            "max_abs_dCD_vs_fp64": max(abs(r["CD"] - q["CD"]) for r, q in zip(rec, plain)),
            "iterations": [[r["iterations"], q["iterations"]] for r, q in zip(rec, plain)]}


def bench_wing(a) -> dict:
    """The swept tapered wing of the accuracy case, 193x97x65 nodes (1,179,648 hexes), M 0.84,
    AoA 3 deg: the FP64 path and the mixed-precision engine (FP32 flux and BC, FP64 update with
    compensated accumulation) run the same RK5 iterations from freestream. Reported: CL, CD and the
    density of both, their differences, and each path's throughput."""
    import numpy as np

    from euler3d_amd import Batch, EulerSolver, Mesh, default_params

    t = time.perf_counter()
    m = Mesh.wing(ni=193, nj=97, nk=65, thickness=0.10, sweep_deg=30.0, taper=0.56, semispan=1.5, farfield=15.0)
    m.preprocess(1.0)
    n = m.info()["n_elem"]
    log(f"wing: {n} cells, mesh in {time.perf_counter() - t:.1f} s")
    iters = a.max_iter
    prm = default_params(mach=0.84, aoa_deg=3.0, gamma=1.4, cfl=1.5, sref=1.0, cref=1.0, max_iter=iters,
                         orient_cl=(1, 1), orient_cd=(0, 1), orient_cm=(2, 1))
    n_faces = m.info()["n_face"]
    with EulerSolver(m, prm) as s:
        s.set_kernel_timing(True)
        t0 = time.perf_counter()
        r64 = s.run()
        wall64 = time.perf_counter() - t0
        kern_ms, launches, kname = s.kernel_timing()
        rho64 = s.get_state()["rho"][:n]
    bpc = algorithmic_bytes_per_cell_stage(n, n_faces)
    achieved = bpc * n / (kern_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "kernel": kname, "kernel_ms": kern_ms, "launches": launches,
            "bytes_per_cell_stage": bpc, "traffic": measured_traffic(kname, n, 0)}
    log(f"fp64: {r64['iterations']} iterations in {wall64:.2f} s, CL {r64['CL']:.9f} CD {r64['CD']:.9f}; "
        f"{kname} {kern_ms * 1e3:.1f} us, {roof['frac']:.3f} of HBM peak")
    with Batch(max_cases=1) as b:
        cid = b.add(m, prm, drop=0.0, max_iter=iters)
        t0 = time.perf_counter()
        while not b.poll()[0]["done"]:
            b.run(200)
        wall32 = time.perf_counter() - t0
        res = b.result(cid)
        rho32 = b.get_state(cid)["rho"][:n]
    log(f"mixed: {len(res['rms'])} iterations in {wall32:.2f} s, CL {res['CL']:.9f} CD {res['CD']:.9f}")
    rows = r64["rows"]
    h32 = res["rms"]
    return {"metric": "mixed-precision error vs FP64 (wing)", "value": abs(res["CL"] - r64["CL"]), "unit": "|dCL|",
            "n_gpus": 1, "higher_is_better": False, "dtype": "f32 flux / f64 update vs f64", "data": "synthetic",
            "config": {"workload": "swept tapered wing, 10% symmetric 4-digit section, LE sweep 30 deg, taper 0.56, "
                                   "semispan 1.5 root chords, O-H grid 193x97x65 nodes, farfield 15 root chords, "
                                   f"M 0.84, AoA 3 deg, Roe, RK5 CFL 1.5, {iters} iterations from freestream",
                       "cells": n},
            "fp64": {"CL": r64["CL"], "CD": r64["CD"], "iterations": r64["iterations"], "rms_first": float(rows[0, 1]),
                     "rms_final": float(rows[-1, 1]), "rms_floor": float(rows[:, 1].min()),
                     "wall_s": wall64, "cell_stage_per_s": 5 * n * r64["iterations"] / wall64, "roofline": roof},
            "mixed": {"CL": res["CL"], "CD": res["CD"], "iterations": len(h32), "rms_first": float(h32[0]),
                      "rms_final": float(h32[-1]), "rms_floor": float(h32.min()),
                      "wall_s": wall32, "cell_stage_per_s": 5 * n * len(res["rms"]) / wall32},
            "abs_dCL": abs(res["CL"] - r64["CL"]), "abs_dCD": abs(res["CD"] - r64["CD"]),
            "linf_drho": float(np.max(np.abs(rho32 - rho64)))}


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
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-other-scheme", action="store_true",
                    help="skip the FP64 leg with the other flux scheme (AUSM when --scheme 0)")
    ap.add_argument("--no-serve", action="store_true", help="skip the bounded served leg")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run one plain device-to-device copy of this many MiB (a known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--mode", choices=("plain", "batch", "serve", "wing"), default="plain",
                    help="plain: the FP64 stage benchmark (default); batch: 64 uneven airfoil cases; "
                         "serve: Poisson request streams; wing: mixed precision against FP64 on the wing")
    ap.add_argument("--cases", type=int, default=64)
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--max-cells", type=float, default=2e6)
    ap.add_argument("--lambdas", type=float, nargs="+", default=[1, 2, 5, 10, 20, 50])
    ap.add_argument("--drop", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=5000)
    ap.add_argument("--admit-workers", type=int, default=8)
    ap.add_argument("--no-stall", action="store_true",
                    help="serve without the stall test (stalled cases are not finished on the FP64 path)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs one resident mesh per GPU and measures one GPU only; "
              f"--gpus {a.gpus} is not supported (multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if a.copy_calib_mib:
        import ctypes

        from euler3d_amd import native

        ms = ctypes.c_double()
        native.check(native.lib().e3d_copy_probe(0, a.copy_calib_mib << 20, ctypes.byref(ms)))
        log(f"copy calibration: {a.copy_calib_mib} MiB read + {a.copy_calib_mib} MiB written in {ms.value:.3f} ms")

    if a.mode == "batch":
        print(json.dumps(bench_batch(a)), flush=True)
        return
    if a.mode == "serve":
        bench_serve(a)
        return
    if a.mode == "wing":
        print(json.dumps(bench_wing(a)), flush=True)
        return

    from euler3d_amd import EulerSolver, Mesh, default_params

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
    if not a.no_other_scheme:
        out["other_scheme"] = other_scheme_leg(mesh, pd, 1 - a.scheme, n_cells, n_faces, a.steps, a.warmup)
        log(f"{out['other_scheme']['scheme']}: {out['other_scheme']['value']:.3e} cell-stage/s, "
            f"{out['other_scheme']['roofline']['kernel']} {out['other_scheme']['roofline']['kernel_ms'] * 1e3:.1f} us, "
            f"{out['other_scheme']['roofline']['frac']:.3f} of HBM peak")
    if not a.no_mixed:
        out["mixed_precision"] = mixed_precision_leg(mesh, prm, n_cells, n_faces, a.steps, a.warmup)
        log(f"mixed precision: {out['mixed_precision']['value']:.3e} cell-stage/s, kMixStage "
            f"{out['mixed_precision']['roofline']['kernel_ms'] * 1e3:.1f} us")
    if not a.no_serve:
        out["served"] = served_leg(a)
        log(f"served leg: {out['served']['value']:.0f} solves/h ({out['served']['vs_plain_solves_per_h']:.2f}x plain), "
            f"p50 {out['served']['served']['p50_s']:.2f} s p99 {out['served']['served']['p99_s']:.2f} s")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, pd, n_cells)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} cell-stage/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
