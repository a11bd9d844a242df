"""The batched and served workloads at their own sizes, on the mixed-precision engine.

* 64 seeded uneven airfoil cases (n ~ U{65..257} nodes a side, seed 0) resident in one engine with
  the autotuned launch shape, 200 iterations: every case's CL, CD and density within 1e-4 of the
  FP64 path after the same iterations, and a sample of cases bitwise the same alone.
* a Poisson stream of eight solve requests from the three mesh families, 16k to 1.7M cells (seed 3),
  capped at 100 iterations: every served result bitwise what the case gives alone, and within
  1e-4 of the FP64 path in CL and CD.
"""
from __future__ import annotations

import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from euler3d_amd import Batch, EulerSolver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(cases):
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda c: c.build(), cases))


def fp64(mesh, prm, iters):
    with EulerSolver(mesh, prm) as s:
        s.iterate(iters)
        cf, cm = s.aero()
        rho = s.get_state()["rho"][: mesh.info()["n_elem"]]
    cl = cf[prm.orient_cl[0]] * prm.orient_cl[1]
    cd = cf[prm.orient_cd[0]] * prm.orient_cd[1]
    return cl, cd, rho


@pytest.mark.timeout(600)
def test_gpu_airfoil_batch_64_cases():
    from euler3d_amd.autotune import CANDIDATES, autotune
    from euler3d_amd.workloads import airfoil_batch

    iters = 200
    cases = airfoil_batch(64, seed=0)
    built = build(cases)
    cells = [m.info()["n_elem"] for m, _ in built]
    say(f"[gpu] airfoil batch: 64 cases, {sum(cells)} cells ({min(cells)}..{max(cells)})")
    tune = autotune(built, iters=2, warmup=1, use_cache=False)
    assert len(tune["table"]) == len(CANDIDATES) and tune["best"] in CANDIDATES
    assert all(r["ms_per_iteration"] > 0 for r in tune["table"])
    say(f"[gpu] autotuned launch shape {tune['best']}")
    w, x, nt = tune["best"]
    with Batch(max_cases=64, wg_per_cu=w, xcds=x, nontemporal=nt) as b:
        with ThreadPoolExecutor(8) as pool:
            ids = list(pool.map(lambda mp: b.add(mp[0], mp[1], drop=0.0, max_iter=iters), built))
        b.run(iters)
        st = [b.status(i) for i in ids]
        res = [b.result(i) for i in ids]
        rho = [b.get_state(i)["rho"][:n] for i, n in zip(ids, cells)]
    assert all(s["outcome"] == "capped" and s["iterations"] == iters for s in st)
    worst = [0.0, 0.0, 0.0]
    for k, (m, p) in enumerate(built):
        cl, cd, r64 = fp64(m, p, iters)
        d = (abs(res[k]["CL"] - cl), abs(res[k]["CD"] - cd), float(np.max(np.abs(rho[k] - r64))))
        worst = [max(a, b) for a, b in zip(worst, d)]
        assert d[0] <= 1e-4 and d[1] <= 1e-4 and d[2] <= 1e-4, (k, cells[k], d)
    say(f"[gpu] airfoil batch vs FP64: max |dCL| {worst[0]:.2e} |dCD| {worst[1]:.2e} Linf rho {worst[2]:.2e}")
    for k in (0, 31, 63):
        with Batch() as b:
            i = b.add(built[k][0], built[k][1], drop=0.0, max_iter=iters)
            b.run(iters)
            alone = b.result(i)
        assert np.array_equal(alone["rms"].view(np.uint64), res[k]["rms"].view(np.uint64)), k
        assert alone["CL"] == res[k]["CL"] and alone["CD"] == res[k]["CD"], k


@pytest.mark.timeout(600)
def test_gpu_served_stream_10k_to_2m_cells():
    from euler3d_amd.serve import Server, summarize
    from euler3d_amd.workloads import poisson_arrivals, request_mix

    iters = 100
    cases = request_mix(8, seed=3, min_cells=1e4, max_cells=2e6)
    built = build(cases)
    cells = [m.info()["n_elem"] for m, _ in built]
    say(f"[gpu] served stream: {len(cases)} requests, {min(cells)}..{max(cells)} cells, "
        f"kinds {[c.kind for c in cases]}")
    assert min(cells) < 2e4 and max(cells) > 1.5e6 and {c.kind for c in cases} == {"naca", "bump", "wing"}
    srv = Server(hbm_budget=200e9, max_cases=8, drop=1e-6, max_iter=iters, admit_workers=4, poll_ms=5.0)
    try:
        rec = srv.serve([(m, p, c.kind) for (m, p), c in zip(built, cases)], poisson_arrivals(8, 20.0, seed=0))
    finally:
        srv.close()
    s = summarize(rec)
    say(f"[gpu] served: p50 {s['p50_s']:.2f} s p99 {s['p99_s']:.2f} s, outcomes {[r['outcome'] for r in rec]}")
    assert s["failed"] == 0 and s["rejected"] == 0
    for k, ((m, p), r) in enumerate(zip(built, rec)):
        with Batch() as b:
            i = b.add(m, p, drop=1e-6, max_iter=iters)
            b.run(iters)
            alone, ast = b.result(i), b.status(i)
        assert r["iterations"] == ast["iterations"] and r["outcome"] == ast["outcome"], k
        assert r["CL"] == alone["CL"] and r["CD"] == alone["CD"], k
        cl, cd, _ = fp64(m, p, ast["iterations"])
        assert abs(r["CL"] - cl) <= 1e-4 and abs(r["CD"] - cd) <= 1e-4, (k, r["CL"], cl, r["CD"], cd)
