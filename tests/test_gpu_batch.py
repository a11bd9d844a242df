"""Mixed-precision multi-case engine (e3d_batch_*) on the GPU.

Accuracy bar (FP32 flux/BC in perturbation form, FP64 update): against the FP64 path, which is
itself bit-exact with the reference, after the same number of RK5 iterations from freestream,
CL and CD agree within 1e-4 and density within 1e-4 (L-infinity); the measured errors are printed.
Engine bar: a case's result is bitwise the same alone and in a mix, with and without graph
replay and for every launch shape; finished cases stop advancing; admission respects slots and
the HBM budget.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from euler3d_amd import Batch, E3DError, EulerSolver, Mesh, default_params
from euler3d_amd.native import E3D_ERR_BUSY

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b) -> bool:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def naca(ni=65, nj=33, thickness=0.12, camber=0.0):
    return Mesh.naca(ni=ni, nj=nj, nk=2, thickness=thickness, camber=camber, camber_pos=0.4 if camber else 0.0,
                     wall_spacing=2e-3, farfield_radius=20.0, dz=0.1).preprocess(1.0)


def naca_params(mach=0.8, aoa=1.25, scheme=0, **kw):
    return default_params(mach=mach, aoa_deg=aoa, gamma=1.4, cfl=1.5, sref=0.1, cref=1.0, scheme=scheme,
                          orient_cl=(1, 1), orient_cd=(0, 1), orient_cm=(2, 1), **kw)


def fp64_reference(mesh, prm, iters):
    with EulerSolver(mesh, prm) as s:
        sums = s.iterate(iters)
        cf, cm = s.aero()
        st = s.get_state()
    n = mesh.info()["n_elem"]
    cl = cf[prm.orient_cl[0]] * prm.orient_cl[1]
    cd = cf[prm.orient_cd[0]] * prm.orient_cd[1]
    return {"rms": np.sqrt(sums / n), "CL": cl, "CD": cd, "state": st}


def mixed_run(mesh, prm, iters, **bopts):
    with Batch(**bopts) as b:
        cid = b.add(mesh, prm, drop=0.0, max_iter=iters)
        b.run(iters)
        st = b.status(cid)
        r = b.result(cid)
        r["state"] = b.get_state(cid)
        r["iterations"] = st["iterations"]
    return r


def compare(tag, mesh, ref, mix, tol_coeff=1e-4, tol_rho=1e-4):
    n = mesh.info()["n_elem"]
    drho = np.max(np.abs(ref["state"]["rho"][:n] - mix["state"]["rho"][:n]))
    dcl, dcd = abs(ref["CL"] - mix["CL"]), abs(ref["CD"] - mix["CD"])
    drms = np.max(np.abs(ref["rms"] - mix["rms"]) / ref["rms"])
    say(f"[gpu] {tag}: |dCL|={dcl:.3e} |dCD|={dcd:.3e} Linf(rho)={drho:.3e} max rel d(rms)={drms:.3e} "
        f"CL={ref['CL']:.6f} CD={ref['CD']:.6f}")
    assert dcl <= tol_coeff and dcd <= tol_coeff, (dcl, dcd)
    assert drho <= tol_rho, drho


@pytest.mark.parametrize("scheme", [0, 1])
def test_mixed_close_to_fp64_naca(scheme):
    say(f"[gpu] mixed vs fp64 naca 65x33 scheme={scheme}")
    m = naca()
    prm = naca_params(scheme=scheme)
    it = 400
    ref = fp64_reference(m, prm, it)
    mix = mixed_run(m, prm, it)
    assert mix["iterations"] == it and len(mix["rms"]) == it
    compare(f"naca scheme={scheme}", m, ref, mix)


def test_mixed_close_to_fp64_wing_and_bump():
    say("[gpu] mixed vs fp64 wing 33x17x13 and bump 16x4x4")
    w = Mesh.wing(ni=33, nj=17, nk=13).preprocess(1.0)
    pw = default_params(mach=0.84, aoa_deg=3.0, gamma=1.4, cfl=1.5, sref=1.0, cref=1.0, orient_cl=(1, 1),
                        orient_cd=(0, 1), orient_cm=(2, 1))
    compare("wing", w, fp64_reference(w, pw, 200), mixed_run(w, pw, 200))
    bm = Mesh.bump(nx=16, ny=4, nz=4, jitter=0.1, seed=0, permute=True).preprocess(1.0)
    pb = default_params(mach=0.5, aoa_deg=0.0, gamma=1.4, cfl=1.5, sref=3.0, cref=1.0, orient_cl=(1, 1),
                        orient_cd=(0, 1), orient_cm=(2, 1))
    compare("bump", bm, fp64_reference(bm, pb, 200), mixed_run(bm, pb, 200))


def test_mixed_close_to_fp64_hybrid():
    """Cells of 4, 5 and 6 faces (tetrahedra, pyramids, wedges, hexahedra) in the same blocks."""
    from conftest import hybrid_mesh

    say("[gpu] mixed vs fp64 hybrid 16x12 columns")
    m = hybrid_mesh(16, 12)
    prm = default_params(mach=0.5, aoa_deg=2.0, gamma=1.4, cfl=1.5, sref=192.0, cref=1.0, orient_cl=(2, 1),
                         orient_cd=(0, 1), orient_cm=(1, 1))
    for scheme in (0, 1):
        prm.scheme = scheme
        compare(f"hybrid scheme={scheme}", m, fp64_reference(m, prm, 200), mixed_run(m, prm, 200))


def test_uniform_flow_is_preserved_exactly():
    say("[gpu] freestream box stays freestream")
    m = Mesh.box(4, 3, 2, tags=("farfield",) * 6).preprocess(1.0)
    prm = naca_params(mach=0.6, aoa=2.0)
    r = mixed_run(m, prm, 20)
    n = m.info()["n_elem"]
    assert np.all(r["state"]["rho"][:n] == 1.0)
    assert np.all(r["rms"] == 0.0)


def test_alone_vs_mix_is_bitwise_identical():
    say("[gpu] alone vs in a mix")
    a_mesh, b_mesh, c_mesh = naca(), naca(49, 25, thickness=0.09, camber=0.02), naca(81, 41, thickness=0.15)
    pa, pb, pc = naca_params(), naca_params(mach=0.5, aoa=3.0), naca_params(mach=0.7, aoa=-1.0, scheme=1)
    alone = mixed_run(a_mesh, pa, 150)
    with Batch() as b:
        ib = b.add(b_mesh, pb, drop=0.0, max_iter=1000)
        b.run(40)
        ia = b.add(a_mesh, pa, drop=0.0, max_iter=150)
        b.run(30)
        ic = b.add(c_mesh, pc, drop=0.0, max_iter=1000)
        b.run(200)  # a finishes after 150 of its own iterations and stops
        sa = b.status(ia)
        ra = b.result(ia)
        st = b.get_state(ia)
        b.remove(ib)
        b.run(5)
        assert b.status(ic)["iterations"] == 205
    assert sa["done"] and sa["iterations"] == 150
    assert bits_equal(ra["rms"], alone["rms"])
    for k in ("rho", "u", "v", "w", "p", "E", "H", "M"):
        assert bits_equal(st[k], alone["state"][k]), k
    assert ra["CL"] == alone["CL"] and ra["CD"] == alone["CD"]


def test_graph_and_launch_shape_do_not_change_results():
    say("[gpu] graph replay and launch shapes")
    m = naca()
    prm = naca_params()
    base = mixed_run(m, prm, 60)
    for opts in (dict(use_graph=False), dict(wg_per_cu=1, xcds=1), dict(wg_per_cu=7, nontemporal=False)):
        r = mixed_run(m, prm, 60, **opts)
        assert bits_equal(r["rms"], base["rms"]), opts
        assert bits_equal(r["state"]["rho"], base["state"]["rho"]), opts


@pytest.mark.parametrize("classes", ["merge", "split"])
def test_merged_and_split_launch_classes_same_bits(classes, monkeypatch):
    # a hexahedral wing (large blocks: its own launch class when split), a tetrahedral bump and an
    # O-grid, each alone and all three in one engine with one launch per stage or one per class
    say(f"[gpu] launch classes {classes}")
    meshes = [Mesh.wing(ni=25, nj=13, nk=9).preprocess(1.0), Mesh.bump(8, 2, 2, seed=3).preprocess(1.0), naca()]
    prms = [naca_params(mach=0.7, aoa=2.0), naca_params(mach=0.5, aoa=0.0), naca_params()]
    iters = 25
    alone = [mixed_run(m, p, iters) for m, p in zip(meshes, prms)]
    monkeypatch.setenv("E3D_MIX_CLASSES", classes)
    with Batch() as b:
        ids = [b.add(m, p, drop=0.0, max_iter=iters) for m, p in zip(meshes, prms)]
        b.run(iters)
        for i, a in zip(ids, alone):
            assert b.status(i)["iterations"] == iters
            assert bits_equal(b.result(i)["rms"], a["rms"])
            st = b.get_state(i)
            for k in ("rho", "u", "v", "w", "p", "E", "H", "M"):
                assert bits_equal(st[k], a["state"][k]), k


def test_converged_case_stops_advancing():
    say("[gpu] convergence stop")
    m = naca()
    with Batch() as b:
        cid = b.add(m, naca_params(), drop=0.5, max_iter=5000)
        b.run(300)
        s1 = b.status(cid)
        st1 = b.get_state(cid)
        b.run(20)
        s2 = b.status(cid)
        st2 = b.get_state(cid)
    assert s1["done"] and s1["iterations"] < 300
    assert s1["rms"] < 0.5 * s1["rms0"]
    assert s2["iterations"] == s1["iterations"]
    assert bits_equal(st1["rho"], st2["rho"])


def test_admission_budget_and_backfill():
    say("[gpu] admission")
    m = naca(33, 17)
    bound = Batch.case_bytes(m, max_iter=100)
    with Batch() as b:
        b.add(m, naca_params(), drop=0.0, max_iter=100)
        need = b.info()["bytes"]
    assert 0 < need <= bound
    with Batch(hbm_budget=2.5 * need, max_cases=8) as b:
        ids = [b.add(m, naca_params(), drop=0.0, max_iter=100) for _ in range(2)]
        with pytest.raises(E3DError) as e:
            b.add(m, naca_params(), drop=0.0, max_iter=100)
        assert e.value.code == E3D_ERR_BUSY
        b.run(10)
        b.remove(ids[0])
        ids.append(b.add(m, naca_params(), drop=0.0, max_iter=100))
        b.run(10)
        assert b.status(ids[1])["iterations"] == 20
        assert b.status(ids[2])["iterations"] == 10
        assert b.info()["cases"] == 2
    with Batch(max_cases=1) as b:
        b.add(m, naca_params(), drop=0.0, max_iter=10)
        with pytest.raises(E3DError) as e:
            b.add(m, naca_params(), drop=0.0, max_iter=10)
        assert e.value.code == E3D_ERR_BUSY


def test_server_results_equal_alone_results():
    """Continuous batching: requests admitted while others run, retired and backfilled, each give
    bitwise the same coefficients and iteration count as the same case solved alone."""
    say("[gpu] server vs alone")
    from euler3d_amd.serve import Server, summarize
    from euler3d_amd.workloads import poisson_arrivals, request_mix

    cases = request_mix(6, seed=3, min_cells=2e3, max_cells=4e4)
    built = [c.build() for c in cases]
    srv = Server(max_cases=3, drop=1e-3, max_iter=300, admit_workers=2, poll_ms=2.0)
    try:
        rec = srv.serve([(m, p, c.kind) for (m, p), c in zip(built, cases)], poisson_arrivals(6, 50.0, seed=3))
    finally:
        srv.close()
    s = summarize(rec)
    assert s["requests"] == 6 and s["p99_s"] >= s["p50_s"] > 0
    for (m, p), r in zip(built, rec):
        with Batch() as b:
            cid = b.add(m, p, drop=1e-3, max_iter=300)
            b.run(300)
            st = b.status(cid)
            alone = b.result(cid)
        assert r["iterations"] == st["iterations"]
        assert r["CL"] == alone["CL"] and r["CD"] == alone["CD"], (r, alone["CL"], alone["CD"])


def test_nan_case_is_retired_with_an_error():
    """A case whose residual turns NaN (here: a NaN Mach number through the ABI, so its freestream
    and state are NaN) is retired with E3D_CASE_FAILED and E3D_ERR_NUMERIC from status and result,
    while the other cases of the mix are bitwise what they are alone."""
    import ctypes as C

    from euler3d_amd import native as N

    say("[gpu] NaN case in a mix")
    a_mesh, c_mesh = naca(), naca(49, 25, thickness=0.09, camber=0.02)
    pa, pc = naca_params(), naca_params(mach=0.6, aoa=2.0, scheme=1)
    pbad = naca_params(mach=float("nan"))
    alone_a, alone_c = mixed_run(a_mesh, pa, 60), mixed_run(c_mesh, pc, 60)
    with Batch() as b:
        ia = b.add(a_mesh, pa, drop=0.0, max_iter=60)
        ib = b.add(naca(41, 21), pbad, drop=0.0, max_iter=60)
        ic = b.add(c_mesh, pc, drop=0.0, max_iter=60)
        b.run(60)
        sb = b.status(ib)
        rb = b.result(ib)
        done, it = C.c_int(), C.c_int()
        r0, r1 = C.c_double(), C.c_double()
        rc = N.lib().e3d_batch_status(b._h, ib, C.byref(done), C.byref(it), C.byref(r0), C.byref(r1))
        polled = {s["id"]: s for s in b.poll()}
        ra, rcres = b.result(ia), b.result(ic)
        sta, stc = b.get_state(ia), b.get_state(ic)
    assert rc == N.E3D_ERR_NUMERIC and done.value == 3
    assert sb["outcome"] == "failed" and sb["iterations"] == 1 and rb["failed"]
    assert polled[ib]["outcome"] == "failed" and polled[ia]["outcome"] == "capped"
    assert not ra["failed"] and not rcres["failed"]
    for r, st, alone in ((ra, sta, alone_a), (rcres, stc, alone_c)):
        assert bits_equal(r["rms"], alone["rms"])
        for k in ("rho", "u", "v", "w", "p", "E", "H", "M"):
            assert bits_equal(st[k], alone["state"][k]), k
        assert r["CL"] == alone["CL"] and r["CD"] == alone["CD"]


def test_case_larger_than_the_budget_is_refused_outright():
    say("[gpu] oversized case")
    m = naca(33, 17)
    with Batch() as b:
        b.add(m, naca_params(), drop=0.0, max_iter=100)
        need = b.info()["bytes"]
    with Batch(hbm_budget=0.5 * need) as b:
        with pytest.raises(E3DError) as e:
            b.add(m, naca_params(), drop=0.0, max_iter=100)
        assert e.value.code == 1 and "whole HBM budget" in str(e.value)
