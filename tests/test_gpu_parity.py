"""HIP path against the reference's golden vectors and the oracle, through the C ABI.

Bar: bit-exact FP64 for BC ghost states, per-face fluxes, residuals, local time steps and states
after RK5 iterations (same operation order, no contraction, IEEE division and sqrt). Reductions
(RMS sum, force integration) use a fixed tree instead of the reference's sequential sum; they are
compared at 1e-12 relative, inside the stated residual (1e-10) and CL/CD (1e-9) tolerances.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
from conftest import FIELDS, INTEGRATOR_CASES, SMALL_CASES, golden_mesh, golden_params, golden_seeds, \
    integrator_params, load_golden, native_params, write_par
from oracle_ctypes import Oracle

from euler3d_amd import EulerSolver, Mesh

pytestmark = pytest.mark.gpu


def bits_equal(a, b) -> bool:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_solver(case, tmp_path, **over):
    g = load_golden(case)
    pd = golden_params(g)
    m = Mesh.read(str(write_par(case, tmp_path / f"{case}.par"))).preprocess(pd["cref"])
    return g, pd, m, EulerSolver(m, native_params(pd, **over))


@pytest.mark.parametrize("case", SMALL_CASES)
def test_gpu_bc_flux_residual_dt_iterate(case, tmp_path):
    say(f"[gpu] bc/flux/residual/dt/iterate case={case}")
    g, pd, m, s = make_solver(case, tmp_path)
    with s:
        for seed in golden_seeds(g):
            pre = f"flux{seed}_"
            s.set_state({k: g[pre + "in_" + k] for k in FIELDS})
            s.update_bc()
            st = s.get_state()
            for k in FIELDS:
                assert bits_equal(st[k], g[pre + "bc_" + k]), (case, seed, "bc", k)
            F, FS = s.face_flux()
            assert bits_equal(F, g[pre + "face_flux"]), (case, seed, "F")
            assert bits_equal(FS, g[pre + "face_flux_scaled"]), (case, seed, "FS")
            assert bits_equal(s.residual(), g[pre + "residual"]), (case, seed, "R")
            assert bits_equal(s.timestep(), g[pre + "dt"]), (case, seed, "dt")
            # Run's body from the seeded state: updateBC, computeResidual, RungeKutta
            s.set_state({k: g[pre + "in_" + k] for k in FIELDS})
            sums = s.iterate(1)
            st = s.get_state()
            for k in FIELDS:
                assert bits_equal(st[k], g[pre + "rk_" + k]), (case, seed, "rk", k)
            ref = float(g[pre + "rk_rms_sum"][0])
            assert abs(sums[0] - ref) <= 1e-12 * abs(ref), (sums[0], ref)


@pytest.mark.parametrize("cells", ["auto", "128", "256"])
@pytest.mark.parametrize("case", SMALL_CASES)
def test_gpu_run_matches_reference(case, cells, tmp_path, monkeypatch):
    """A whole run against the reference, with the fused stage's blocks of the size the library
    picks and of both sizes forced."""
    say(f"[gpu] run case={case} cells={cells}")
    if cells != "auto":
        monkeypatch.setenv("E3D_BLOCK_CELLS", cells)
    g, pd, m, s = make_solver(case, tmp_path)
    with s:
        r = s.run()
        st = s.get_state()
    assert r["iterations"] == int(g["run_iterations"][0])
    for k in FIELDS:
        assert bits_equal(st[k], g["run_" + k]), (case, k)
    rd = g["run_residuals_dat"]
    assert r["rows"].shape == rd.shape
    np.testing.assert_array_equal(r["rows"][:, 0], rd[:, 0])
    np.testing.assert_allclose(r["rows"][:, 1], rd[:, 1], rtol=6e-11)      # 11 digits in residuals.dat
    np.testing.assert_allclose(r["rows"][:, 2:], rd[:, 2:], rtol=6e-11, atol=1e-13)
    for k in ("CL", "CD", "CM"):
        ref = float(g["run_" + k][0])
        assert abs(r[k] - ref) <= 1e-9 * max(abs(ref), 1e-3), (k, r[k], ref)


def test_gpu_full_naca_2000_iterations():
    """129x129x2 O-grid (16,384 hexes), 2000 RK5 iterations against the reference: final density
    and pressure bit-exact, residual history to 11 digits, CL/CD to 1e-9."""
    say("[gpu] naca_full 2000 iterations")
    g = load_golden("naca_full")
    pd = golden_params(g)
    m = Mesh.naca(ni=129, nj=129, nk=2, thickness=0.12, wall_spacing=1e-3, farfield_radius=20.0, dz=0.1)
    m.preprocess(pd["cref"])
    with EulerSolver(m, native_params(pd)) as s:
        r = s.run()
        st = s.get_state()
    n = m.info()["n_elem"]
    assert r["iterations"] == 2000
    assert bits_equal(st["rho"][:n], g["run_rho"])
    assert bits_equal(st["p"][:n], g["run_p"])
    rd = g["run_residuals_dat"]
    np.testing.assert_allclose(r["rows"][:, 1], rd[:, 1], rtol=6e-11)
    for k in ("CL", "CD"):
        ref = float(g["run_" + k][0])
        assert abs(r[k] - ref) <= 1e-9 * abs(ref), (k, r[k], ref)


@pytest.mark.parametrize("case", ["naca_small", "bump_small"])
def test_gpu_ausm_matches_oracle(case, tmp_path):
    """AUSM (no reference implementation): the device and the oracle restate the same published
    method; they must agree bit for bit on fluxes and iterations."""
    say(f"[gpu] ausm case={case}")
    g, pd, m, s = make_solver(case, tmp_path, scheme=1)
    o = Oracle(golden_mesh(g), dict(pd, scheme=1))
    with s:
        pre = "flux0_"
        state = {k: g[pre + "in_" + k] for k in FIELDS}
        s.set_state(state)
        o.set_state(state)
        F, FS = s.face_flux()
        oF, oFS, _ = o.face_flux()
        assert bits_equal(F, oF) and bits_equal(FS, oFS)
        # five iterations from freestream
        s.set_state({k: o.state[k] * 0 + v for k, v in zip(FIELDS, [o.params.rho_inf, o.params.u_inf, o.params.v_inf,
                                                                        o.params.w_inf, o.params.p_inf, o.params.E_inf,
                                                                        o.params.H_inf, o.params.M_inf])})
        o2 = Oracle(golden_mesh(g), dict(pd, scheme=1))
        sums = s.iterate(5)
        osums = o2.iterate(5)
        st = s.get_state()
        for k in ("rho", "u", "v", "w", "p", "E", "H"):
            assert bits_equal(st[k], o2.state[k]), k
        np.testing.assert_allclose(sums, osums, rtol=1e-12)


def test_gpu_bit_reproducible(tmp_path):
    """Two runs of the same case give identical bits (no atomics, fixed reduction order)."""
    say("[gpu] reproducibility")
    m = Mesh.bump(nx=24, ny=8, nz=8, jitter=0.1, seed=0, permute=True).preprocess(1.0)
    pd = golden_params(load_golden("bump_small"))
    outs = []
    for _ in range(2):
        with EulerSolver(m, native_params(pd)) as s:
            sums = s.iterate(20)
            outs.append((sums, s.get_state()))
    assert bits_equal(outs[0][0], outs[1][0])
    for k in FIELDS:
        assert bits_equal(outs[0][1][k], outs[1][1][k])


@pytest.mark.parametrize("scheme", [0, 1])
def test_gpu_reads_between_iterations_change_nothing(scheme, tmp_path):
    """The fused RK stages keep the interior state as conservative W and every reader converts it
    back to primitives. Reading the state, fluxes, residual, time step and coefficients between
    iterations must leave the run's bits as they are."""
    say(f"[gpu] interleaved reads scheme={scheme}")
    m = Mesh.bump(nx=24, ny=8, nz=8, jitter=0.1, seed=0, permute=True).preprocess(1.0)
    pd = golden_params(load_golden("bump_small"))
    with EulerSolver(m, native_params(pd, scheme=scheme)) as s:
        full = s.iterate(10)
        ref = s.get_state()
        cf_ref, cm_ref = s.aero()
    with EulerSolver(m, native_params(pd, scheme=scheme)) as s:
        parts = [s.iterate(3)]
        st = s.get_state()
        cf, cm = s.aero()
        parts.append(s.iterate(2))
        s.face_flux()
        s.residual()
        s.timestep()
        parts.append(s.iterate(1))
        cf2, _ = s.aero()
        parts.append(s.iterate(4))
        cf3, cm3 = s.aero()
        out = s.get_state()
    assert all(np.isfinite(st[k]).all() for k in FIELDS) and np.isfinite(cf).all() and np.isfinite(cf2).all()
    assert bits_equal(np.concatenate(parts), full)
    for k in FIELDS:
        assert bits_equal(out[k], ref[k]), k
    assert bits_equal(cf3, cf_ref) and bits_equal(cm3, cm_ref)


@pytest.mark.parametrize("scheme", [0, 1])
def test_gpu_launch_switches_do_not_change_results(scheme, monkeypatch):
    """Each launch-shape and arithmetic-sequence switch of the fused stage (library IEEE divisions
    instead of the shared-reciprocal and short sequences, cached loads, plain dispatch order, one
    walk direction, Morton cell order, 128-cell blocks, each thread shape of the fixed-stride
    stage) gives the state bits of the default."""
    say(f"[gpu] launch switches scheme={scheme}")
    m = Mesh.bump(nx=24, ny=8, nz=8, jitter=0.1, seed=0, permute=True).preprocess(1.0)
    pd = golden_params(load_golden("bump_small"))

    def run():
        with EulerSolver(m, native_params(pd, scheme=scheme)) as s:
            sums = s.iterate(12)
            return sums, s.get_state()

    base_sums, base = run()
    for var, val in (("E3D_SHARED_RECIP", "0"), ("E3D_NONTEMPORAL", "0"), ("E3D_XCDS", "1"), ("E3D_ALTERNATE", "0"),
                     ("E3D_CELL_ORDER", "morton"), ("E3D_BLOCK_CELLS", "128"), ("E3D_FIXED_SHAPE", "0"),
                     ("E3D_FIXED_SHAPE", "1"), ("E3D_FIXED_SHAPE", "2")):
        with monkeypatch.context() as mp:
            mp.setenv(var, val)
            sums, st = run()
        for k in FIELDS:
            assert bits_equal(st[k], base[k]), (var, val, k)
        # the RMS sums block partials, which follow the cell order and block size
        np.testing.assert_allclose(sums, base_sums, rtol=1e-12, err_msg=var)


@pytest.mark.parametrize("scheme", [0, 1])
def test_gpu_midsize_mesh_matches_oracle(scheme):
    """A permuted 48x12x12 x 6 tet bump (41,472 cells, 162 blocks of 256 spread over the XCDs):
    eight RK5 iterations from freestream on the device and in the oracle give the same state
    bits, through the fused stage with its first stage on primitive records and the rest on
    conservative states."""
    say(f"[gpu] mid-size mesh vs oracle scheme={scheme}")
    names = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit", "faceSurfaces",
             "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost", "farCell", "farFace",
             "symGhost", "symCell", "symFace")
    m = Mesh.bump(nx=48, ny=12, nz=12, jitter=0.1, seed=3, permute=True).preprocess(1.0)
    pd = dict(golden_params(load_golden("bump_small")), scheme=scheme)
    o = Oracle({k: m.get(k) for k in names}, pd)
    with EulerSolver(m, native_params(pd)) as s:
        sums = s.iterate(8)
        st = s.get_state()
    osums = o.iterate(8)
    for k in ("rho", "u", "v", "w", "p", "E", "H"):
        assert bits_equal(st[k], o.state[k]), k
    np.testing.assert_allclose(sums, osums, rtol=1e-12)


def test_gpu_short_sqrt_and_division_sequences_are_ieee():
    """The fused stage's square root, reciprocal and quotient without operand scaling give the
    library's (IEEE) bits on 2^26 pseudo-random inputs each, zeros of either sign included."""
    import ctypes as C

    from euler3d_amd import native as N

    say("[gpu] short sqrt/div sequences")
    bad = (C.c_ulonglong * 3)()
    N.check(N.lib().e3d_math_selfcheck(0, 1 << 26, 12345, bad))
    assert list(bad) == [0, 0, 0], list(bad)


def test_gpu_freestream_preserved_and_conservative(tmp_path):
    """Size-independent properties on a mid-size permuted tet mesh: a uniform state is a steady
    state of the interior scheme (closed cells), and the sum of all cell residuals equals the sum of
    the boundary-face fluxes (each interior face adds and subtracts the same F*S)."""
    say("[gpu] invariants")
    m = Mesh.bump(nx=48, ny=12, nz=12, jitter=0.1, seed=0, permute=True).preprocess(1.0)
    pd = golden_params(load_golden("bump_small"))
    info = m.info()
    with EulerSolver(m, native_params(pd)) as s:
        st = s.get_state()
        F, FS = s.face_flux()
        R = s.residual().reshape(-1, 5)
        nfi = info["n_face_int"]
        bsum = FS.reshape(-1, 5)[nfi:].sum(axis=0)
        np.testing.assert_allclose(R.sum(axis=0), bsum, rtol=1e-9, atol=1e-9)
        # interior fluxes of a uniform state reduce to the physical flux: residual of an interior
        # cell (no boundary face) vanishes to round-off
        f2e = m.get("face2element").reshape(-1, 2)
        bcell = np.zeros(info["n_elem"], bool)
        bcell[f2e[nfi:, 0]] = True
        assert np.max(np.abs(R[~bcell])) < 1e-12
        assert st["rho"].shape == (info["n_elem"] + info["n_ghost"],)


def test_gpu_error_paths(tmp_path):
    say("[gpu] error paths")
    from euler3d_amd import E3DError

    m = Mesh.box(2, 2, 2).preprocess(1.0)
    pd = golden_params(load_golden("block8"))
    with pytest.raises(E3DError, match="TIME_INTEGRATION"):
        EulerSolver(m, native_params(pd, time_integration=3))
    with pytest.raises(E3DError, match="device"):
        EulerSolver(m, native_params(pd), device=99)
    with EulerSolver(m, native_params(pd)) as s:
        with pytest.raises(ValueError):
            s.set_state({k: np.zeros(3) for k in FIELDS})


@pytest.mark.parametrize("path", ["fused", "unfused"])
@pytest.mark.parametrize("case", sorted(INTEGRATOR_CASES))
def test_gpu_other_integrators_match_reference(case, path, tmp_path, monkeypatch):
    """Explicit Euler (TIME_INTEGRATION= 0), implicit Euler (1), residual smoothing, and Run's
    stopping rules and sampling period, against full reference runs: iteration count, final state
    bit for bit, residual history, coefficients. The stop-rule thresholds sit at least 1e-3
    relative away from every RMS of the run, far outside the 1e-12 by which the device's
    fixed-tree RMS sum may differ from the reference's sequential sum."""
    say(f"[gpu] integrator case={case} path={path}")
    if path == "unfused":
        monkeypatch.setenv("E3D_STAGE_PATH", "unfused")
    g = load_golden(case)
    base = INTEGRATOR_CASES[case]
    pd = integrator_params(g)
    m = Mesh.read(str(write_par(base, tmp_path / f"{base}.par"))).preprocess(pd["cref"])
    with EulerSolver(m, native_params(pd)) as s:
        r = s.run()
        st = s.get_state()
    assert r["iterations"] == int(g["run_iterations"][0])
    for k in FIELDS:
        assert bits_equal(st[k], g["run_" + k]), (case, k)
    rd = g["run_residuals_dat"]
    assert r["rows"].shape == rd.shape
    np.testing.assert_allclose(r["rows"], rd, rtol=6e-11, atol=1e-300)
    for k in ("CL", "CD", "CM"):
        ref = float(g["run_" + k][0])
        assert abs(r[k] - ref) <= 1e-12 * max(abs(ref), 1e-300), (k, r[k], ref)


@pytest.mark.parametrize("case", ["naca_small", "naca_euler"])
def test_gpu_checkpoint_resume_is_exact(case, tmp_path):
    """Run n iterations, or n/2, checkpoint, read into a fresh context, n/2 more: same bits."""
    say(f"[gpu] checkpoint case={case}")
    from euler3d_amd import E3DError

    g = load_golden(case)
    base = INTEGRATOR_CASES.get(case, case)
    pd = integrator_params(g)
    m = Mesh.read(str(write_par(base, tmp_path / f"{base}.par"))).preprocess(pd["cref"])
    with EulerSolver(m, native_params(pd)) as s:
        full = s.iterate(30)
        ref = s.get_state()
    ck = tmp_path / "state.ckpt"
    with EulerSolver(m, native_params(pd)) as s:
        first = s.iterate(15)
        s.checkpoint_write(str(ck))
    with EulerSolver(m, native_params(pd)) as s:
        s.checkpoint_read(str(ck))
        second = s.iterate(15)
        st = s.get_state()
        with pytest.raises(E3DError):
            s.checkpoint_read(str(tmp_path / "missing.ckpt"))
    assert bits_equal(np.concatenate([first, second]), full)
    for k in FIELDS:
        assert bits_equal(st[k], ref[k]), k
    other = Mesh.box(3, 3, 3).preprocess(1.0)
    with EulerSolver(other, native_params(pd)) as s:
        with pytest.raises(E3DError, match="not a checkpoint of this mesh"):
            s.checkpoint_read(str(ck))
