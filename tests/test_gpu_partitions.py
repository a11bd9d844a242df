"""Partitioned input on the HIP path: the partition files of an `mpiexec -np N` reference run read
as one resident mesh, against the reference's ranks (tests/golden/<case>.npz, r<rank>_* arrays).

Bar, as for one partition: bit-exact FP64 for ghost states, per-face fluxes (the rank's own face
order), residuals, time steps and states after iterations, for every partition; the RMS and force
sums at 1e-12 relative (fixed-tree reductions against the reference's sequential sums).

Both ways of reading a partition interface run: "direct" (the face reads the other partition's cell)
and "halo" (E3D_EXCHANGE=halo: per-pair messages packed from the send cells into a device buffer and
unpacked into the interface ghosts before every stage, MPIHandler::updateFlowField). With the
exchange the interface ghosts are also compared after updateBC, where the reference holds the
values the other rank sent; directly, they are compared after an update only (their own cell's
copy).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
from conftest import FIELDS, integrator_params, load_golden, native_params
from test_partitions import PART_CASES, composite, join_cells, join_interior, split_faces

from euler3d_amd import EulerSolver

pytestmark = pytest.mark.gpu

RUN_CASES = PART_CASES + ("bump_small_p4_smooth", "wing_small_p2_euler_smooth")


def bits_equal(a, b) -> bool:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def seeds_of(g):
    return sorted({int(k.split("_")[1][4:]) for k in g if k.startswith("r0_flux") and k.endswith("_in_rho")})


@pytest.mark.parametrize("exchange", ["direct", "halo"])
@pytest.mark.parametrize("case", PART_CASES)
def test_gpu_partitioned_bc_flux_residual_dt_iterate(case, exchange, tmp_path, monkeypatch):
    say(f"[gpu] partitioned bc/flux/residual/dt/iterate case={case} exchange={exchange}")
    monkeypatch.setenv("E3D_EXCHANGE", exchange)
    g = load_golden(case)
    pd = integrator_params(g)
    m = composite(case, tmp_path)
    parts = m.partitions()
    nFI = m.info()["n_face_int"]
    mpi = set(m.get("mpiGhost").tolist()) if exchange == "direct" else set()
    with EulerSolver(m, native_params(pd)) as s:
        keep = np.array([i for i in range(s.n_tot) if i not in mpi])
        for sd in seeds_of(g):
            pre = f"flux{sd}_"
            s.set_state({k: join_cells(g, pre + "in_" + k, parts) for k in FIELDS})
            s.update_bc()
            st = s.get_state()
            for k in FIELDS:
                assert bits_equal(st[k][keep], join_cells(g, pre + "bc_" + k, parts)[keep]), (case, sd, "bc", k)
            F, FS = s.face_flux()
            for r, (f, fs) in enumerate(zip(split_faces(F, parts, nFI, 5), split_faces(FS, parts, nFI, 5))):
                assert bits_equal(f, g[f"r{r}_{pre}face_flux"]), (case, sd, r, "F")
                assert bits_equal(fs, g[f"r{r}_{pre}face_flux_scaled"]), (case, sd, r, "FS")
            assert bits_equal(s.residual(), join_interior(g, pre + "residual", parts, 5)), (case, sd, "R")
            assert bits_equal(s.timestep(), join_interior(g, pre + "dt", parts)), (case, sd, "dt")
            s.set_state({k: join_cells(g, pre + "in_" + k, parts) for k in FIELDS})
            sums = s.iterate(1)
            st = s.get_state()
            for k in FIELDS:
                assert bits_equal(st[k], join_cells(g, pre + "rk_" + k, parts)), (case, sd, "rk", k)
            ref = sum(float(g[f"r{r}_{pre}rk_rms_sum"][0]) for r in range(len(parts)))
            assert abs(sums[0] - ref) <= 1e-12 * abs(ref), (sums[0], ref)


@pytest.mark.parametrize("path", ["fused", "fused256", "unfused", "fused-halo", "unfused-halo"])
@pytest.mark.parametrize("case", RUN_CASES)
def test_gpu_partitioned_run_matches_reference(case, path, tmp_path, monkeypatch):
    """A whole multi-rank run on one device: final state of every partition bit for bit, the
    iteration count, the residual history and the coefficients summed per partition."""
    say(f"[gpu] partitioned run case={case} path={path}")
    if path.startswith("unfused"):
        monkeypatch.setenv("E3D_STAGE_PATH", "unfused")
    if path == "fused256":
        monkeypatch.setenv("E3D_BLOCK_CELLS", "256")
    if path.endswith("halo"):
        monkeypatch.setenv("E3D_EXCHANGE", "halo")
    g = load_golden(case)
    pd = integrator_params(g)
    m = composite(case, tmp_path)
    parts = m.partitions()
    with EulerSolver(m, native_params(pd)) as s:
        r = s.run()
        st = s.get_state()
    assert r["iterations"] == int(g["r0_run_iterations"][0])
    for k in FIELDS:
        assert bits_equal(st[k], join_cells(g, "run_" + k, parts)), (case, k)
    rd = g["run_residuals_dat"]
    assert r["rows"].shape == rd.shape
    np.testing.assert_allclose(r["rows"], rd, rtol=6e-11, atol=1e-13)
    for k in ("CL", "CD", "CM"):
        ref = float(g["r0_run_" + k][0])
        assert abs(r[k] - ref) <= 1e-12 * max(abs(ref), 1e-12), (k, r[k], ref)


@pytest.mark.parametrize("exchange", ["direct", "halo"])
def test_gpu_partitioned_ausm_matches_oracle(exchange, tmp_path, monkeypatch):
    """AUSM (no reference implementation) on partitioned input: device and oracle, both on the
    composite mesh, agree bit for bit on the fluxes of a seeded state (after updateBC) and over
    iterations (the interface faces included)."""
    from oracle_ctypes import Oracle
    from test_partitions import composite_arrays

    say(f"[gpu] partitioned ausm case=bump_small_p4 exchange={exchange}")
    monkeypatch.setenv("E3D_EXCHANGE", exchange)
    g = load_golden("bump_small_p4")
    pd = integrator_params(g)
    m = composite("bump_small_p4", tmp_path)
    parts = m.partitions()
    arr = composite_arrays(m)
    with EulerSolver(m, native_params(pd, scheme=1)) as s:
        state = {k: join_cells(g, "flux0_in_" + k, parts) for k in FIELDS}
        o = Oracle(arr, dict(pd, scheme=1))
        s.set_state(state)
        o.set_state(state)
        s.update_bc()
        o.update_bc()
        F, FS = s.face_flux()
        oF, oFS, _ = o.face_flux()
        assert bits_equal(F, oF) and bits_equal(FS, oFS)
        o2 = Oracle(arr, dict(pd, scheme=1))
        s.set_state({k: o2.state[k].copy() for k in FIELDS})
        sums = s.iterate(5)
        osums = o2.iterate(5)
        st = s.get_state()
        for k in ("rho", "u", "v", "w", "p", "E", "H"):
            assert bits_equal(st[k], o2.state[k]), k
        np.testing.assert_allclose(sums, osums, rtol=1e-12)
