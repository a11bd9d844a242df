"""The CPU oracle against the reference's golden vectors (bit-exact).

The fixtures were written by the reference solver itself (tests/golden/make_golden.py); here the
oracle runs on the reference's own connectivity/metrics dump and seeded states and must reproduce
every FP64 bit: BC ghost states, per-face fluxes, residuals, local time steps, one RK5 step, and a
full Run (final state, coefficients, residual history to the 11 digits residuals.dat keeps).
"""
from __future__ import annotations

import numpy as np
import pytest
from conftest import FIELDS, INTEGRATOR_CASES, SMALL_CASES, golden_mesh, golden_params, golden_seeds, \
    integrator_params, load_golden
from oracle_ctypes import Oracle


def bits_equal(a, b) -> bool:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("case", SMALL_CASES)
def test_oracle_flux_bc_residual_dt_rk(case):
    g = load_golden(case)
    mesh, params = golden_mesh(g), golden_params(g)
    for s in golden_seeds(g):
        pre = f"flux{s}_"
        o = Oracle(mesh, params)
        o.set_state({k: g[pre + "in_" + k] for k in FIELDS})
        o.update_bc()
        for k in FIELDS:
            assert bits_equal(o.state[k], g[pre + "bc_" + k]), (case, s, "bc", k)
        F, FS, kind = o.face_flux()
        assert np.array_equal(kind, g[pre + "face_kind"])
        assert bits_equal(F, g[pre + "face_flux"]), (case, s)
        assert bits_equal(FS, g[pre + "face_flux_scaled"]), (case, s)
        assert bits_equal(o.residual(), g[pre + "residual"]), (case, s)
        assert bits_equal(o.timestep(), g[pre + "dt"]), (case, s)
        o.rk5()
        for k in FIELDS:
            assert bits_equal(o.state[k], g[pre + "rk_" + k]), (case, s, "rk", k)
        assert bits_equal(o.R, g[pre + "rk_residual"])
        assert bits_equal([o.rms_sum()], g[pre + "rk_rms_sum"])


@pytest.mark.parametrize("case", SMALL_CASES)
def test_oracle_run(case):
    g = load_golden(case)
    o = Oracle(golden_mesh(g), golden_params(g))
    r = o.run()
    assert r["iterations"] == int(g["run_iterations"][0])
    for k in FIELDS:
        assert bits_equal(o.state[k], g["run_" + k]), (case, k)
    assert bits_equal(o.R, g["run_residual"])
    assert bits_equal([o.rms_sum()], g["run_rms_sum"])
    for k in ("CL", "CD", "CM"):
        assert bits_equal([r[k]], g["run_" + k]), k
    assert bits_equal(r["force_coeff"], g["run_force_coeff"])
    assert bits_equal(r["moment_coeff"], g["run_moment_coeff"])
    rd = g["run_residuals_dat"]
    assert r["rows"].shape == rd.shape
    # residuals.dat keeps 11 significant digits (%.10e, ResidualsFile.hpp:40)
    np.testing.assert_allclose(r["rows"], rd, rtol=6e-11, atol=1e-300)


@pytest.mark.parametrize("case", sorted(INTEGRATOR_CASES))
def test_oracle_run_other_integrators(case):
    """Explicit Euler (TIME_INTEGRATION= 0), implicit Euler (1: dt only), implicit residual
    smoothing (RESIDUAL_SMOOTHING= 1), and Run's stopping rules and sampling period (a MIN_RESIDUAL
    crossed between two samples, MIN_AEROCOEFF_ERROR met twice in a row after an isolated hit,
    SAMPLING= 7 with MAX_ITER not a multiple of it) against full reference runs."""
    g = load_golden(case)
    mesh = golden_mesh(load_golden(INTEGRATOR_CASES[case]))
    o = Oracle(mesh, integrator_params(g))
    r = o.run()
    assert r["iterations"] == int(g["run_iterations"][0])
    for k in FIELDS:
        assert bits_equal(o.state[k], g["run_" + k]), (case, k)
    assert bits_equal(o.R, g["run_residual"])
    for k in ("CL", "CD", "CM"):
        assert bits_equal([r[k]], g["run_" + k]), k
    rd = g["run_residuals_dat"]
    assert r["rows"].shape == rd.shape
    np.testing.assert_allclose(r["rows"], rd, rtol=6e-11, atol=1e-300)


def test_oracle_smoothing_matches_two_jacobi_sweeps():
    """Smoothing pins beyond the fixtures: a zero residual takes one sweep and stays zero; any other
    residual takes exactly two Jacobi sweeps R <- (R0 + 0.25 sum of interior neighbours) / (1 +
    0.25 n) (the reference's convergence loop stops after its second pass), checked against an
    independent numpy evaluation of the two sweeps."""
    g = load_golden("naca_small")
    mesh = golden_mesh(g)
    pd = dict(golden_params(g), residual_smoothing=1)
    o = Oracle(mesh, pd)
    o.R[:] = 0.0
    assert o.smooth() == 1 and np.all(o.R == 0.0)
    rng = np.random.default_rng(3)
    R0 = rng.standard_normal((o.n_elem, 5))
    o.R[:] = R0.ravel()
    assert o.smooth() == 2
    start, e2e = mesh["element2faceStart"], mesh["element2element"]
    nface = np.diff(start)
    rows = np.repeat(np.arange(o.n_elem), nface)
    inner = e2e < o.n_elem

    def sweep(x):
        acc = np.zeros_like(x)
        np.add.at(acc, rows[inner], 0.25 * x[e2e[inner]])
        return (R0 + acc) / (1 + 0.25 * nface)[:, None]

    np.testing.assert_allclose(o.R.reshape(-1, 5), sweep(sweep(R0)), rtol=1e-14, atol=1e-15)


def test_oracle_threads_keep_every_bit():
    """The multi-threaded oracle (parallel face fluxes, per-cell ascending-face gather) gives the
    sequential restatement's bits, so the CPU baseline timed on all cores is the same computation."""
    import oracle_ctypes

    g = load_golden("bump_small")
    outs = []
    for th in (1, 4):
        oracle_ctypes.set_threads(th)
        try:
            o = Oracle(golden_mesh(g), golden_params(g))
            sums = o.iterate(3)
            outs.append((sums, {k: o.state[k].copy() for k in FIELDS}, o.R.copy()))
        finally:
            oracle_ctypes.set_threads(1)
    assert bits_equal(outs[0][0], outs[1][0]) and bits_equal(outs[0][2], outs[1][2])
    for k in FIELDS:
        assert bits_equal(outs[0][1][k], outs[1][1][k]), k


def test_stop_rule_fixtures_reach_their_branches():
    """The stop-rule fixtures exercise what they are for: naca_minres stops on the residual between
    two samples (its last residuals.dat row is the break row, one iteration past a sample that is
    not on the period); naca_coeffconv has an isolated sample under MIN_AEROCOEFF_ERROR that resets
    the counter before two in a row stop the run; bump_sampling7 ends between two samples."""
    g = load_golden("naca_minres")
    rows, (period, min_res, _) = g["run_residuals_dat"], g["stop"]
    it = int(g["run_iterations"][0])
    assert rows[-1, 0] == it and it % int(period) != 0 and rows[-1, 1] < min_res <= rows[:-1, 1].min()
    assert np.all(rows[:-1, 0] % period == 0)
    g = load_golden("naca_coeffconv")
    pd = golden_params(g)
    rows, tol = g["run_residuals_dat"], pd["min_aerocoeff_error"]
    (a0, s0), (a1, s1), (a2, s2) = pd["orient"]
    c = np.stack([rows[:, 2 + a0] * s0, rows[:, 2 + a1] * s1, rows[:, 5 + a2] * s2], 1)
    crit = (np.abs(np.diff(np.vstack([np.zeros(3), c]), axis=0)) < tol).all(1)
    assert crit[-1] and crit[-2] and not crit[-3]
    assert any(crit[k] and not crit[k + 1] for k in range(len(crit) - 3)), "no isolated hit before the stop"
    assert int(g["run_iterations"][0]) < pd["max_iter"]
    g = load_golden("bump_sampling7")
    assert int(g["run_iterations"][0]) == 40 and g["run_residuals_dat"][-1, 0] == 35
