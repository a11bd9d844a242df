"""The CPU oracle against the reference's golden vectors (bit-exact).

The fixtures were written by the reference solver itself (tests/golden/make_golden.py); here the
oracle runs on the reference's own connectivity/metrics dump and seeded states and must reproduce
every FP64 bit: BC ghost states, per-face fluxes, residuals, local time steps, one RK5 step, and a
full Run (final state, coefficients, residual history to the 11 digits residuals.dat keeps).
"""
from __future__ import annotations

import numpy as np
import pytest
from conftest import FIELDS, SMALL_CASES, golden_mesh, golden_params, golden_seeds, load_golden
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
