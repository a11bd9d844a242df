"""The device's AUSM flux through e3d_face_flux, pinned by what the mathematics fixes.

The reference declares AUSM without defining it (include/solver/ConvectiveFluxesSchemes.h:27-30).
The CPU tests pin the oracle's restatement (tests/test_ausm_invariants.py); these run the same
kinds of check on the HIP flux of a real mesh, whose jittered tetrahedra give faces of every
orientation:
* consistency: a uniform state gives F(U, U, n) = physical flux on every interior face;
* upwinding: where M_L, M_R > 1 the flux is F(U_L), where M_L, M_R < -1 it is F(U_R);
* conservation: the residual sums to the boundary faces' F.S;
and the device flux equals the oracle's bit for bit on the same states.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
from oracle_ctypes import Oracle

from euler3d_amd import EulerSolver, Mesh, default_params

pytestmark = pytest.mark.gpu

GAMMA = 1.4
NAMES = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit", "faceSurfaces",
         "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost", "farCell", "farFace",
         "symGhost", "symCell", "symFace")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bump():
    m = Mesh.bump(48, 12, 12, jitter=0.1, seed=0, permute=True)
    m.preprocess(1.0)
    return m


def full_state(rho, u, v, w, p) -> dict:
    E = p / ((GAMMA - 1.0) * rho) + (u * u + v * v + w * w) / 2.0
    H = E + p / rho
    M = np.sqrt(GAMMA * (p / rho)) / np.sqrt(u * u + v * v + w * w)
    return dict(rho=rho, u=u, v=v, w=w, p=p, E=E, H=H, M=M)


def physical_flux(s, idx, n):
    V = s["u"][idx] * n[:, 0] + s["v"][idx] * n[:, 1] + s["w"][idx] * n[:, 2]
    r = s["rho"][idx]
    return np.stack([r * V, r * s["u"][idx] * V + s["p"][idx] * n[:, 0], r * s["v"][idx] * V + s["p"][idx] * n[:, 1],
                     r * s["w"][idx] * V + s["p"][idx] * n[:, 2], r * s["H"][idx] * V], axis=1)


def flux_scale(s, idx):
    a = np.sqrt(GAMMA * s["p"][idx] / s["rho"][idx])
    q = np.sqrt(s["u"][idx] ** 2 + s["v"][idx] ** 2 + s["w"][idx] ** 2)
    return (s["rho"][idx] * (a + q) * (1.0 + q + np.abs(s["H"][idx])) + s["p"][idx])[:, None]


def solver(mesh):
    return EulerSolver(mesh, default_params(mach=0.5, aoa_deg=0.0, cfl=1.5, scheme=1))


def test_gpu_ausm_consistency(bump):
    info = bump.info()
    ntot, nfi = info["n_elem"] + info["n_ghost"], info["n_face_int"]
    n = bump.get("faceNormalsUnit").reshape(-1, 3)[:nfi]
    f2e = bump.get("face2element").reshape(-1, 2)[:nfi]
    rng = np.random.default_rng(11)
    with solver(bump) as s:
        for trial, mach in enumerate((1e-3, 0.3, 0.8, 0.999, 1.0, 1.5, 3.0)):
            say(f"[gpu] ausm consistency trial={trial} mach={mach}")
            rho, p = rng.uniform(0.3, 3.0), rng.uniform(0.3, 3.0)
            d = rng.normal(size=3)
            d *= mach * np.sqrt(GAMMA * p / rho) / np.linalg.norm(d)
            one = np.ones(ntot)
            st = full_state(rho * one, d[0] * one, d[1] * one, d[2] * one, p * one)
            s.set_state(st)
            F, _ = s.face_flux()
            F = F.reshape(-1, 5)[:nfi]
            ref = physical_flux(st, f2e[:, 0], n)
            err = np.abs(F - ref) / flux_scale(st, f2e[:, 0])
            assert err.max() < 1e-14, (mach, err.max())


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_gpu_ausm_supersonic_upwinding(bump, sign):
    say(f"[gpu] ausm supersonic upwinding sign={sign}")
    info = bump.info()
    ne, ntot, nfi = info["n_elem"], info["n_elem"] + info["n_ghost"], info["n_face_int"]
    n = bump.get("faceNormalsUnit").reshape(-1, 3)[:nfi]
    f2e = bump.get("face2element").reshape(-1, 2)[:nfi]
    rng = np.random.default_rng(12)
    rho = rng.uniform(0.8, 1.25, ntot)
    p = rng.uniform(0.8, 1.25, ntot)
    d = np.array([0.8, 0.5, 0.33])
    d /= np.linalg.norm(d)
    vel = d[None, :] * 4.0 * np.sqrt(GAMMA) + rng.normal(scale=0.3, size=(ntot, 3))
    st = full_state(rho, vel[:, 0].copy(), vel[:, 1].copy(), vel[:, 2].copy(), p)
    a = np.sqrt(GAMMA * st["p"] / st["rho"])
    ML = (vel[f2e[:, 0]] * n).sum(1) / a[f2e[:, 0]]
    MR = (vel[f2e[:, 1]] * n).sum(1) / a[f2e[:, 1]]
    sel = (sign * ML > 1.0) & (sign * MR > 1.0)
    assert sel.sum() > 0.2 * nfi, sel.sum()
    with solver(bump) as s:
        s.set_state(st)
        F, _ = s.face_flux()
    F = F.reshape(-1, 5)[:nfi][sel]
    up = f2e[sel, 0] if sign > 0 else f2e[sel, 1]
    ref = physical_flux(st, up, n[sel])
    err = np.abs(F - ref) / flux_scale(st, up)
    assert err.max() < 1e-14, err.max()
    assert ne > 0


def test_gpu_ausm_conservation_and_oracle(bump):
    """Sum of the residual over the cells = sum of F.S over the boundary faces, on a random
    subsonic state after the ghosts are updated; fluxes and residual bitwise equal the oracle's."""
    say("[gpu] ausm conservation")
    info = bump.info()
    ne, ntot, nfi = info["n_elem"], info["n_elem"] + info["n_ghost"], info["n_face_int"]
    rng = np.random.default_rng(13)
    d = rng.normal(size=(ntot, 3))
    d *= (rng.uniform(0.0, 0.9, ntot) * np.sqrt(GAMMA))[:, None] / np.linalg.norm(d, axis=1)[:, None]
    st = full_state(rng.uniform(0.7, 1.4, ntot), d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(),
                    rng.uniform(0.7, 1.4, ntot))
    with solver(bump) as s:
        s.set_state(st)
        s.update_bc()
        F, FS = s.face_flux()
        R = s.residual()
        bc = s.get_state()
    FS5 = FS.reshape(-1, 5)
    lhs = R.reshape(-1, 5).sum(0)
    rhs = FS5[nfi:].sum(0)
    tol = 1e-12 * np.abs(FS5).sum(0).max()
    assert np.abs(lhs - rhs).max() < tol, (lhs, rhs)
    o = Oracle({k: bump.get(k) for k in NAMES}, dict(mach=0.5, scheme=1))
    o.set_state(st)
    o.update_bc()
    for k in ("rho", "u", "v", "w", "p", "H"):
        assert np.array_equal(bc[k].view(np.uint64), o.state[k].view(np.uint64)), k
    oF, oFS, _ = o.face_flux()
    assert np.array_equal(F.view(np.uint64), oF.view(np.uint64))
    assert np.array_equal(R.view(np.uint64), o.residual().view(np.uint64))
    assert ne > 0
