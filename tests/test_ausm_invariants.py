"""AUSM flux of the oracle pinned by what the mathematics fixes (CPU).

The reference only declares AUSM (include/solver/ConvectiveFluxesSchemes.h:27-30), so there is no
fixture to compare with. The oracle restates Liou & Steffen, J. Comput. Phys. 107 (1993) 23-39:
M = V/a with each side's own sound speed, interface Mach m = M+(M_L) + M-(M_R), pressure
p = P+(M_L) p_L + P-(M_R) p_R, F = m (Phi_L + Phi_R)/2 - |m| (Phi_R - Phi_L)/2 + p n with
Phi = rho a (1, u, v, w, H). These checks pin it:

* the splittings, restricted to |M| <= 1, are polynomials of degree 2 (M+-) and 3 (P+-) that join
  the supersonic branches with matching value and slope at M = +-1; those conditions determine
  the polynomials uniquely, and M+ + M- = M, P+ + P- = 1, M-(M) = -M+(-M), P-(M) = P+(-M);
* consistency: F(U, U, n) is the physical flux;
* upwinding: M_L, M_R > 1 gives F(U_L) and M_L, M_R < -1 gives F(U_R);
* mirror antisymmetry: F(L, R, n) = -F(R, L, -n), bit for bit;
* rotation invariance: rotating both velocities and the normal rotates the momentum flux;
* conservation: on a closed mesh the residual sums to the boundary fluxes, and a uniform state
  gives a zero residual.

The same checks run against oracle builds with a planted mistake (OR_AUSM_MUTANT 1-4: swapped
pressure splittings, one sound speed for both sides, no |m| term, a wrong pressure polynomial);
each mutant must fail at least one of them.
"""
from __future__ import annotations

import ctypes as C
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle_ctypes import DP, IP, OrMesh, OrParams, OrState, Oracle

GAMMA = 1.4


def _load(path) -> C.CDLL:
    L = C.CDLL(str(path))
    m, p, s = C.POINTER(OrMesh), C.POINTER(OrParams), C.POINTER(OrState)
    L.or_face_flux.restype = None
    L.or_face_flux.argtypes = [m, p, s, DP, DP, IP]
    L.or_ausm_split.restype = None
    L.or_ausm_split.argtypes = [C.c_double, DP]
    return L


def _build(out, mutant: int):
    cmd = ["gcc", "-std=c99", "-O2", "-ffp-contract=off", "-fPIC", "-shared", f"-DOR_AUSM_MUTANT={mutant}",
           str(ROOT / "oracle" / "e3d_oracle.c"), "-o", str(out), "-lm", "-fopenmp"]
    subprocess.run(cmd, check=True, capture_output=True)
    return _load(out)


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    """oracle builds: 0 = the oracle as shipped, 1-4 = planted mistakes"""
    d = tmp_path_factory.mktemp("ausm")
    return {k: _build(d / f"oracle_ausm_{k}.so", k) for k in range(5)}


def split(L, M: float) -> np.ndarray:
    out = np.empty(4)
    L.or_ausm_split(float(M), out.ctypes.data_as(DP))
    return out


def ausm(L, sl: dict, sr: dict, n: np.ndarray) -> np.ndarray:
    """AUSM flux of face i between state sl[i] (elem0) and sr[i] (elem1), unit normal n[i]."""
    nf = len(n)
    st = {k: np.empty(2 * nf) for k in ("rho", "u", "v", "w", "p", "E", "H", "M")}
    for k in st:
        st[k][0::2] = sl[k]
        st[k][1::2] = sr[k]
    keep = []

    def ip(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(IP)

    def dp(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(DP)

    m = OrMesh()
    m.nElem, m.nFace, m.nFaceInt, m.nGhost = 2 * nf, nf, nf, 0
    m.face2element = ip(np.arange(2 * nf))
    m.faceUnitNormal = dp(n.reshape(-1))
    m.faceArea = dp(np.ones(nf))
    q = OrParams()
    q.gamma = GAMMA
    q.scheme = 1
    s = OrState(*[st[k].ctypes.data_as(DP) for k in ("rho", "u", "v", "w", "p", "E", "H", "M")])
    F = np.empty(5 * nf)
    FS = np.empty(5 * nf)
    kind = np.empty(nf, dtype=np.int32)
    L.or_face_flux(C.byref(m), C.byref(q), C.byref(s), F.ctypes.data_as(DP), FS.ctypes.data_as(DP),
                   kind.ctypes.data_as(IP))
    assert (kind == 0).all()
    return F.reshape(nf, 5)


def make_state(rho, vel, p) -> dict:
    u, v, w = vel[:, 0], vel[:, 1], vel[:, 2]
    E = p / ((GAMMA - 1.0) * rho) + (u * u + v * v + w * w) / 2.0
    H = E + p / rho
    M = np.sqrt(GAMMA * (p / rho)) / np.sqrt(u * u + v * v + w * w)
    return dict(rho=rho, u=u.copy(), v=v.copy(), w=w.copy(), p=p, E=E, H=H, M=M)


def unit(rng, n) -> np.ndarray:
    x = rng.normal(size=(n, 3))
    return x / np.linalg.norm(x, axis=1)[:, None]


def random_states(rng, n, mach_lo=0.0, mach_hi=3.0) -> dict:
    rho = rng.uniform(0.3, 3.0, n)
    p = rng.uniform(0.3, 3.0, n)
    a = np.sqrt(GAMMA * p / rho)
    vel = unit(rng, n) * (rng.uniform(mach_lo, mach_hi, n) * a)[:, None]
    return make_state(rho, vel, p)


def physical_flux(s: dict, n: np.ndarray) -> np.ndarray:
    V = s["u"] * n[:, 0] + s["v"] * n[:, 1] + s["w"] * n[:, 2]
    r = s["rho"]
    return np.stack([r * V, r * s["u"] * V + s["p"] * n[:, 0], r * s["v"] * V + s["p"] * n[:, 1],
                     r * s["w"] * V + s["p"] * n[:, 2], r * s["H"] * V], axis=1)


def flux_scale(s: dict) -> np.ndarray:
    """size of the terms an AUSM flux sums, per face (for rounding tolerances)"""
    a = np.sqrt(GAMMA * s["p"] / s["rho"])
    q = np.sqrt(s["u"] ** 2 + s["v"] ** 2 + s["w"] ** 2)
    return (s["rho"] * (a + q) * (1.0 + q + np.abs(s["H"])) + s["p"])[:, None]


# ---- the checks, each a function of an oracle build -------------------------------------------

def check_splittings(L) -> None:
    Ms = np.concatenate([np.linspace(-1.0, 1.0, 41), [-0.3137, 0.7071, 1e-9, -1e-9]])
    for M in Ms:
        mp, mm, pp, pm = split(L, M)
        assert abs(mp + mm - M) <= 4e-16 * max(1.0, abs(M)), ("M+ + M- != M", M)
        assert abs(pp + pm - 1.0) <= 4e-16, ("P+ + P- != 1", M)
        mp2, mm2, pp2, pm2 = split(L, -M)
        assert mm == -mp2 and pm == pp2, ("mirror symmetry of the splittings", M)
    # degree: a polynomial of degree k has vanishing (k+1)-th differences on an even grid
    h = 0.25
    grid = np.arange(-1.0, 1.0 + 1e-12, h)
    vals = np.array([split(L, M) for M in grid])
    for col, deg in ((0, 2), (1, 2), (2, 3), (3, 3)):
        d = np.diff(vals[:, col], deg + 1)
        assert np.abs(d).max() < 1e-13, ("degree of the subsonic splitting", col)
    # values and one-sided slopes at the joins: value and slope of the supersonic branch
    e = 1e-6
    for M0 in (-1.0, 1.0):
        inner = (split(L, M0) - split(L, M0 - np.sign(M0) * e)) / e * np.sign(M0)   # slope from |M| < 1
        outer = (split(L, M0 + np.sign(M0) * e) - split(L, M0)) / e * np.sign(M0)   # slope from |M| > 1
        sup = split(L, M0 + np.sign(M0) * 1e-12)
        assert np.allclose(split(L, M0), sup, atol=1e-11), ("value jump at the join", M0)
        assert np.allclose(inner, outer, atol=1e-5), ("slope jump at the join", M0, inner, outer)
    # the supersonic branches themselves
    for M in (1.5, 3.0, -1.5, -3.0):
        mp, mm, pp, pm = split(L, M)
        assert (mp, mm) == ((M, 0.0) if M > 0 else (0.0, M))
        assert (pp, pm) == ((1.0, 0.0) if M > 0 else (0.0, 1.0))


def check_consistency(L, seed=0) -> None:
    rng = np.random.default_rng(seed)
    n = 3000
    s = random_states(rng, n, 0.0, 3.0)
    nv = unit(rng, n)
    F = ausm(L, s, s, nv)
    ref = physical_flux(s, nv)
    err = np.abs(F - ref) / flux_scale(s)
    assert err.max() < 1e-14, ("consistency", err.max())


def supersonic_pairs(rng, n, sign):
    nv = unit(rng, n)
    sides = []
    for _ in range(2):
        rho = rng.uniform(0.3, 3.0, n)
        p = rng.uniform(0.3, 3.0, n)
        a = np.sqrt(GAMMA * p / rho)
        tang = unit(rng, n)
        tang -= (tang * nv).sum(1)[:, None] * nv
        vel = nv * (sign * rng.uniform(1.02, 4.0, n) * a)[:, None] + tang * rng.uniform(0, 2, n)[:, None] * a[:, None]
        sides.append(make_state(rho, vel, p))
    return sides[0], sides[1], nv


def check_upwinding(L, seed=1) -> None:
    rng = np.random.default_rng(seed)
    for sign in (1.0, -1.0):
        sl, sr, nv = supersonic_pairs(rng, 2000, sign)
        for s in (sl, sr):
            M = (s["u"] * nv[:, 0] + s["v"] * nv[:, 1] + s["w"] * nv[:, 2]) / np.sqrt(GAMMA * s["p"] / s["rho"])
            assert (sign * M > 1.0).all()
        F = ausm(L, sl, sr, nv)
        up = sl if sign > 0 else sr
        err = np.abs(F - physical_flux(up, nv)) / flux_scale(up)
        assert err.max() < 1e-14, ("supersonic upwinding", sign, err.max())


def check_mirror(L, seed=2) -> None:
    rng = np.random.default_rng(seed)
    n = 3000
    sl, sr = random_states(rng, n, 0.0, 2.5), random_states(rng, n, 0.0, 2.5)
    nv = unit(rng, n)
    F = ausm(L, sl, sr, nv)
    G = ausm(L, sr, sl, -nv)
    assert np.array_equal(F, -G), ("mirror antisymmetry", np.abs(F + G).max())


def check_rotation(L, seed=3) -> None:
    rng = np.random.default_rng(seed)
    n = 2000
    sl, sr = random_states(rng, n, 0.0, 2.5), random_states(rng, n, 0.0, 2.5)
    nv = unit(rng, n)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))

    def rot(s):
        vel = np.stack([s["u"], s["v"], s["w"]], 1) @ Q.T
        return make_state(s["rho"], vel, s["p"])

    F = ausm(L, sl, sr, nv)
    G = ausm(L, rot(sl), rot(sr), nv @ Q.T)
    scale = np.maximum(flux_scale(sl), flux_scale(sr))
    want = F.copy()
    want[:, 1:4] = F[:, 1:4] @ Q.T
    err = np.abs(G - want) / scale
    assert err.max() < 1e-12, ("rotation invariance", err.max())


CHECKS = {"splittings": check_splittings, "consistency": check_consistency, "upwinding": check_upwinding,
          "mirror": check_mirror, "rotation": check_rotation}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_ausm_invariant(libs, name):
    CHECKS[name](libs[0])


# which checks each planted mistake must fail (it may fail more)
MUTANT_CAUGHT_BY = {1: ("upwinding",), 2: ("mirror",), 3: ("upwinding",), 4: ("splittings", "consistency")}


@pytest.mark.parametrize("mutant", sorted(MUTANT_CAUGHT_BY))
def test_planted_ausm_mistake_is_caught(libs, mutant):
    for name in MUTANT_CAUGHT_BY[mutant]:
        with pytest.raises(AssertionError):
            CHECKS[name](libs[mutant])


def test_ausm_conservation_on_closed_mesh():
    """Sum over cells of R = sum over boundary faces of F.S (interior faces cancel), and a uniform
    state gives R = 0 to rounding in every cell away from the boundary (the cells are closed), on
    the 8x4x4-hex (768-tet) bump with AUSM."""
    from euler3d_amd import Mesh

    mesh = Mesh.bump(8, 4, 4, jitter=0.1, seed=0, permute=True)
    mesh.preprocess(1.0)
    names = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit",
             "faceSurfaces", "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost",
             "farCell", "farFace", "symGhost", "symCell", "symFace")
    arrays = {k: mesh.get(k) for k in names}
    o = Oracle(arrays, dict(mach=0.5, scheme=1))
    o.update_bc()
    R0 = o.residual().reshape(-1, 5)
    scale0 = np.abs(o.face_flux()[1]).max()
    # cells whose faces are all interior: R = F(U, U, n) summed over a closed surface = 0
    start, e2f = arrays["element2faceStart"], arrays["element2face"]
    nfi = int(np.count_nonzero(arrays["face2element"][1::2] < len(arrays["cellVolumes"])))
    inner = np.array([(e2f[start[c]:start[c + 1]] < nfi).all() for c in range(len(start) - 1)])
    assert inner.sum() > 100
    assert np.abs(R0[inner]).max() < 1e-13 * scale0, "freestream is not preserved"
    rng = np.random.default_rng(7)
    st = {k: v.copy() for k, v in o.state.items()}
    ne = o.n_elem
    s = random_states(rng, ne, 0.0, 0.9)
    for k in ("rho", "u", "v", "w", "p", "E", "H", "M"):
        st[k][:ne] = s[k]
    o.set_state(st)
    o.update_bc()
    R = o.residual().reshape(-1, 5)
    F, FS, kind = o.face_flux()
    FS = FS.reshape(-1, 5)
    f2e = arrays["face2element"].reshape(-1, 2)
    bnd = f2e[:, 1] >= ne
    lhs = R.sum(0)
    rhs = FS[bnd].sum(0)
    tol = 1e-12 * np.abs(FS).sum(0).max()
    assert np.abs(lhs - rhs).max() < tol, (lhs, rhs)
