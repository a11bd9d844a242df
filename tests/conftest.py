"""Shared test helpers: golden fixtures, the oracle built on them, and the product's mesh reader."""
from __future__ import annotations

import gzip
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(GOLDEN))

FIELDS = ("rho", "u", "v", "w", "p", "E", "H", "M")
SMALL_CASES = ("block8", "naca9x9x5", "naca_small", "bump_small", "wing_small", "hybrid_small")
# run-mode fixtures -> the fixture whose mesh they use: the other integrators, then Run's stopping
# rules and sampling period (MIN_RESIDUAL, MIN_AEROCOEFF_ERROR, SAMPLING)
INTEGRATOR_CASES = {"naca_euler": "naca_small", "naca_implicit": "naca_small", "bump_smooth": "bump_small",
                    "wing_euler_smooth": "wing_small", "naca_minres": "naca_small", "naca_coeffconv": "naca_small",
                    "bump_sampling7": "bump_small"}
STOP_CASES = ("naca_minres", "naca_coeffconv", "bump_sampling7")
MESH_KEYS_I = ("face2element", "face2nodeStart", "face2node", "element2faceStart", "element2face",
               "element2element", "wallGhost", "wallCell", "wallFace", "farGhost", "farCell", "farFace",
               "symGhost", "symCell", "symFace")
MESH_KEYS_F = ("faceNormals", "faceNormalsUnit", "faceSurfaces", "faceCenters", "cellVolumes", "cellCentroids")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def load_golden(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz") as d:
        return {k: d[k] for k in d.files}


def golden_mesh(g: dict) -> dict:
    return {k[5:]: v for k, v in g.items() if k.startswith("mesh_")}


def axis_sign(code: int) -> tuple[int, int]:
    """MESH_ORIENTATION code -> (axis, sign), SimConfig.cpp:185-230."""
    return code // 2, 1 if code % 2 == 0 else -1


def golden_params(g: dict) -> dict:
    p = g["params"]
    ocl, ocd = axis_sign(int(p[9])), axis_sign(int(p[10]))
    ocm = (3 - ocl[0] - ocd[0], -ocl[1] * ocd[1])
    stop = g["stop"] if "stop" in g else (1, 0.0, 0.0)   # SAMPLING, MIN_RESIDUAL, MIN_AEROCOEFF_ERROR
    return dict(mach=float(p[0]), aoa_deg=float(p[1]), gamma=float(p[2]), cfl=float(p[3]), sref=float(p[4]),
                cref=float(p[5]), ref_point=(float(p[6]), float(p[7]), float(p[8])), orient=(ocl, ocd, ocm),
                max_iter=int(p[11]), sampling=int(stop[0]), min_residual=float(stop[1]),
                min_aerocoeff_error=float(stop[2]), scheme=0)


def integrator_params(g: dict) -> dict:
    """golden_params plus TIME_INTEGRATION / RESIDUAL_SMOOTHING of the fixture (RK5, none if absent)."""
    pd = golden_params(g)
    ti, sm = (int(x) for x in g["integrator"]) if "integrator" in g else (2, 0)
    pd.update(time_integration=ti, residual_smoothing=sm)
    return pd


def golden_seeds(g: dict) -> list[int]:
    return sorted({int(k[4:].split("_")[0]) for k in g if k.startswith("flux") and k.endswith("_in_rho")})


def write_par(name: str, path: Path) -> Path:
    path.write_bytes(gzip.decompress((GOLDEN / f"{name}.par.gz").read_bytes()))
    return path


def native_params(pd: dict, **over):
    from euler3d_amd import default_params

    kw = dict(mach=pd["mach"], aoa_deg=pd["aoa_deg"], gamma=pd["gamma"], cfl=pd["cfl"], sref=pd["sref"],
              cref=pd["cref"], ref_point=pd["ref_point"], orient_cl=pd["orient"][0], orient_cd=pd["orient"][1],
              orient_cm=pd["orient"][2], max_iter=pd["max_iter"], sampling=pd["sampling"],
              min_residual=pd["min_residual"], min_aerocoeff_error=pd["min_aerocoeff_error"], scheme=pd["scheme"],
              time_integration=pd.get("time_integration", 2), residual_smoothing=pd.get("residual_smoothing", 0))
    kw.update(over)
    return default_params(**kw)


def hybrid_mesh(nx: int, ny: int, jitter: float = 0.08, seed: int = 3):
    """The hybrid mesh of the hybrid_small fixture (hexahedra, pyramids, tetrahedra and wedges in
    nx x ny columns, tests/golden/make_golden.py:hybrid_text) at another size, read and
    preprocessed by the product."""
    import tempfile

    from make_golden import hybrid_text

    from euler3d_amd import Mesh

    with tempfile.TemporaryDirectory() as td:
        path = Path(td) / "hybrid.par"
        path.write_text(hybrid_text(nx, ny, jitter, seed))
        return Mesh.read(str(path), par=True).preprocess(1.0)


@pytest.fixture(scope="session")
def gpu_available() -> bool:
    from euler3d_amd.native import has_device

    return has_device()
