"""Full-size parity: the device against the oracle on the benchmark meshes themselves.

* the 3L x L x L bump channel, 256 x 64 x 64 hexes split into 6 tets (6,291,456 cells, 12,656,640
  faces, jitter 0.1, permuted numbering, seed 0), Roe and AUSM, M 0.5: two RK5 iterations from
  freestream, every state array bit for bit, RMS sums within 1e-12 (fixed-tree against sequential
  sum); and two runs of five iterations bitwise equal;
* the swept tapered wing, 193 x 97 x 65 nodes (1,179,648 hexes), M 0.84, AoA 3 deg, Roe: two RK5
  iterations, the same bar.

The oracle runs on the host cores (OpenMP over the face fluxes and the per-cell gathers, the same
bits as its sequential form, tests/test_oracle_golden.py).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest
import oracle_ctypes
from oracle_ctypes import Oracle

from euler3d_amd import EulerSolver, Mesh, default_params

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p", "E", "H", "M")
NAMES = ("face2element", "element2faceStart", "element2face", "faceNormals", "faceNormalsUnit", "faceSurfaces",
         "faceCenters", "cellVolumes", "wallGhost", "wallCell", "wallFace", "farGhost", "farCell", "farFace",
         "symGhost", "symCell", "symFace")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b) -> bool:
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def host_threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def oracle_run(arrays, params: dict, iters: int):
    oracle_ctypes.set_threads(host_threads())
    try:
        o = Oracle(arrays, params)
        sums = o.iterate(iters)
        return sums, o.state
    finally:
        oracle_ctypes.set_threads(1)


@pytest.fixture(scope="module")
def bump_full():
    say("[gpu] building the 256x64x64 x 6 bump")
    m = Mesh.bump(256, 64, 64, jitter=0.1, seed=0, permute=True).preprocess(1.0)
    assert m.info()["n_elem"] == 6291456 and m.info()["n_face"] == 12656640
    return m, {k: m.get(k) for k in NAMES}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("scheme", [0, 1])
def test_gpu_full_bump_matches_oracle_and_repeats(bump_full, scheme):
    m, arrays = bump_full
    prm = default_params(mach=0.5, aoa_deg=0.0, cfl=1.5, scheme=scheme)
    with EulerSolver(m, prm) as s:
        start = s.get_state()
        say(f"[gpu] full bump scheme={scheme}: 2 iterations on the device")
        sums = s.iterate(2)
        st = s.get_state()
        say(f"[gpu] full bump scheme={scheme}: oracle on {host_threads()} threads")
        osums, ost = oracle_run(arrays, dict(mach=0.5, scheme=scheme), 2)
        for k in FIELDS:
            assert bits_equal(st[k], ost[k]), (scheme, k)
        np.testing.assert_allclose(sums, osums, rtol=1e-12, atol=0)
        if scheme == 0:
            say("[gpu] full bump: two runs of 5 iterations")
            runs = []
            for _ in range(2):
                s.set_state(start)
                runs.append((s.iterate(5), s.get_state()))
            assert bits_equal(runs[0][0], runs[1][0])
            for k in FIELDS:
                assert bits_equal(runs[0][1][k], runs[1][1][k]), k


@pytest.mark.timeout(600)
def test_gpu_full_wing_matches_oracle():
    say("[gpu] building the 193x97x65 wing")
    m = Mesh.wing(ni=193, nj=97, nk=65, thickness=0.10, sweep_deg=30.0, taper=0.56, semispan=1.5, farfield=15.0)
    m.preprocess(1.0)
    assert m.info()["n_elem"] == 1179648
    pd = dict(mach=0.84, aoa_deg=3.0, cfl=1.5)
    with EulerSolver(m, default_params(**pd)) as s:
        sums = s.iterate(2)
        st = s.get_state()
    say("[gpu] full wing: oracle")
    osums, ost = oracle_run({k: m.get(k) for k in NAMES}, pd, 2)
    for k in FIELDS:
        assert bits_equal(st[k], ost[k]), k
    np.testing.assert_allclose(sums, osums, rtol=1e-12, atol=0)
