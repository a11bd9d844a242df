"""The persistent, software-pipelined RK stage against the one-block-per-workgroup stage.

The pipelined kernel runs by default only where a block's LDS leaves room for one workgroup per CU
(hexahedral meshes with large halos, such as the full-size wing, whose state the full-size test
compares bit for bit with the oracle). Here E3D_PIPE forces it on and off for every plan it fits,
on tetrahedral, hexahedral, O-grid and hybrid (all four element types) meshes, Roe and AUSM, with both XCD orders, from a seeded
(non-uniform) state and from freestream: states and RMS sums after several RK5 iterations must be
the same bits. The hexahedral wing's plan fits one workgroup per CU, so its one-block-per-workgroup
runs also cover the 768- and 512-thread forms of that kernel. The iterations cross the first stage
(W0 stored), the last (ghosts and RMS partials stored) and the alternating block walk.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import hybrid_mesh

from euler3d_amd import EulerSolver, Mesh, default_params

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p", "E", "H", "M")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b) -> bool:
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


MESHES = {
    "bump": lambda: (Mesh.bump(48, 12, 12, jitter=0.1, seed=3, permute=True), dict(mach=0.5)),
    "wing": lambda: (Mesh.wing(ni=65, nj=33, nk=25, thickness=0.10, sweep_deg=30.0, taper=0.56, semispan=1.5,
                               farfield=15.0), dict(mach=0.84, aoa_deg=3.0)),
    "naca": lambda: (Mesh.naca(ni=129, nj=65, nk=2, thickness=0.12, wall_spacing=1e-3, farfield_radius=20.0,
                               dz=0.1), dict(mach=0.8, aoa_deg=1.25)),
    # hexahedra, pyramids, tetrahedra and wedges: cells of 4, 5 and 6 faces in one block
    "hybrid": lambda: (hybrid_mesh(16, 12), dict(mach=0.5, aoa_deg=2.0)),
}


def run(monkeypatch, mesh, prm, pipe: int, xcds: int, start=None, iters=7, wide_threads=768):
    monkeypatch.setenv("E3D_PIPE", str(pipe))
    monkeypatch.setenv("E3D_XCDS", str(xcds))
    monkeypatch.setenv("E3D_WIDE_THREADS", str(wide_threads))
    with EulerSolver(mesh, prm) as s:
        if start is not None:
            s.set_state(start)
        sums = s.iterate(iters)
        return sums, s.get_state()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("scheme", [0, 1])
def test_gpu_pipelined_stage_same_bits(monkeypatch, name, scheme):
    m, pd = MESHES[name]()
    if name != "hybrid":
        m.preprocess(1.0)
    prm = default_params(cfl=1.5, scheme=scheme, **pd)
    say(f"[gpu] pipelined stage {name} ({m.info()['n_elem']} cells) scheme={scheme}")
    # a non-uniform start: a few iterations from freestream on the default path
    with EulerSolver(m, prm) as s:
        s.iterate(3)
        start = s.get_state()
    for xcds in (8, 1):
        for st in (None, start):
            ref_sums, ref = run(monkeypatch, m, prm, 0, xcds, st)
            for pipe, wt in ((1, 768), (0, 512)):
                sums, got = run(monkeypatch, m, prm, pipe, xcds, st, wide_threads=wt)
                assert bits_equal(sums, ref_sums), (name, scheme, xcds, pipe, wt)
                for k in FIELDS:
                    assert bits_equal(got[k], ref[k]), (name, scheme, xcds, pipe, wt, k)
