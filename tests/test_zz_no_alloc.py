"""e3d_iterate_noalloc (the allocation-free variant of e3d_iterate) against e3d_iterate, through the
C ABI: two calls in a row on one device, the second on a smaller mesh with fewer iterations,
each giving bitwise the same RMS sums and state as e3d_iterate on a fresh context of the same case;
and the size limit of its static scratch."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from euler3d_amd import EulerSolver, Mesh, default_params
from euler3d_amd import native as N

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p", "E", "H", "M")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(mesh, prm, n, noalloc):
    with EulerSolver(mesh, prm) as s:
        sums = np.empty(n)
        fn = N.lib().e3d_iterate_noalloc if noalloc else N.lib().e3d_iterate
        N.check(fn(s._h, n, N.dptr(sums)))
        return sums, s.get_state()


def test_noalloc_iterate_matches_iterate():
    naca = Mesh.naca(ni=33, nj=17, nk=2, thickness=0.12, wall_spacing=1e-2, farfield_radius=20.0, dz=0.1)
    naca.preprocess(1.0)
    box = Mesh.box(3, 2, 2, tags=("farfield",) * 6).preprocess(1.0)
    p1 = default_params(mach=0.8, aoa_deg=1.25, cfl=1.5)
    p2 = default_params(mach=1.5, aoa_deg=1.0, cfl=0.5)
    for tag, mesh, prm, n in (("naca 512 hexes, 30 iterations", naca, p1, 30), ("box 12 hexes, 4 iterations", box, p2, 4)):
        say(f"[gpu] noalloc vs iterate: {tag}")
        a_sums, a_st = run(mesh, prm, n, noalloc=True)
        b_sums, b_st = run(mesh, prm, n, noalloc=False)
        assert np.array_equal(a_sums.view(np.uint64), b_sums.view(np.uint64)), tag
        for k in FIELDS:
            assert np.array_equal(a_st[k].view(np.uint64), b_st[k].view(np.uint64)), (tag, k)


def test_noalloc_iterate_rejects_more_iterations_than_its_scratch():
    box = Mesh.box(2, 2, 2, tags=("farfield",) * 6).preprocess(1.0)
    with EulerSolver(box, default_params(mach=0.5)) as s:
        rc = N.lib().e3d_iterate_noalloc(s._h, 65537, None)
        assert rc == N.E3D_ERR_ARG
        assert "65536" in N.lib().e3d_last_error().decode()
        sums = np.empty(2)
        N.check(N.lib().e3d_iterate_noalloc(s._h, 2, N.dptr(sums)))
        assert np.all(np.isfinite(sums))
    assert C.sizeof(C.c_double) == 8
