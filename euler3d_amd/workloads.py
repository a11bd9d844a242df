"""Seeded synthetic cases for the batched and served solves.

Three families, each a generator of this package:
    naca  2D-extruded NACA 4-digit O-grid (one cell layer, symmetry planes), hexes
    bump  3L x L x L channel with a circular-arc bump, hexes split into 6 tets
    wing  swept tapered wing in an O-H box, hexes
A Case records its generator arguments and flow parameters; Case.build() makes the preprocessed
mesh and the solver parameters. Draws use numpy's PCG64 generator with the given seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .native import Params, default_params
from .solver import Mesh


@dataclass
class Case:
    kind: str
    gen: dict
    mach: float
    aoa_deg: float
    scheme: int = 0
    sref: float = 1.0
    meta: dict = field(default_factory=dict)

    def cells(self) -> int:
        g = self.gen
        if self.kind == "naca":
            return (g["ni"] - 1) * (g["nj"] - 1) * (g["nk"] - 1)
        if self.kind == "bump":
            return 6 * g["nx"] * g["ny"] * g["nz"]
        return (g["ni"] - 1) * (g["nj"] - 1) * (g["nk"] - 1)

    def params(self) -> Params:
        return default_params(mach=self.mach, aoa_deg=self.aoa_deg, gamma=1.4, cfl=1.5, sref=self.sref, cref=1.0,
                              scheme=self.scheme, orient_cl=(1, 1), orient_cd=(0, 1), orient_cm=(2, 1))

    def mesh(self) -> Mesh:
        make = {"naca": Mesh.naca, "bump": Mesh.bump, "wing": Mesh.wing}[self.kind]
        return make(**self.gen).preprocess(1.0)

    def build(self) -> tuple[Mesh, Params]:
        return self.mesh(), self.params()


def naca_case(n: int, thickness: float, camber: float, mach: float, aoa: float, scheme: int = 0) -> Case:
    """n x n x 2 node O-grid, wall spacing 1e-3 chord, farfield at 20 chords, dz = 0.1 chord."""
    gen = dict(ni=n, nj=n, nk=2, thickness=thickness, camber=camber, camber_pos=0.4 if camber > 0 else 0.0,
               wall_spacing=1e-3, farfield_radius=20.0, dz=0.1)
    return Case("naca", gen, mach, aoa, scheme, sref=0.1)


def airfoil_batch(n_cases: int = 64, seed: int = 0) -> list[Case]:
    """Uneven 2D-extruded airfoil cases: n ~ U{65..257} nodes a side, thickness U[0.06, 0.15],
    camber U[0, 0.04], M U[0.3, 0.85], AoA U[-2, 4] degrees."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_cases):
        n = int(rng.integers(65, 258))
        out.append(naca_case(n, float(rng.uniform(0.06, 0.15)), float(rng.uniform(0.0, 0.04)),
                             float(rng.uniform(0.3, 0.85)), float(rng.uniform(-2.0, 4.0))))
    return out


def sized_case(kind: str, cells: float, rng: np.random.Generator) -> Case:
    """A case of the given family with about `cells` cells."""
    if kind == "naca":
        n = max(33, int(round(math.sqrt(cells))) + 1)
        return naca_case(n, float(rng.uniform(0.06, 0.15)), float(rng.uniform(0.0, 0.04)),
                         float(rng.uniform(0.3, 0.85)), float(rng.uniform(-2.0, 4.0)))
    if kind == "bump":
        ny = max(2, int(round((cells / 24.0) ** (1.0 / 3.0))))
        gen = dict(nx=4 * ny, ny=ny, nz=ny, jitter=0.1, seed=int(rng.integers(0, 1 << 31)), permute=True)
        return Case("bump", gen, float(rng.uniform(0.3, 0.6)), 0.0, 0, sref=3.0)
    s = (cells / (192.0 * 96.0 * 64.0)) ** (1.0 / 3.0)
    gen = dict(ni=max(17, int(round(192 * s)) + 1), nj=max(9, int(round(96 * s)) + 1), nk=max(7, int(round(64 * s)) + 1),
               thickness=0.10, sweep_deg=30.0, taper=0.56, semispan=1.5, farfield=15.0)
    return Case("wing", gen, float(rng.uniform(0.6, 0.84)), float(rng.uniform(0.0, 3.0)), 0, sref=1.0)


def request_mix(n: int, seed: int = 0, min_cells: float = 1e4, max_cells: float = 2e6,
                kinds=("naca", "naca", "bump", "wing")) -> list[Case]:
    """Solve requests from the three families (airfoils twice as likely), sizes log-uniform in
    [min_cells, max_cells]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = kinds[int(rng.integers(0, len(kinds)))]
        cells = math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))
        out.append(sized_case(kind, cells, rng))
    return out


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    """Arrival times (s) of a Poisson stream of the given rate (requests/s)."""
    rng = np.random.default_rng(seed + 1)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
