"""Host-side mirror of the reference's solver interface over the C ABI.

Reference seam (Euler3D):
    SimConfig(file)                       src/parser/SimConfig.cpp:15-236       -> SimConfig
    Mesh<MeshPartition>(file).solveConnectivity(); Metrics(mesh, rank, config)
                                          include/mesh/Mesh.hpp:30-105, src/mesh/Metrics.cpp:9-56 -> Mesh
    FlowField(config, mesh) + EulerSolver(...).Run()
                                          src/solver/FlowField.cpp:8-101, src/solver/EulerSolver.cpp:21-167 -> EulerSolver

Errors raise native.E3DError (the reference exits the process instead). Every compute call runs
on the HIP device through libe3d_hip.so; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as N

STATE_FIELDS = ("rho", "u", "v", "w", "p", "E", "H", "M")

I32_ARRAYS = ("face2element", "face2nodeStart", "face2node", "element2faceStart", "element2face",
              "element2element", "wallGhost", "wallCell", "wallFace", "farGhost", "farCell", "farFace",
              "symGhost", "symCell", "symFace", "ghostKind", "elemVtk", "elemStart", "elemNodes", "mpiGhost",
              "mpiCell", "partStart", "haloStart", "haloFrom", "haloTo", "haloSrc", "haloDst", "deviceOrder")
F64_ARRAYS = ("faceNormals", "faceNormalsUnit", "faceSurfaces", "faceCenters", "cellVolumes",
              "cellCentroids", "xyz")


class Mesh:
    """A mesh read from SU2/.par or generated, and (after preprocess) its faces, ghosts and metrics."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self.cref = None

    # ---- construction
    @classmethod
    def read(cls, path: str, par: bool = True) -> "Mesh":
        h = C.c_void_p()
        N.check(N.lib().e3d_mesh_read(str(path).encode(), int(par), C.byref(h)))
        return cls(h.value)

    @classmethod
    def read_partitions(cls, paths) -> "Mesh":
        """The partition files of one multi-rank run (rank r reads paths[r]) as one mesh."""
        arr = (C.c_char_p * len(paths))(*[str(p).encode() for p in paths])
        h = C.c_void_p()
        N.check(N.lib().e3d_mesh_read_partitions(len(paths), arr, C.byref(h)))
        return cls(h.value)

    @classmethod
    def box(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, tags=("farfield",) * 6) -> "Mesh":
        h = C.c_void_p()
        N.check(N.lib().e3d_mesh_gen_box(nx, ny, nz, lx, ly, lz, ",".join(tags).encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def naca(cls, ni=129, nj=129, nk=2, thickness=0.12, camber=0.0, camber_pos=0.0, wall_spacing=1e-3,
             farfield_radius=20.0, dz=0.1) -> "Mesh":
        h = C.c_void_p()
        N.check(N.lib().e3d_mesh_gen_naca(ni, nj, nk, thickness, camber, camber_pos, wall_spacing,
                                          farfield_radius, dz, C.byref(h)))
        return cls(h.value)

    @classmethod
    def bump(cls, nx=256, ny=64, nz=64, jitter=0.1, seed=0, permute=True) -> "Mesh":
        h = C.c_void_p()
        N.check(N.lib().e3d_mesh_gen_bump(nx, ny, nz, jitter, seed, int(permute), C.byref(h)))
        return cls(h.value)

    @classmethod
    def wing(cls, ni=193, nj=97, nk=65, thickness=0.10, sweep_deg=30.0, taper=0.56, semispan=1.5,
             farfield=15.0) -> "Mesh":
        h = C.c_void_p()
        N.check(N.lib().e3d_mesh_gen_wing(ni, nj, nk, thickness, sweep_deg, taper, semispan, farfield,
                                          C.byref(h)))
        return cls(h.value)

    # ---- operations
    def write(self, path: str, precision: int = 16, par: bool = True) -> None:
        N.check(N.lib().e3d_mesh_write(self._h, str(path).encode(), precision, int(par)))

    def write_partitions(self, paths, elem2part=None, precision: int = 10) -> None:
        """E3D_PRE: the partition files of this (preprocessed, single) mesh; elem2part None uses
        recursive coordinate bisection."""
        arr = (C.c_char_p * len(paths))(*[str(p).encode() for p in paths])
        ep = None
        if elem2part is not None:
            ep = np.ascontiguousarray(elem2part, dtype=np.int32)
            if ep.shape != (self.info()["n_elem"],):
                raise ValueError("one partition id per cell is required")
        N.check(N.lib().e3d_mesh_partition_write(self._h, len(paths), None if ep is None else N.iptr(ep), arr,
                                                 precision))

    def partition_assign(self, n_part: int) -> np.ndarray:
        out = np.empty(self.info()["n_elem"], dtype=np.int32)
        N.check(N.lib().e3d_mesh_partition_assign(self._h, n_part, N.iptr(out)))
        return out

    def partitions(self) -> list[dict]:
        """Where each partition sits in the (preprocessed) mesh: cells, nodes, faces, ghosts, walls."""
        n = C.c_int()
        N.check(N.lib().e3d_mesh_partition_count(self._h, C.byref(n)))
        out = []
        for r in range(n.value):
            q = N.PartitionInfo()
            N.check(N.lib().e3d_mesh_partition_info(self._h, r, C.byref(q)))
            out.append({k: getattr(q, k) for k, _ in N.PartitionInfo._fields_})
        return out

    def preprocess(self, cref: float = 1.0) -> "Mesh":
        N.check(N.lib().e3d_mesh_preprocess(self._h, float(cref)))
        self.cref = cref
        return self

    def info(self) -> dict:
        m = N.MeshInfo()
        N.check(N.lib().e3d_mesh_get_info(self._h, C.byref(m)))
        return {k: getattr(m, k) for k, _ in N.MeshInfo._fields_}

    def get(self, name: str) -> np.ndarray:
        n = C.c_longlong()
        if name in I32_ARRAYS:
            N.check(N.lib().e3d_mesh_get_i32(self._h, name.encode(), None, 0, C.byref(n)))
            out = np.empty(n.value, dtype=np.int32)
            N.check(N.lib().e3d_mesh_get_i32(self._h, name.encode(), N.iptr(out), n.value, C.byref(n)))
            return out
        if name in F64_ARRAYS:
            N.check(N.lib().e3d_mesh_get_f64(self._h, name.encode(), None, 0, C.byref(n)))
            out = np.empty(n.value, dtype=np.float64)
            N.check(N.lib().e3d_mesh_get_f64(self._h, name.encode(), N.dptr(out), n.value, C.byref(n)))
            return out
        raise KeyError(name)

    def arrays(self) -> dict:
        """The mesh's own arrays (the planned deviceOrder is computed on request, not included)."""
        return {k: self.get(k) for k in I32_ARRAYS + F64_ARRAYS if k != "deviceOrder"}

    def close(self) -> None:
        if self._h:
            N.lib().e3d_mesh_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class SimConfig:
    """Parsed .e3d file (SimConfig): solver parameters, the partition files and the output dir."""
    params: N.Params
    mesh_path: str
    output_dir: str
    partition_files: tuple = ()

    @classmethod
    def read(cls, path: str) -> "SimConfig":
        p = N.Params()
        mesh = C.create_string_buffer(4096)
        out = C.create_string_buffer(4096)
        N.check(N.lib().e3d_config_read(str(path).encode(), C.byref(p), mesh, 4096, out, 4096))
        n = C.c_int()
        N.check(N.lib().e3d_config_partition(str(path).encode(), 0, None, 0, C.byref(n)))
        files = []
        for i in range(n.value):
            buf = C.create_string_buffer(4096)
            N.check(N.lib().e3d_config_partition(str(path).encode(), i, buf, 4096, None))
            files.append(buf.value.decode())
        return cls(p, mesh.value.decode(), out.value.decode(), tuple(files))


class EulerSolver:
    """One case resident on a HIP device (FlowField + EulerSolver of the reference)."""

    def __init__(self, mesh: Mesh, params: N.Params, device: int = 0):
        h = C.c_void_p()
        N.check(N.lib().e3d_create(mesh._h, C.byref(params), device, C.byref(h)))
        self._h = h
        info = mesh.info()
        self.n_elem = info["n_elem"]
        self.n_face = info["n_face"]
        self.n_ghost = info["n_ghost"]
        self.n_tot = self.n_elem + self.n_ghost
        self.params = params

    def set_state(self, state: dict) -> None:
        arrs = [np.ascontiguousarray(state[k], dtype=np.float64) for k in STATE_FIELDS]
        for a in arrs:
            if a.shape != (self.n_tot,):
                raise ValueError(f"state arrays must have {self.n_tot} values")
        ptrs = (N.DP * 8)(*[N.dptr(a) for a in arrs])
        N.check(N.lib().e3d_set_state(self._h, ptrs))

    def get_state(self) -> dict:
        out = {k: np.empty(self.n_tot) for k in STATE_FIELDS}
        ptrs = (N.DP * 8)(*[N.dptr(out[k]) for k in STATE_FIELDS])
        N.check(N.lib().e3d_get_state(self._h, ptrs))
        return out

    def update_bc(self) -> None:
        N.check(N.lib().e3d_update_bc(self._h))

    def face_flux(self) -> tuple[np.ndarray, np.ndarray]:
        F = np.empty(5 * self.n_face)
        FS = np.empty(5 * self.n_face)
        N.check(N.lib().e3d_face_flux(self._h, N.dptr(F), N.dptr(FS)))
        return F, FS

    def residual(self) -> np.ndarray:
        R = np.empty(5 * self.n_elem)
        N.check(N.lib().e3d_residual(self._h, N.dptr(R)))
        return R

    def timestep(self) -> np.ndarray:
        dt = np.empty(self.n_elem)
        N.check(N.lib().e3d_timestep(self._h, N.dptr(dt)))
        return dt

    def iterate(self, n: int) -> np.ndarray:
        sums = np.empty(max(n, 1))
        N.check(N.lib().e3d_iterate(self._h, int(n), N.dptr(sums)))
        return sums[:n]

    def aero(self) -> tuple[np.ndarray, np.ndarray]:
        cf = np.empty(3)
        cm = np.empty(3)
        N.check(N.lib().e3d_aero(self._h, N.dptr(cf), N.dptr(cm)))
        return cf, cm

    def wall_cp(self) -> np.ndarray:
        """cp of every wall face as the last force update left it (zeros before any)."""
        n = C.c_longlong(0)
        N.check(N.lib().e3d_wall_cp(self._h, None, 0, C.byref(n)))
        cp = np.empty(n.value)
        N.check(N.lib().e3d_wall_cp(self._h, N.dptr(cp), n.value, C.byref(n)))
        return cp

    def run(self, max_rows: int | None = None) -> dict:
        rows_n = max_rows if max_rows is not None else self.params.max_iter + 1
        rows = np.zeros(8 * max(rows_n, 1))
        nrows = C.c_int()
        iters = C.c_int()
        coeffs = np.empty(3)
        cf = np.empty(3)
        cm = np.empty(3)
        N.check(N.lib().e3d_run(self._h, N.dptr(rows), rows_n, C.byref(nrows), C.byref(iters), N.dptr(coeffs),
                                N.dptr(cf), N.dptr(cm)))
        k = min(nrows.value, rows_n)
        return {"rows": rows[: 8 * k].reshape(k, 8), "iterations": iters.value, "CL": coeffs[0], "CD": coeffs[1],
                "CM": coeffs[2], "force_coeff": cf, "moment_coeff": cm}

    def checkpoint_write(self, path: str) -> None:
        """Exact FP64 checkpoint (state, explicit Euler conservative state, Mach bookkeeping)."""
        N.check(N.lib().e3d_checkpoint_write(self._h, str(path).encode()))

    def checkpoint_read(self, path: str) -> None:
        N.check(N.lib().e3d_checkpoint_read(self._h, str(path).encode()))

    def elapsed_ms(self) -> float:
        ms = C.c_double()
        N.check(N.lib().e3d_last_elapsed_ms(self._h, C.byref(ms)))
        return ms.value

    def set_kernel_timing(self, enable: bool) -> None:
        N.check(N.lib().e3d_set_kernel_timing(self._h, int(enable)))

    def kernel_timing(self) -> tuple[float, int, str]:
        """(mean ms per launch, launches, kernel name) of the stage kernel since timing was enabled."""
        ms = C.c_double()
        n = C.c_int()
        name = C.create_string_buffer(128)
        N.check(N.lib().e3d_kernel_timing(self._h, C.byref(ms), C.byref(n), name, 128))
        return ms.value, n.value, name.value.decode()

    def close(self) -> None:
        if self._h:
            N.lib().e3d_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Mixed-precision multi-case engine: many cases resident on one device, advanced together by
    one launch per RK stage (no counterpart in the reference, which solves one case per
    process). A case's result does not depend on the other resident cases."""

    def __init__(self, device: int = 0, hbm_budget: float = 0.0, max_cases: int = 256, wg_per_cu: int = 0,
                 xcds: int = 8, nontemporal: bool = True, use_graph: bool = True):
        o = N.BatchOpts()
        N.check(N.lib().e3d_batch_opts_default(C.byref(o)))
        o.hbm_budget = float(hbm_budget)
        o.max_cases = int(max_cases)
        o.wg_per_cu = int(wg_per_cu)
        o.xcds = int(xcds)
        o.nontemporal = int(nontemporal)
        o.use_graph = int(use_graph)
        h = C.c_void_p()
        N.check(N.lib().e3d_batch_create(int(device), C.byref(o), C.byref(h)))
        self._h = h
        self._n = {}

    @staticmethod
    def case_opts(drop: float = 1e-6, max_iter: int = 5000, hist_cap: int | None = None,
                  stall: tuple | None = None) -> N.CaseOpts:
        """stall: (window, ratio, level) of the stall test, None: off. The case retires as "stalled"
        when the lowest RMS of a window of `window` iterations is not below ratio x the previous
        window's lowest while below level x the first RMS."""
        co = N.CaseOpts()
        co.drop = float(drop)
        co.max_iter = int(max_iter)
        co.hist_cap = int(max_iter if hist_cap is None else hist_cap)
        if stall is not None:
            co.stall_window, co.stall_ratio, co.stall_level = int(stall[0]), float(stall[1]), float(stall[2])
        return co

    @staticmethod
    def case_bytes(mesh: Mesh, drop: float = 1e-6, max_iter: int = 5000) -> float:
        b = C.c_double()
        co = Batch.case_opts(drop, max_iter)
        N.check(N.lib().e3d_batch_case_bytes(mesh._h, C.byref(co), C.byref(b)))
        return b.value

    def add(self, mesh: Mesh, params: N.Params, drop: float = 1e-6, max_iter: int = 5000,
            hist_cap: int | None = None, stall: tuple | None = None) -> int:
        """Admit a case; returns its id. Raises E3DError with code E3D_ERR_BUSY when full, E3D_ERR_ARG
        when it can never fit. stall: see case_opts."""
        co = self.case_opts(drop, max_iter, hist_cap, stall)
        cid = C.c_int()
        N.check(N.lib().e3d_batch_add(self._h, mesh._h, C.byref(params), C.byref(co), C.byref(cid)))
        info = mesh.info()
        self._n[cid.value] = (info["n_elem"], info["n_ghost"], co.hist_cap)
        return cid.value

    def run(self, n: int) -> None:
        N.check(N.lib().e3d_batch_run(self._h, int(n)))

    def wait(self) -> None:
        N.check(N.lib().e3d_batch_wait(self._h))

    def status(self, cid: int) -> dict:
        """done: finished; outcome: running / converged / capped / failed (non-finite residual) /
        stalled (residual at the FP32 flux floor; finish on the FP64 path)."""
        done, it = C.c_int(), C.c_int()
        r0, r1 = C.c_double(), C.c_double()
        rc = N.lib().e3d_batch_status(self._h, int(cid), C.byref(done), C.byref(it), C.byref(r0), C.byref(r1))
        if rc != N.E3D_ERR_NUMERIC:
            N.check(rc)
        return {"done": bool(done.value), "outcome": N.CASE_OUTCOMES[done.value], "iterations": it.value,
                "rms0": r0.value, "rms": r1.value}

    def poll(self) -> list[dict]:
        """Wait for the queued iterations; progress of every resident case."""
        cap = 1024
        while True:
            ids, done, it = (np.empty(cap, dtype=np.int32) for _ in range(3))
            rms = np.empty(cap)
            n = C.c_int()
            N.check(N.lib().e3d_batch_poll(self._h, N.iptr(ids), N.iptr(done), N.iptr(it), N.dptr(rms), cap,
                                           C.byref(n)))
            if n.value <= cap:
                return [{"id": int(ids[k]), "done": bool(done[k]), "outcome": N.CASE_OUTCOMES[int(done[k])],
                         "iterations": int(it[k]), "rms": float(rms[k])} for k in range(n.value)]
            cap = n.value

    def result(self, cid: int) -> dict:
        cap = self._n[cid][2]
        hist = np.empty(max(cap, 1))
        nh = C.c_int()
        coeffs, cf, cm = np.empty(3), np.empty(3), np.empty(3)
        rc = N.lib().e3d_batch_result(self._h, int(cid), N.dptr(hist), cap, C.byref(nh), N.dptr(coeffs),
                                      N.dptr(cf), N.dptr(cm))
        if rc != N.E3D_ERR_NUMERIC:
            N.check(rc)
        return {"rms": hist[: nh.value].copy(), "CL": coeffs[0], "CD": coeffs[1], "CM": coeffs[2],
                "force_coeff": cf, "moment_coeff": cm, "failed": rc == N.E3D_ERR_NUMERIC}

    def get_state(self, cid: int) -> dict:
        n_elem, n_ghost, _ = self._n[cid]
        out = {k: np.empty(n_elem + n_ghost) for k in STATE_FIELDS}
        ptrs = (N.DP * 8)(*[N.dptr(out[k]) for k in STATE_FIELDS])
        N.check(N.lib().e3d_batch_get_state(self._h, int(cid), ptrs))
        return out

    def remove(self, cid: int) -> None:
        N.check(N.lib().e3d_batch_remove(self._h, int(cid)))
        self._n.pop(cid, None)

    def set_tuning(self, wg_per_cu: int, xcds: int, nontemporal: bool) -> None:
        N.check(N.lib().e3d_batch_set_tuning(self._h, int(wg_per_cu), int(xcds), int(nontemporal)))

    def info(self) -> dict:
        n, used, budget = C.c_int(), C.c_double(), C.c_double()
        N.check(N.lib().e3d_batch_info(self._h, C.byref(n), C.byref(used), C.byref(budget)))
        return {"cases": n.value, "bytes": used.value, "budget": budget.value}

    def set_timing(self, enable: bool) -> None:
        N.check(N.lib().e3d_batch_set_timing(self._h, int(enable)))

    def timing(self) -> tuple[float, int, bool]:
        """(mean ms per timed launch, launches, whether a launch is a whole iteration)"""
        ms, n, per = C.c_double(), C.c_int(), C.c_int()
        N.check(N.lib().e3d_batch_timing(self._h, C.byref(ms), C.byref(n), C.byref(per)))
        return ms.value, n.value, bool(per.value)

    def close(self) -> None:
        if self._h:
            N.lib().e3d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
