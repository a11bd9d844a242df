"""ctypes binding of the CPU oracle (oracle/lib/libe3d_oracle.so). TEST INFRASTRUCTURE ONLY:
imported by tests/, the smoke check and the benchmark's CPU-baseline leg, never by the product.

The oracle takes the mesh as plain arrays (the names the golden fixtures and
euler3d_amd.Mesh.get use), so it can run on the reference's own connectivity/metrics dumps as well
as on the product's preprocessing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "lib" / "libe3d_oracle.so"

IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)


class OrMesh(C.Structure):
    _fields_ = [("nElem", C.c_int), ("nFace", C.c_int), ("nFaceInt", C.c_int), ("nGhost", C.c_int),
                ("face2element", IP), ("element2faceStart", IP), ("element2face", IP),
                ("faceNormal", DP), ("faceUnitNormal", DP), ("faceArea", DP), ("faceCenter", DP),
                ("cellVolume", DP), ("nWall", C.c_int), ("nFar", C.c_int), ("nSym", C.c_int),
                ("wallGhost", IP), ("wallCell", IP), ("wallFace", IP),
                ("farGhost", IP), ("farCell", IP), ("farFace", IP),
                ("symGhost", IP), ("symCell", IP), ("symFace", IP),
                ("nMpi", C.c_int), ("mpiGhost", IP), ("mpiCell", IP),
                ("nPart", C.c_int), ("partElemStart", IP), ("partWallStart", IP)]


class OrParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("mach", C.c_double), ("aoa_deg", C.c_double),
                ("sref", C.c_double), ("cref", C.c_double), ("ref_point", C.c_double * 3),
                ("orient", (C.c_int * 2) * 3), ("scheme", C.c_int), ("max_iter", C.c_int), ("sampling", C.c_int),
                ("min_residual", C.c_double), ("min_aerocoeff_error", C.c_double),
                ("rho_inf", C.c_double), ("p_inf", C.c_double), ("u_inf", C.c_double), ("v_inf", C.c_double),
                ("w_inf", C.c_double), ("E_inf", C.c_double), ("H_inf", C.c_double), ("M_inf", C.c_double),
                ("time_integration", C.c_int), ("residual_smoothing", C.c_int), ("element2element", IP)]


class OrState(C.Structure):
    _fields_ = [(k, DP) for k in ("rho", "u", "v", "w", "p", "E", "H", "M")]


FIELDS = ("rho", "u", "v", "w", "p", "E", "H", "M")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run python -m euler3d_amd.build")
        L = C.CDLL(str(LIB))
        m, p, s = C.POINTER(OrMesh), C.POINTER(OrParams), C.POINTER(OrState)
        sig = {
            "or_set_threads": (None, [C.c_int]),
            "or_freestream": (None, [p]),
            "or_init_state": (None, [m, p, s]),
            "or_update_bc": (None, [m, p, s]),
            "or_face_flux": (None, [m, p, s, DP, DP, IP]),
            "or_residual": (None, [m, p, s, DP]),
            "or_timestep": (None, [m, p, s, DP]),
            "or_rk5": (None, [m, p, s, DP]),
            "or_rms_sum": (C.c_double, [DP, C.c_int]),
            "or_aero": (None, [m, p, s, DP, DP]),
            "or_run": (C.c_int, [m, p, s, DP, DP, C.c_int, IP, DP, DP, DP]),
            "or_smooth": (C.c_int, [m, p, DP]),
            "or_euler": (None, [m, p, s, DP, DP]),
            "or_init_conservative": (None, [m, p, s, DP]),
            "or_run_any": (C.c_int, [m, p, s, DP, DP, DP, C.c_int, IP, DP, DP, DP]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Threads of the oracle's parallel loops (1: sequential); results are bitwise the same."""
    lib().or_set_threads(int(n))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(IP)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(DP)


class Oracle:
    """Mesh arrays + parameters + an FP64 state, advanced by the C oracle."""

    def __init__(self, arrays: dict, params: dict):
        self._keep = []
        m = OrMesh()
        f2e = np.asarray(arrays["face2element"])
        m.nFace = len(f2e) // 2
        m.nElem = len(arrays["cellVolumes"])
        if "nFaceInt" in arrays:  # partitioned meshes: interface faces have an interior elem1
            m.nFaceInt = int(arrays["nFaceInt"])
            m.nGhost = m.nFace - m.nFaceInt
        else:
            m.nGhost = int(f2e[1::2].max()) + 1 - m.nElem if m.nFace else 0
            nbnd = int(np.count_nonzero(f2e[1::2] >= m.nElem))
            m.nFaceInt = m.nFace - nbnd
        for name, field in (("face2element", "face2element"), ("element2faceStart", "element2faceStart"),
                            ("element2face", "element2face")):
            a, ptr = _i(arrays[name])
            self._keep.append(a)
            setattr(m, field, ptr)
        for name, field in (("faceNormals", "faceNormal"), ("faceNormalsUnit", "faceUnitNormal"),
                            ("faceSurfaces", "faceArea"), ("faceCenters", "faceCenter"),
                            ("cellVolumes", "cellVolume")):
            a, ptr = _d(arrays[name])
            self._keep.append(a)
            setattr(m, field, ptr)
        for grp in ("wall", "far", "sym"):
            for part in ("Ghost", "Cell", "Face"):
                a, ptr = _i(arrays[grp + part])
                self._keep.append(a)
                setattr(m, grp + part, ptr)
            setattr(m, "n" + grp.capitalize(), len(arrays[grp + "Ghost"]))
        if len(arrays.get("mpiGhost", ())):
            for name in ("mpiGhost", "mpiCell"):
                a, ptr = _i(arrays[name])
                self._keep.append(a)
                setattr(m, name, ptr)
            m.nMpi = len(arrays["mpiGhost"])
        if "partElemStart" in arrays and len(arrays["partElemStart"]) > 2:
            for name in ("partElemStart", "partWallStart"):
                a, ptr = _i(arrays[name])
                self._keep.append(a)
                setattr(m, name, ptr)
            m.nPart = len(arrays["partElemStart"]) - 1
        self.mesh = m
        q = OrParams()
        q.gamma = params.get("gamma", 1.4)
        q.cfl = params.get("cfl", 1.5)
        q.mach = params["mach"]
        q.aoa_deg = params.get("aoa_deg", 0.0)
        q.sref = params.get("sref", 1.0)
        q.cref = params.get("cref", 1.0)
        for k, v in enumerate(params.get("ref_point", (0.0, 0.0, 0.0))):
            q.ref_point[k] = v
        orient = params.get("orient", ((1, 1), (0, 1), (2, 1)))
        for i in range(3):
            q.orient[i][0], q.orient[i][1] = orient[i]
        q.scheme = params.get("scheme", 0)
        q.max_iter = params.get("max_iter", 1)
        q.sampling = params.get("sampling", 1)
        q.min_residual = params.get("min_residual", 0.0)
        q.min_aerocoeff_error = params.get("min_aerocoeff_error", 0.0)
        q.time_integration = params.get("time_integration", 2)
        q.residual_smoothing = params.get("residual_smoothing", 0)
        if "element2element" in arrays:
            a, ptr = _i(arrays["element2element"])
            self._keep.append(a)
            q.element2element = ptr
        elif q.residual_smoothing:
            raise ValueError("residual smoothing needs the element2element array")
        lib().or_freestream(C.byref(q))
        self.params = q
        self.n_elem, self.n_face, self.n_ghost = m.nElem, m.nFace, m.nGhost
        self.n_tot = m.nElem + m.nGhost
        self.state = {k: np.empty(self.n_tot) for k in FIELDS}
        self._st = OrState(*[self.state[k].ctypes.data_as(DP) for k in FIELDS])
        lib().or_init_state(C.byref(m), C.byref(q), C.byref(self._st))
        self.R = np.zeros(5 * self.n_elem)
        # explicit Euler's conservative state (FlowField's rho u, rho v, rho w, rho E arrays)
        self.W = np.empty(5 * self.n_elem)
        lib().or_init_conservative(C.byref(m), C.byref(q), C.byref(self._st), self.W.ctypes.data_as(DP))

    def set_state(self, st: dict) -> None:
        for k in FIELDS:
            self.state[k][:] = st[k]

    def update_bc(self):
        lib().or_update_bc(C.byref(self.mesh), C.byref(self.params), C.byref(self._st))

    def face_flux(self):
        F = np.empty(5 * self.n_face)
        FS = np.empty(5 * self.n_face)
        kind = np.empty(self.n_face, dtype=np.int32)
        lib().or_face_flux(C.byref(self.mesh), C.byref(self.params), C.byref(self._st), F.ctypes.data_as(DP),
                           FS.ctypes.data_as(DP), kind.ctypes.data_as(IP))
        return F, FS, kind

    def residual(self):
        lib().or_residual(C.byref(self.mesh), C.byref(self.params), C.byref(self._st), self.R.ctypes.data_as(DP))
        return self.R.copy()

    def timestep(self):
        dt = np.empty(self.n_elem)
        lib().or_timestep(C.byref(self.mesh), C.byref(self.params), C.byref(self._st), dt.ctypes.data_as(DP))
        return dt

    def rk5(self):
        """RungeKutta from the residual currently in self.R (as Run calls it)."""
        lib().or_rk5(C.byref(self.mesh), C.byref(self.params), C.byref(self._st), self.R.ctypes.data_as(DP))

    def iterate(self, n: int) -> np.ndarray:
        """n iterations of Run's body (updateBC, computeResidual, RungeKutta); R_rho^2 sums."""
        out = np.empty(n)
        for i in range(n):
            self.update_bc()
            self.residual()
            self.rk5()
            out[i] = lib().or_rms_sum(self.R.ctypes.data_as(DP), self.n_elem)
        return out

    def smooth(self) -> int:
        """smoothResiduals on self.R; the number of sweeps (-1: a NaN residual)."""
        return lib().or_smooth(C.byref(self.mesh), C.byref(self.params), self.R.ctypes.data_as(DP))

    def euler(self):
        """EulerExplicit from the residual in self.R."""
        lib().or_euler(C.byref(self.mesh), C.byref(self.params), C.byref(self._st), self.W.ctypes.data_as(DP),
                       self.R.ctypes.data_as(DP))

    def iterate_any(self, n: int) -> np.ndarray:
        """n iterations of Run's body with the configured integrator and smoothing."""
        out = np.empty(n)
        for i in range(n):
            self.update_bc()
            self.residual()
            if self.params.residual_smoothing and self.smooth() < 0:
                raise FloatingPointError("NaN residual in the smoothing")
            if self.params.time_integration == 2:
                if self.params.residual_smoothing:
                    raise NotImplementedError("use run() for RK5 with smoothing")
                self.rk5()
            elif self.params.time_integration == 0:
                self.euler()
            out[i] = lib().or_rms_sum(self.R.ctypes.data_as(DP), self.n_elem)
        return out

    def rms_sum(self) -> float:
        return lib().or_rms_sum(self.R.ctypes.data_as(DP), self.n_elem)

    def aero(self):
        cf = np.empty(3)
        cm = np.empty(3)
        lib().or_aero(C.byref(self.mesh), C.byref(self.params), C.byref(self._st), cf.ctypes.data_as(DP),
                      cm.ctypes.data_as(DP))
        return cf, cm

    def run(self, max_rows: int | None = None) -> dict:
        n = max_rows if max_rows is not None else self.params.max_iter + 1
        rows = np.zeros(8 * n)
        iters = C.c_int()
        co, cf, cm = np.empty(3), np.empty(3), np.empty(3)
        nrows = lib().or_run_any(C.byref(self.mesh), C.byref(self.params), C.byref(self._st),
                                 self.W.ctypes.data_as(DP), self.R.ctypes.data_as(DP), rows.ctypes.data_as(DP), n,
                                 C.byref(iters), co.ctypes.data_as(DP), cf.ctypes.data_as(DP), cm.ctypes.data_as(DP))
        if nrows < 0:
            raise FloatingPointError("NaN residual in the smoothing")
        k = min(nrows, n)
        return {"rows": rows[: 8 * k].reshape(k, 8), "iterations": iters.value, "CL": co[0], "CD": co[1],
                "CM": co[2], "force_coeff": cf, "moment_coeff": cm}
