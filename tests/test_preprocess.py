"""Host preprocessing, file formats and configuration (CPU).

Connectivity, ghost classification and metrics of the product are compared bit for bit with the
reference's own dumps (Connectivity.cpp, Mesh.hpp:471-707, Metrics.cpp); the .e3d reader with the
reference's parser test answers (test/Test_Parser.cpp:137-178); the SU2 writer with the layout
the reference's writer test pins (test/Test_SU2Writer.cpp:17-60).
"""
from __future__ import annotations

import hashlib
from pathlib import Path

import numpy as np
import pytest
from conftest import MESH_KEYS_F, MESH_KEYS_I, SMALL_CASES, golden_mesh, load_golden, write_par

from euler3d_amd import E3DError, Mesh, SimConfig


@pytest.mark.parametrize("case", SMALL_CASES)
def test_preprocess_matches_reference(case, tmp_path):
    g = load_golden(case)
    ref = golden_mesh(g)
    m = Mesh.read(str(write_par(case, tmp_path / "m.par"))).preprocess(float(g["params"][5]))
    info = m.info()
    assert [info["n_elem"], info["n_face"], info["n_face_int"]] == list(ref["counts"][:3])
    for k in MESH_KEYS_I:
        assert np.array_equal(m.get(k), ref[k]), k
    for k in MESH_KEYS_F:
        a, b = m.get(k), ref[k]
        assert a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64)), k


def test_generator_full_naca_is_deterministic(tmp_path):
    """The 129x129x2 O-grid the full-size fixture was made from is regenerated byte for byte."""
    g = load_golden("naca_full")
    m = Mesh.naca(ni=129, nj=129, nk=2, thickness=0.12, wall_spacing=1e-3, farfield_radius=20.0, dz=0.1)
    p = tmp_path / "naca.par"
    m.write(str(p), precision=16, par=True)
    assert hashlib.sha256(p.read_bytes()).hexdigest().encode() == bytes(g["par_sha256"])
    m.preprocess(1.0)
    info = m.info()
    # 128 x 128 x 1 hexes; faces 32,640 interior + 33,024 boundary (128 wall, 128 farfield, 32,768 sym)
    assert info["n_elem"] == 16384
    assert info["n_face"] == 65664 and info["n_face_int"] == 32640
    assert (info["n_wall"], info["n_far"], info["n_sym"]) == (128, 128, 32768)


def test_bump_counts_small():
    m = Mesh.bump(nx=8, ny=4, nz=4, jitter=0.1, seed=0, permute=True).preprocess(1.0)
    info = m.info()
    assert info["n_elem"] == 8 * 4 * 4 * 6
    # every tet has 4 faces: 4 N = 2 interior + boundary
    assert 4 * info["n_elem"] == 2 * info["n_face_int"] + info["n_ghost"]
    # two triangles per boundary quad: inlet/outlet 2*(ny*nz) quads, walls 2*(nx*nz + nx*ny)
    assert info["n_far"] == 2 * 2 * 4 * 4
    assert info["n_wall"] == 2 * 2 * (8 * 4 + 8 * 4)
    vol = m.get("cellVolumes")
    assert np.all(vol > 0)
    # the channel volume minus the bump's cross-section area (circular segment, h = 0.1, chord 1)
    R = (0.25 + 0.01) / 0.2
    seg = R * R * np.arcsin(0.5 / R) - 0.5 * (R - 0.1)
    assert abs(vol.sum() - (3.0 - seg)) < 5e-3


def test_closed_cells_sum_normals_to_zero():
    """Each cell's outward area vectors sum to zero (closed surface) for tets and planar hexes."""
    for m in (Mesh.bump(nx=6, ny=3, nz=3, jitter=0.1, seed=3, permute=True), Mesh.box(3, 2, 2)):
        m.preprocess(1.0)
        f2e = m.get("face2element").reshape(-1, 2)
        N = m.get("faceNormals").reshape(-1, 3)
        n_elem = m.info()["n_elem"]
        acc = np.zeros((n_elem, 3))
        np.add.at(acc, f2e[:, 0], N)
        inner = f2e[:, 1] < n_elem
        np.add.at(acc, f2e[inner, 1], -N[inner])
        assert np.max(np.abs(acc)) < 1e-12


def test_config_reader_reference_answers(tmp_path):
    """Values the reference's parser test expects for test/ConfigFiles/ConfigFIle_test.e3d."""
    text = """START
INITIAL_MESH= StructuredBlock_64x64.su2
PARTITION_FILES= 1
part_#.par
PRE_LOG= prelog.txt
MESH_ORIENTATION_CL= 5
MESH_ORIENTATION_CD= 1
MESH_REF_POINT_X= 1.0
MESH_REF_POINT_Y= 2.0
MESH_REF_POINT_Z= 3.0
SREF= 2.5
CREF= 0.5
SAMPLING= 500
SPEED_OPTION= 1
SPEED_VALUE= 250
# Options : ROE -> 0
AOA= 1
AIRFLOW_PRESSURE= 101325
AIRFLOW_TEMPERATURE= 288.15
AIRFLOW_VISCOSITY= 1.853e-5
AIRFLOW_DENSITY= 1.2886
GAMMA= 1.4
GAS_CONSTANT= 287.058
SPECIFIC_HEAT= 1004.7
SCHEME= 1
TIME_INTEGRATION= 2
CFL= 3
MIN_RESIDUAL= 1e-10
MAX_ITER= 99999
SOLVER_LOG= solverlog.txt
OUTPUT_FILE= out/output.dat
POST_LOG= solverlog.txt
END
"""
    p = tmp_path / "c.e3d"
    p.write_text(text)
    c = SimConfig.read(str(p))
    q = c.params
    assert abs(q.mach - 0.7346523151) < 1e-10
    assert (q.orient_cl[0], q.orient_cl[1]) == (2, -1)
    assert (q.orient_cd[0], q.orient_cd[1]) == (0, -1)
    assert (q.orient_cm[0], q.orient_cm[1]) == (1, -1)
    assert tuple(q.ref_point) == (1.0, 2.0, 3.0)
    assert (q.sref, q.cref, q.sampling) == (2.5, 0.5, 500)
    assert (q.cfl, q.min_residual, q.max_iter, q.scheme, q.time_integration) == (3.0, 1e-10, 99999, 1, 2)
    assert q.aoa_deg == 1.0 and q.gamma == 1.4
    assert c.mesh_path == "part_0.par"
    assert c.output_dir == "out"


def test_config_errors(tmp_path):
    p = tmp_path / "bad.e3d"
    p.write_text("START\nMESH_ORIENTATION_CL= 9\nMESH_ORIENTATION_CD= 0\nOUTPUT_FILE= a.dat\nEND\n")
    with pytest.raises(E3DError):
        SimConfig.read(str(p))
    p.write_text("START\nPARTITION_FILES= 2\np#.par\nMESH_ORIENTATION_CL= 2\nMESH_ORIENTATION_CD= 0\n"
                 "SPEED_VALUE= 0.5\nSAMPLING= 1\nCFL= 1\nMAX_ITER= 1\nOUTPUT_FILE= a.dat\nEND\n")
    c = SimConfig.read(str(p))  # several partitions: every file of the pattern (SimConfig.cpp:66-87)
    assert c.partition_files == ("p0.par", "p1.par") and c.mesh_path == "p0.par"
    with pytest.raises(E3DError):
        SimConfig.read(str(tmp_path / "missing.e3d"))


def test_su2_writer_layout(tmp_path):
    """SU2Writer layout (Test_SU2Writer.cpp:17-60): headers, setw(20) %e coordinates, node index."""
    m = Mesh.box(1, 1, 1)
    p = tmp_path / "b.su2"
    m.write(str(p), precision=3, par=False)
    lines = p.read_text().split("\n")
    i = lines.index("NPOIN= 8")
    assert lines[i - 3:i] == ["%", "% Node coordinates", "%"]
    assert lines[i + 2].split() == ["1.000e+00", "0.000e+00", "0.000e+00", "1"]
    assert lines[i + 1] == f"{'0.000e+00':>20}{'0.000e+00':>20}{'0.000e+00':>20}{0:>10}"
    j = lines.index("NELEM= 1")
    assert lines[j + 1].split() == ["12", "0", "1", "3", "2", "4", "5", "7", "6", "0"]
    assert "NMARK= 1" in lines and "MARKER_TAG= farfield" in lines and "MARKER_ELEMS= 6" in lines
    assert "NINTERNAL_MARKER= 0" not in lines
    m.write(str(p), precision=16, par=True)
    assert "NINTERNAL_MARKER= 0" in p.read_text().split("\n")
    # round trip at full precision
    m2 = Mesh.read(str(p))
    assert np.array_equal(m2.get("xyz"), m.get("xyz"))


def test_reader_errors(tmp_path):
    p = tmp_path / "x.su2"
    p.write_text("NDIME= 2\n")
    with pytest.raises(E3DError):
        Mesh.read(str(p), par=False)
    p.write_text("NDIME= 3\nNELEM= 0\nNPOIN= 0\nNMARK= 0\n")
    with pytest.raises(E3DError, match="NINTERNAL_MARKER"):
        Mesh.read(str(p), par=True)


def test_unknown_boundary_tag_rejected():
    m = Mesh.box(2, 2, 2, tags=("farfield", "farfield", "wall", "wall", "sym1", "sym"))
    with pytest.raises(E3DError, match="sym1"):
        m.preprocess(1.0)


@pytest.mark.parametrize("dims", [(24, 12, 9), (25, 13, 9), (30, 16, 11), (33, 17, 12)])
def test_wing_generator_cells_are_closed(dims):
    """Every generated wing cell is closed (sum of n S over its faces vanishes) and every node lies
    in the box, for odd and even node counts."""
    ni, nj, nk = dims
    m = Mesh.wing(ni=ni, nj=nj, nk=nk).preprocess(1.0)
    info = m.info()
    n = info["n_elem"]
    xyz = m.get("xyz").reshape(-1, 3)
    assert np.all(np.isfinite(xyz)) and np.all(np.abs(xyz[:, 1]) <= 15.0 + 1e-12)
    f2e = m.get("face2element").reshape(-1, 2)
    v = m.get("faceNormalsUnit").reshape(-1, 3) * m.get("faceSurfaces")[:, None]
    C = np.zeros((n, 3))
    A = np.zeros(n)
    np.add.at(C, f2e[:, 0], v)
    inner = f2e[:, 1] < n
    np.subtract.at(C, f2e[inner, 1], v[inner])
    np.add.at(A, f2e[:, 0], m.get("faceSurfaces"))
    assert np.max(np.linalg.norm(C, axis=1) / A) < 1e-12
    assert np.all(m.get("cellVolumes") > 0)

