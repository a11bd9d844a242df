"""Partitioned input: several .par files read as one mesh, against the reference's own ranks.

Fixtures (tests/golden/<case>.npz, made by tests/golden/make_golden.py) hold, per MPI rank r of a
`mpiexec -np N` run of the reference harness, that rank's connectivity, ghost lists and metrics
(r<r>_mesh_*), seeded states before and after updateBC (r<r>_flux<seed>_*), and the final state
of a run (r<r>_run_*). Here the product's reader and preprocessing (host only) are checked against
them: every partition numbered as its rank numbers it, and every interface face pointed at the
cell whose state the reference's ghost exchange delivers to that face's ghost.
"""
from __future__ import annotations

import gzip
from pathlib import Path

import numpy as np
import pytest
from conftest import GOLDEN, MESH_KEYS_F, load_golden

from euler3d_amd import Mesh

PART_CASES = ("block8_p4", "naca_small_p3", "bump_small_p4", "hybrid_small_p3")
BND_MPI = 4


def write_parts(name: str, d: Path) -> list[Path]:
    paths = []
    r = 0
    while (GOLDEN / f"{name}.p{r}.par.gz").exists():
        p = d / f"{name}_p{r}.par"
        p.write_bytes(gzip.decompress((GOLDEN / f"{name}.p{r}.par.gz").read_bytes()))
        paths.append(p)
        r += 1
    return paths


def composite(name: str, d: Path) -> Mesh:
    return Mesh.read_partitions([str(p) for p in write_parts(name, d)]).preprocess(1.0)


def to_local(parts, n_elem, cell):
    """composite cell/ghost id -> (partition, local id in that partition's reference numbering)"""
    if cell < n_elem:
        for r, q in enumerate(parts):
            if q["elem0"] <= cell < q["elem0"] + q["n_elem"]:
                return r, cell - q["elem0"]
    k = cell - n_elem
    for r, q in enumerate(parts):
        if q["ghost0"] <= k < q["ghost0"] + q["n_ghost"]:
            return r, q["n_elem"] + k - q["ghost0"]
    raise AssertionError(cell)


@pytest.mark.parametrize("case", PART_CASES + ("bump_small_p4_smooth", "wing_small_p2_euler_smooth"))
def test_device_order_keeps_partition_ranges(case, tmp_path):
    """The solver's device cell order (bisection blocks, cells read by other blocks grouped last in
    each block) is a permutation that leaves every partition in its own range of device positions:
    the per-partition smoothing decisions and coefficient sums reduce over those ranges. With
    288-cell partitions some 256-cell blocks straddle two partitions."""
    m = composite(case, tmp_path)
    parts = m.partitions()
    order = m.get("deviceOrder")
    n = m.info()["n_elem"]
    assert np.array_equal(np.sort(order), np.arange(n))
    for q in parts:
        lo, hi = q["elem0"], q["elem0"] + q["n_elem"]
        seg = order[lo:hi]
        assert seg.min() >= lo and seg.max() < hi, (case, q)


@pytest.mark.parametrize("case", PART_CASES)
def test_partitions_numbered_as_each_rank(case, tmp_path):
    g = load_golden(case)
    m = composite(case, tmp_path)
    parts = m.partitions()
    info = m.info()
    nE, nFI = info["n_elem"], info["n_face_int"]
    assert len(parts) == int(g["parts"][0])
    f2e = m.get("face2element").reshape(-1, 2)
    f2nS, f2n = m.get("face2nodeStart"), m.get("face2node")
    e2fS, e2f, e2e = m.get("element2faceStart"), m.get("element2face"), m.get("element2element")
    kind = m.get("ghostKind")
    arrays = {k: m.get(k) for k in ("wallGhost", "wallCell", "wallFace", "farGhost", "farCell", "farFace", "symGhost",
                                    "symCell", "symFace", "mpiGhost", "mpiCell")}
    metrics = {k: m.get(k) for k in MESH_KEYS_F}
    for r, q in enumerate(parts):
        ref = {k[len(f"r{r}_mesh_"):]: v for k, v in g.items() if k.startswith(f"r{r}_mesh_")}
        ne, nfi, ng = q["n_elem"], q["n_face_int"], q["n_ghost"]
        assert ref["counts"][0] == ne and ref["counts"][2] == nfi and ref["counts"][1] == nfi + ng

        def face_g(f):
            return q["face_int0"] + f if f < nfi else nFI + q["ghost0"] + (f - nfi)

        def cell_g(c):
            return q["elem0"] + c if c < ne else nE + q["ghost0"] + (c - ne)

        faces = np.array([face_g(f) for f in range(nfi + ng)])
        # face2element: interior faces and non-interface boundary faces as the rank numbers them
        ref_f2e = ref["face2element"].reshape(-1, 2)
        ours = f2e[faces]
        is_mpi = np.zeros(nfi + ng, dtype=bool)
        is_mpi[nfi:] = kind[q["ghost0"]:q["ghost0"] + ng] == BND_MPI
        np.testing.assert_array_equal(ours[:, 0], [cell_g(c) for c in ref_f2e[:, 0]])
        np.testing.assert_array_equal(ours[~is_mpi, 1], [cell_g(c) for c in ref_f2e[~is_mpi, 1]])
        # interface faces now read a cell of another partition
        for f in np.nonzero(is_mpi)[0]:
            s, _ = to_local(parts, nE, ours[f, 1])
            assert s != r and ours[f, 1] < nE
        # face nodes in local numbering, cell -> face and cell -> cell lists
        for f in range(nfi + ng):
            gf = faces[f]
            loc = f2n[f2nS[gf]:f2nS[gf + 1]] - q["node0"]
            np.testing.assert_array_equal(loc, ref["face2node"][ref["face2nodeStart"][f]:ref["face2nodeStart"][f + 1]])
        sl = slice(e2fS[q["elem0"]], e2fS[q["elem0"] + ne])
        np.testing.assert_array_equal(e2f[sl], [face_g(f) for f in ref["element2face"]])
        np.testing.assert_array_equal(e2e[sl], [cell_g(c) for c in ref["element2element"]])
        # ghost lists, interface ghost list, metrics
        for k, conv in (("Ghost", cell_g), ("Cell", cell_g), ("Face", face_g)):
            for kind_name in ("wall", "far", "sym", "mpi"):
                key = kind_name + k
                if key not in ref:
                    continue
                ours_l = [conv(x) for x in ref[key]]
                lo = {"wall": q["wall0"]}.get(kind_name)
                if kind_name == "wall":
                    np.testing.assert_array_equal(arrays[key][q["wall0"]:q["wall0"] + q["n_wall"]], ours_l)
                else:
                    assert set(ours_l) <= set(arrays[key].tolist()), key
                    assert lo is None
        for k, v in metrics.items():
            width = 3 if v.size in (3 * len(f2e), 3 * nE) else 1
            if k in ("cellVolumes", "cellCentroids"):
                idx = np.arange(q["elem0"], q["elem0"] + ne)
            else:
                idx = faces
            mine = v.reshape(-1, width)[idx].ravel()
            np.testing.assert_array_equal(mine.view(np.uint64), ref[k].view(np.uint64), err_msg=f"{k} r{r}")


@pytest.mark.parametrize("case", PART_CASES)
def test_interface_faces_read_the_exchanged_cell(case, tmp_path):
    """After updateBC, the reference's interface ghosts hold the states the pairwise exchange
    delivered (MPIHandler.cpp:84-267). With seeded, all-different cell states, each such ghost
    must equal the cell the composite's interface face reads: that pins the pairing, including
    cells with several faces toward the same partition."""
    g = load_golden(case)
    m = composite(case, tmp_path)
    parts = m.partitions()
    nE, nFI = m.info()["n_elem"], m.info()["n_face_int"]
    f2e = m.get("face2element").reshape(-1, 2)
    seeds = sorted({int(k.split("_")[1][4:]) for k in g if k.startswith("r0_flux") and k.endswith("_in_rho")})
    assert seeds
    checked = 0
    for sd in seeds:
        for r, q in enumerate(parts):
            mg = g[f"r{r}_mesh_mpiGhost"]
            for gl in mg:
                f = nFI + q["ghost0"] + (gl - q["n_elem"])
                s, loc = to_local(parts, nE, f2e[f, 1])
                for fld in ("rho", "u", "v", "w", "E", "p", "H"):
                    got = g[f"r{r}_flux{sd}_bc_{fld}"][gl]
                    want = g[f"r{s}_flux{sd}_in_{fld}"][loc]
                    assert got.view(np.uint64) == want.view(np.uint64), (sd, r, gl, fld)
                checked += 1
    assert checked > 0


def test_partition_writer_against_reference_e3d_pre(tmp_path):
    """E3D_PRE's own 4-partition output of the 8-hex block (test/mesh/StructuredBlock_8_p*.par,
    kept as the block8_p4 fixture): with the same element-to-partition assignment, the writer
    gives the same elements, local node numbering, coordinates, markers and interface blocks (the
    reference files predate the current SU2Writer layout, so values are compared, not bytes)."""
    ref_files = write_parts("block8_p4", tmp_path)
    ref_parsed = [parse_par(p) for p in ref_files]
    # the global mesh: the reference's StructuredBlock_8 (the block8 fixture)
    src = tmp_path / "block8.par"
    src.write_bytes(gzip.decompress((GOLDEN / "block8.par.gz").read_bytes()))
    gm = Mesh.read(str(src)).preprocess(1.0)
    gxyz = gm.get("xyz").reshape(-1, 3)
    gel = gm.get("elemNodes").reshape(-1, 8)
    # assignment: each reference partition element -> the global element with the same nodes
    key = {tuple(sorted(map(tuple, gxyz[e]))): i for i, e in enumerate(gel)}
    assign = np.full(len(gel), -1, dtype=np.int32)
    for r, pp in enumerate(ref_parsed):
        for el in pp["elems"]:
            nodes = [pp["xyz"][k] for k in el[1:9]]
            assign[key[tuple(sorted(map(tuple, nodes)))]] = r
    assert (assign >= 0).all()
    out = [tmp_path / f"ours_p{r}.par" for r in range(4)]
    gm.write_partitions([str(p) for p in out], elem2part=assign, precision=16)
    for r in range(4):
        a, b = ref_parsed[r], parse_par(out[r])
        assert a["elems"] == b["elems"], r
        np.testing.assert_allclose(a["xyz"], b["xyz"], rtol=0, atol=1e-15)
        # the reference files list marker and interface faces in another order and rotation than
        # the current Partition.cpp does: compare the face sets
        assert face_sets(a["markers"]) == face_sets(b["markers"]), r
        assert face_sets(a["iface"]) == face_sets(b["iface"]), r


def face_sets(blocks: dict) -> dict:
    """{key: sorted faces}, a face as its leading ids and its node set (layout-independent)"""
    out = {}
    for k, rows in blocks.items():
        faces = []
        for row in rows:
            lead = row[:3] if len(row) in (6, 7) and row[2] in (5, 9) else row[:1]
            faces.append((tuple(lead), tuple(sorted(row[len(lead):]))))
        out[k] = sorted(faces)
    return out


def parse_par(path: Path) -> dict:
    """Values of an SU2/.par file, independent of its column layout."""
    lines = [ln.strip() for ln in path.read_text().splitlines()]
    out = {"elems": [], "xyz": [], "markers": {}, "iface": {}}
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("NELEM="):
            n = int(ln.split("=")[1])
            out["elems"] = [tuple(int(x) for x in lines[i + 1 + k].split()[:9]) for k in range(n)]
            i += n
        elif ln.startswith("NPOIN="):
            n = int(ln.split("=")[1])
            out["xyz"] = np.array([[float(x) for x in lines[i + 1 + k].split()[:3]] for k in range(n)])
            i += n
        elif ln.startswith("MARKER_TAG="):
            tag = ln.split("=")[1].strip()
            n = int(lines[i + 1].split("=")[1])
            out["markers"][tag] = [tuple(int(x) for x in lines[i + 2 + k].split()) for k in range(n)]
            i += n + 1
        elif ln.startswith("INTERNAL_MARKER="):
            pid = int(ln.split("=")[1])
            n = int(lines[i + 1].split("=")[1])
            out["iface"][pid] = [tuple(int(x) for x in lines[i + 2 + k].split()) for k in range(n)]
            i += n + 1
        i += 1
    return out


@pytest.mark.parametrize("n_part", [2, 3, 5])
def test_partitioner_round_trip(n_part, tmp_path):
    """Splitting then reading back gives every cell once, consistent interface blocks, and the
    same faces, volumes and boundary lists as the single mesh (up to numbering)."""
    src = tmp_path / "bump.par"
    src.write_bytes(gzip.decompress((GOLDEN / "bump_small.par.gz").read_bytes()))
    gm = Mesh.read(str(src)).preprocess(1.0)
    assign = gm.partition_assign(n_part)
    assert sorted(np.bincount(assign, minlength=n_part)) == sorted(np.bincount(assign, minlength=n_part))
    assert np.bincount(assign).max() - np.bincount(assign).min() <= 1
    paths = [tmp_path / f"b_p{r}.par" for r in range(n_part)]
    gm.write_partitions([str(p) for p in paths])
    cm = Mesh.read_partitions([str(p) for p in paths]).preprocess(1.0)
    gi, ci = gm.info(), cm.info()
    assert ci["n_elem"] == gi["n_elem"]
    assert (ci["n_wall"], ci["n_far"], ci["n_sym"]) == (gi["n_wall"], gi["n_far"], gi["n_sym"])
    n_mpi = int((cm.get("ghostKind") == BND_MPI).sum())
    # each cut face is a boundary face in both partitions
    assert ci["n_face_int"] + n_mpi // 2 == gi["n_face_int"]
    assert n_mpi % 2 == 0
    np.testing.assert_allclose(np.sort(cm.get("cellVolumes")), np.sort(gm.get("cellVolumes")), rtol=1e-9)


# ---- the oracle on the composite mesh against the reference's ranks

FIELDS8 = ("rho", "u", "v", "w", "p", "E", "H", "M")


def composite_arrays(m: Mesh) -> dict:
    """Oracle input for a partitioned mesh (the product's preprocessing)."""
    from conftest import MESH_KEYS_I

    arr = {k: m.get(k) for k in MESH_KEYS_I + MESH_KEYS_F}
    parts = m.partitions()
    arr["nFaceInt"] = m.info()["n_face_int"]
    arr["mpiGhost"], arr["mpiCell"] = m.get("mpiGhost"), m.get("mpiCell")
    arr["partElemStart"] = m.get("partStart")
    arr["partWallStart"] = np.array([q["wall0"] for q in parts] + [parts[-1]["wall0"] + parts[-1]["n_wall"]])
    return arr


def join_cells(g: dict, key: str, parts) -> np.ndarray:
    """Per-rank arrays over (interior cells, ghosts) -> the composite's (all interiors, all ghosts)."""
    inner = [g[f"r{r}_{key}"][:q["n_elem"]] for r, q in enumerate(parts)]
    ghost = [g[f"r{r}_{key}"][q["n_elem"]:q["n_elem"] + q["n_ghost"]] for r, q in enumerate(parts)]
    return np.concatenate(inner + ghost)


def join_interior(g: dict, key: str, parts, width: int = 1) -> np.ndarray:
    return np.concatenate([g[f"r{r}_{key}"][:width * q["n_elem"]] for r, q in enumerate(parts)])


def split_faces(values: np.ndarray, parts, n_face_int: int, width: int) -> list:
    """composite per-face values -> per-rank arrays in each rank's face order"""
    v = values.reshape(-1, width)
    out = []
    for q in parts:
        idx = np.concatenate([np.arange(q["face_int0"], q["face_int0"] + q["n_face_int"]),
                              n_face_int + q["ghost0"] + np.arange(q["n_ghost"])])
        out.append(v[idx].ravel())
    return out


def oracle_params(g: dict) -> dict:
    from conftest import integrator_params

    return integrator_params(g)


@pytest.mark.parametrize("case", PART_CASES)
def test_oracle_partitioned_flux_and_step(case, tmp_path):
    from oracle_ctypes import Oracle

    g = load_golden(case)
    m = composite(case, tmp_path)
    parts = m.partitions()
    nFI = m.info()["n_face_int"]
    mpi_ghost = set(m.get("mpiGhost").tolist())
    nE = m.info()["n_elem"]
    seeds = sorted({int(k.split("_")[1][4:]) for k in g if k.startswith("r0_flux") and k.endswith("_in_rho")})
    for sd in seeds:
        o = Oracle(composite_arrays(m), oracle_params(g))
        o.set_state({k: join_cells(g, f"flux{sd}_in_{k}", parts) for k in FIELDS8})
        o.update_bc()
        keep = np.array([i for i in range(o.n_tot) if i not in mpi_ghost])
        for k in FIELDS8:
            want = join_cells(g, f"flux{sd}_bc_{k}", parts)
            np.testing.assert_array_equal(o.state[k][keep].view(np.uint64), want[keep].view(np.uint64), err_msg=k)
        F, FS, kind = o.face_flux()
        for r, (f, fs, kd) in enumerate(zip(split_faces(F, parts, nFI, 5), split_faces(FS, parts, nFI, 5),
                                            split_faces(kind, parts, nFI, 1))):
            np.testing.assert_array_equal(kd, g[f"r{r}_flux{sd}_face_kind"])
            np.testing.assert_array_equal(f.view(np.uint64), g[f"r{r}_flux{sd}_face_flux"].view(np.uint64))
            np.testing.assert_array_equal(fs.view(np.uint64), g[f"r{r}_flux{sd}_face_flux_scaled"].view(np.uint64))
        R = o.residual()
        np.testing.assert_array_equal(R.view(np.uint64), join_interior(g, f"flux{sd}_residual", parts, 5).view(np.uint64))
        dt = o.timestep()
        np.testing.assert_array_equal(dt.view(np.uint64), join_interior(g, f"flux{sd}_dt", parts).view(np.uint64))
        o.rk5()
        for k in FIELDS8:
            want = join_cells(g, f"flux{sd}_rk_{k}", parts)
            np.testing.assert_array_equal(o.state[k].view(np.uint64), want.view(np.uint64), err_msg=f"rk {k}")
        total = sum(float(g[f"r{r}_flux{sd}_rk_rms_sum"][0]) for r in range(len(parts)))
        assert abs(o.rms_sum() - total) <= 1e-13 * abs(total)
    assert nE == sum(q["n_elem"] for q in parts)


@pytest.mark.parametrize("case", PART_CASES + ("bump_small_p4_smooth", "wing_small_p2_euler_smooth"))
def test_oracle_partitioned_run(case, tmp_path):
    """A full multi-rank run: final states bit for bit (per partition), the coefficients summed
    per partition in rank order, the residual history to the RMS tolerance."""
    from oracle_ctypes import Oracle

    g = load_golden(case)
    m = composite(case, tmp_path)
    parts = m.partitions()
    pd = oracle_params(g)
    arr = composite_arrays(m)
    o = Oracle(arr, pd)
    out = o.run()
    for k in FIELDS8:
        want = join_cells(g, f"run_{k}", parts)
        np.testing.assert_array_equal(o.state[k].view(np.uint64), want.view(np.uint64), err_msg=k)
    assert out["iterations"] == int(g["r0_run_iterations"][0])
    for k in ("CL", "CD", "CM"):
        ref = float(g[f"r0_run_{k}"][0])
        assert out[k] == ref or abs(out[k] - ref) <= 1e-13 * max(abs(ref), 1e-12), (k, out[k], ref)
    rd = g["run_residuals_dat"]
    assert out["rows"].shape == rd.shape
    np.testing.assert_allclose(out["rows"][:, 1:], rd[:, 1:], rtol=6e-11, atol=1e-13)


def test_e3d_pre_executable(tmp_path):
    """`E3D_PRE <case.e3d>` (src/preprocessor.cpp:16-45): INITIAL_MESH split into the
    PARTITION_FILES, the SU2Writer layout at its default 10 digits; the same bytes as the library
    writer, and the files read back as one mesh of the same cells."""
    import subprocess
    import sys

    from conftest import ROOT

    sys.path.insert(0, str(GOLDEN))
    from make_golden import CONFIG

    src = tmp_path / "naca.su2"
    src.write_bytes(gzip.decompress((GOLDEN / "naca_small.par.gz").read_bytes()))
    for n in (1, 3):
        d = tmp_path / f"n{n}"
        d.mkdir()
        cfg = d / "case.e3d"
        cfg.write_text(CONFIG.format(min_residual=0, min_aero=0, mesh=str(src), par=str(d / "part_#.par"), out=str(d), ocl=2, ocd=0, rx=0.0, ry=0.0,
                                     rz=0.0, sref=1.0, cref=1.0, sampling=1, mach=0.5, aoa=0.0, cfl=1.0, iters=1, ti=2,
                                     smooth=0).replace("PARTITION_FILES= 1", f"PARTITION_FILES= {n}"))
        r = subprocess.run([str(ROOT / "euler3d_amd" / "bin" / "E3D_PRE"), str(cfg)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        gm = Mesh.read(str(src)).preprocess(1.0)
        lib_paths = [d / f"lib_{k}.par" for k in range(n)]
        gm.write_partitions([str(p) for p in lib_paths], precision=10)
        for k in range(n):
            assert (d / f"part_{k}.par").read_bytes() == lib_paths[k].read_bytes()
        if n == 1:
            assert "NINTERNAL_MARKER= 0\n" in (d / "part_0.par").read_text()
            assert Mesh.read(str(d / "part_0.par")).preprocess(1.0).info()["n_elem"] == gm.info()["n_elem"]
        else:
            cm = Mesh.read_partitions([str(d / f"part_{k}.par") for k in range(n)]).preprocess(1.0)
            assert cm.info()["n_elem"] == gm.info()["n_elem"]
    r = subprocess.run([str(ROOT / "euler3d_amd" / "bin" / "E3D_PRE")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Usage" in r.stderr


@pytest.mark.parametrize("case", PART_CASES)
def test_interface_exchange_plan(case, tmp_path):
    """The exchange lists (one message per sending/receiving partition pair, MPIHandler::
    updateFlowField): every interface ghost receives exactly once, from the cell that the composite's
    interface face reads (the pairing of the reference's two sorted lists, pinned by the fixtures
    through that face); each message's cells belong to its sender and its ghosts to its receiver;
    the two directions of a pair carry the same number of values, as MPI_Sendrecv requires."""
    m = composite(case, tmp_path)
    info = m.info()
    ne, nfi = info["n_elem"], info["n_face_int"]
    f2e = m.get("face2element").reshape(-1, 2)
    start, frm, to = m.get("haloStart"), m.get("haloFrom"), m.get("haloTo")
    src, dst = m.get("haloSrc"), m.get("haloDst")
    mpi = m.get("mpiGhost")
    part_start = m.get("partStart")
    parts = m.partitions()
    assert len(start) == len(frm) + 1 == len(to) + 1 and start[0] == 0 and start[-1] == len(src) == len(dst)
    assert sorted(dst.tolist()) == sorted(mpi.tolist())
    for g, c in zip(dst, src):
        assert f2e[nfi + (g - ne), 1] == c
    owner = lambda cell: int(np.searchsorted(part_start, cell, side="right") - 1)  # noqa: E731
    sizes = {}
    for k in range(len(frm)):
        a, b = start[k], start[k + 1]
        assert b > a
        assert all(owner(c) == frm[k] for c in src[a:b])
        r = parts[to[k]]
        assert all(r["ghost0"] <= g - ne < r["ghost0"] + r["n_ghost"] for g in dst[a:b])
        sizes[(int(frm[k]), int(to[k]))] = b - a
    for (s, r), n in sizes.items():
        if s != r:
            assert sizes[(r, s)] == n, (s, r)
