// Mesh preprocessing: face numbering, ghost cells, ghost classification and FP64 metrics.
//
// The result is numbered exactly as the reference numbers it, because the per-cell residual
// accumulation order (ascending face id) and the face orientation (elem0 = owner) decide the FP64
// bits of every residual:
//   * local faces of each cell type                include/connectivity/Connectivity.hpp:37-45
//   * interior faces: created in (owner cell, owner local face) order, owner = the lower cell id;
//     the face's nodes are the owner's local face nodes  src/connectivity/Connectivity.cpp:110-210
//   * ghost cells and boundary faces: the unmatched (cell, local face) slots in that same flat
//     order, ghost k = nElem + k, boundary face nFaceInt + k   Connectivity.cpp:212-260
//   * wall / farfield / symmetry lists by marker order, matched on sorted face-node sets
//                                                            include/mesh/Mesh.hpp:594-707
//   * metrics and normal reorientation                       src/mesh/Metrics.cpp:9-309
// The reference finds neighbours with a per-face O(nNode) scratch vector and matches ghosts with
// a quadratic scan (Connectivity.cpp:155-156, Mesh.hpp:594-707). Here both are a sort of face
// keys, O(N log N), which gives the same matches on conforming meshes.
#include "e3d_host.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace e3d {

namespace {

struct LocalFaces {
	int nFace;
	int nNode[6];
	int node[6][4];
};

// Connectivity.hpp:37-45 (tet, hex, wedge, pyramid)
const LocalFaces kTet = {4, {3, 3, 3, 3}, {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}}};
const LocalFaces kHex = {6, {4, 4, 4, 4, 4, 4}, {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}}};
const LocalFaces kWedge = {5, {3, 4, 4, 4, 3}, {{0, 1, 2}, {0, 2, 5, 3}, {0, 3, 4, 1}, {1, 4, 5, 2}, {3, 5, 4}}};
const LocalFaces kPyr = {5, {4, 3, 3, 3, 3}, {{0, 3, 2, 1}, {0, 1, 4}, {1, 2, 4}, {2, 3, 4}, {3, 0, 4}}};

const LocalFaces &facesOf(int vtk) {
	switch (vtk) {
		case VTK_TET: return kTet;
		case VTK_HEX: return kHex;
		case VTK_WEDGE: return kWedge;
		case VTK_PYRAMID: return kPyr;
		default: throw std::runtime_error("VTK not recognised! (cell type " + std::to_string(vtk) + ")");
	}
}

struct FaceKey {
	std::array<int, 4> k; // sorted node ids, -1 padded at the end for triangles
	int64_t slot;         // element2face slot (or marker element index)
	bool operator<(const FaceKey &o) const { return k < o.k || (k == o.k && slot < o.slot); }
};

std::array<int, 4> sortedKey(const int *nodes, int n) {
	std::array<int, 4> k{-1, -1, -1, -1};
	for (int i = 0; i < n; i++) k[i] = nodes[i];
	std::sort(k.begin(), k.begin() + n);
	return k;
}

// ---- Vector3 arithmetic in the reference's evaluation order (include/utils/Vector3.h:18-125)
struct V3 {
	double x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 dv(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Metrics.cpp:60-62
inline double triArea(V3 a, V3 b) { return length(mul(cross(a, b), 0.5)); }

// Metrics.cpp:311-331 (Tartaglia / Cayley-Menger on edge lengths)
double tetVolume(V3 p0, V3 p1, V3 p2, V3 p3) {
	double U = length(sub(p0, p1));
	double V = length(sub(p1, p2));
	double W = length(sub(p2, p0));
	double u = length(sub(p2, p3));
	double v = length(sub(p0, p3));
	double w = length(sub(p3, p1));
	double uP = u * u, vP = v * v, wP = w * w, UP = U * U, VP = V * V, WP = W * W;
	double X = vP + wP - UP, Y = wP + uP - VP, Z = uP + vP - WP;
	double a = 4 * (uP * vP * wP) - uP * (X * X) - vP * (Y * Y) - wP * (Z * Z) + X * Y * Z;
	double vol = std::sqrt(a);
	vol /= 12.0;
	return vol;
}
inline V3 tetCentroid(V3 a, V3 b, V3 c, V3 d) { return dv(add(add(add(a, b), c), d), 4.0); }

std::string lower(std::string s) {
	for (auto &ch : s) ch = (char) std::tolower((unsigned char) ch);
	return s;
}

}// namespace

int localFaceCount(int vtk) { return facesOf(vtk).nFace; }
int localFaceNodes(int vtk, int f, int out[4]) {
	const LocalFaces &lf = facesOf(vtk);
	for (int i = 0; i < lf.nNode[f]; i++) out[i] = lf.node[f][i];
	return lf.nNode[f];
}

SolverMesh preprocess(const MeshFile &m, double cref) {
	SolverMesh s;
	const int nElem = m.nElem();
	s.nElem = nElem;

	// ---- element2face slots
	s.element2faceStart.resize(nElem + 1, 0);
	for (int e = 0; e < nElem; e++) s.element2faceStart[e + 1] = s.element2faceStart[e] + facesOf(m.elemVtk[e]).nFace;
	const int nSlot = s.element2faceStart[nElem];

	// ---- neighbour of every slot by sorting sorted-node keys
	std::vector<FaceKey> keys(nSlot);
	for (int e = 0; e < nElem; e++) {
		const LocalFaces &lf = facesOf(m.elemVtk[e]);
		const int *en = &m.elemNodes[m.elemStart[e]];
		for (int f = 0; f < lf.nFace; f++) {
			int nodes[4];
			for (int i = 0; i < lf.nNode[f]; i++) nodes[i] = en[lf.node[f][i]];
			int slot = s.element2faceStart[e] + f;
			keys[slot] = {sortedKey(nodes, lf.nNode[f]), slot};
		}
	}
	std::vector<int> slotElem(nSlot);
	for (int e = 0; e < nElem; e++)
		for (int j = s.element2faceStart[e]; j < s.element2faceStart[e + 1]; j++) slotElem[j] = e;
	std::vector<int> partner(nSlot, -1);
	{
		std::vector<FaceKey> sorted = keys;
		std::sort(sorted.begin(), sorted.end());
		for (int i = 0; i < nSlot;) {
			int j = i + 1;
			while (j < nSlot && sorted[j].k == sorted[i].k) j++;
			if (j - i == 2) {
				partner[sorted[i].slot] = (int) sorted[i + 1].slot;
				partner[sorted[i + 1].slot] = (int) sorted[i].slot;
			} else if (j - i > 2) {
				throw std::runtime_error("non-manifold mesh: a face is shared by more than two cells");
			}
			i = j;
		}
	}

	// ---- interior faces in creation order (Connectivity.cpp:110-210)
	s.element2face.assign(nSlot, -1);
	s.element2element.assign(nSlot, -1);
	int faceCount = 0;
	std::vector<int> f2e;
	std::vector<int> f2nStart{0};
	std::vector<int> f2n;
	f2e.reserve(nSlot + 2);
	f2n.reserve((size_t) nSlot * 2);
	for (int e = 0; e < nElem; e++) {
		const LocalFaces &lf = facesOf(m.elemVtk[e]);
		const int *en = &m.elemNodes[m.elemStart[e]];
		for (int f = 0; f < lf.nFace; f++) {
			int slot = s.element2faceStart[e] + f;
			int ps = partner[slot];
			if (ps < 0) continue;
			int nb = slotElem[ps];
			s.element2element[slot] = nb;
			if (nb < e) {
				s.element2face[slot] = s.element2face[ps];
			} else {
				s.element2face[slot] = faceCount;
				s.element2face[ps] = faceCount;
				f2e.push_back(e);
				f2e.push_back(nb);
				for (int i = 0; i < lf.nNode[f]; i++) f2n.push_back(en[lf.node[f][i]]);
				f2nStart.push_back((int) f2n.size());
				faceCount++;
			}
		}
	}
	s.nFaceInt = faceCount;

	// ---- ghost cells and boundary faces (Connectivity.cpp:212-260)
	int ghost = nElem;
	for (int j = 0; j < nSlot; j++) {
		if (s.element2element[j] != -1) continue;
		int e = slotElem[j];
		int f = j - s.element2faceStart[e];
		const LocalFaces &lf = facesOf(m.elemVtk[e]);
		const int *en = &m.elemNodes[m.elemStart[e]];
		s.element2element[j] = ghost;
		s.element2face[j] = faceCount;
		f2e.push_back(e);
		f2e.push_back(ghost);
		for (int i = 0; i < lf.nNode[f]; i++) f2n.push_back(en[lf.node[f][i]]);
		f2nStart.push_back((int) f2n.size());
		faceCount++;
		ghost++;
	}
	s.nFace = faceCount;
	s.nGhost = ghost - nElem;
	s.face2element = std::move(f2e);
	s.face2nodeStart = std::move(f2nStart);
	s.face2node = std::move(f2n);

	if (m.nBoundaryElems() + m.nInterfaceElems() != s.nGhost)
		throw std::runtime_error("boundary marker and interface elements (" +
		                         std::to_string(m.nBoundaryElems() + m.nInterfaceElems()) +
		                         ") do not match the mesh's boundary faces (" + std::to_string(s.nGhost) + ")");

	// ---- partition interfaces (Mesh::updateMPIGhostCells, Mesh.hpp:471-584): an interface
	// element's ghost is its cell's only ghost neighbour, or, when the cell has several, the one
	// whose face has the element's nodes
	std::vector<uint8_t> inMpi(s.nGhost, 0);
	s.interfaceGhost.resize(m.interfaces.size());
	for (size_t im = 0; im < m.interfaces.size(); im++) {
		const InterfaceMarker &mk = m.interfaces[im];
		s.interfaceGhost[im].assign(mk.thisElem.size(), -1);
		for (size_t i = 0; i < mk.thisElem.size(); i++) {
			const int e = mk.thisElem[i];
			if (e < 0 || e >= nElem) throw std::runtime_error("interface element refers to cell " + std::to_string(e) + " outside the partition");
			int cand[6], nc = 0;
			for (int j = s.element2faceStart[e]; j < s.element2faceStart[e + 1]; j++)
				if (s.element2element[j] >= nElem) cand[nc++] = s.element2element[j];
			int g = -1;
			if (nc == 1) {
				g = cand[0];
			} else if (nc > 1) {
				std::vector<int> want(mk.nodes.begin() + mk.nodeStart[i], mk.nodes.begin() + mk.nodeStart[i + 1]);
				std::sort(want.begin(), want.end());
				for (int c = 0; c < nc && g < 0; c++) {
					int f = s.nFaceInt + (cand[c] - nElem);
					std::vector<int> have(s.face2node.begin() + s.face2nodeStart[f], s.face2node.begin() + s.face2nodeStart[f + 1]);
					std::sort(have.begin(), have.end());
					if (have == want) g = cand[c];
				}
			}
			if (g < 0)
				throw std::runtime_error("interface element " + std::to_string(i) + " toward partition " + std::to_string(mk.partner) +
				                         " (cell " + std::to_string(e) + ") matches no boundary face");
			if (inMpi[g - nElem]) throw std::runtime_error("two interface elements resolve to ghost " + std::to_string(g));
			inMpi[g - nElem] = 1;
			s.interfaceGhost[im][i] = g;
			s.mpiGhost.push_back(g);
			s.mpiCell.push_back(e);
		}
	}

	// ---- ghost classification (Mesh.hpp:594-707): marker order, first matching boundary face
	{
		std::vector<FaceKey> bkeys(s.nGhost);
		for (int k = 0; k < s.nGhost; k++) {
			int f = s.nFaceInt + k;
			int a = s.face2nodeStart[f];
			bkeys[k] = {sortedKey(&s.face2node[a], s.face2nodeStart[f + 1] - a), f};
		}
		std::sort(bkeys.begin(), bkeys.end());
		auto lookup = [&](const std::array<int, 4> &k) -> int {
			FaceKey probe{k, -1};
			auto it = std::lower_bound(bkeys.begin(), bkeys.end(), probe);
			if (it != bkeys.end() && it->k == k) return (int) it->slot;
			return -1;
		};
		s.ghostKind.assign(s.nGhost, BND_OTHER);
		std::vector<uint8_t> inWall(s.nGhost, 0), inSym(s.nGhost, 0), inFar(s.nGhost, 0);
		auto classify = [&](int kind, std::vector<int> &g, std::vector<int> &c, std::vector<int> &fl,
		                    std::vector<uint8_t> &mark) {
			for (auto &mk : m.markers) {
				std::string t = lower(mk.tag);
				bool match = (kind == BND_WALL && (t == "airfoil" || t == "wall")) || (kind == BND_FARFIELD && t == "farfield") ||
				             (kind == BND_SYM && t == "sym");
				if (!match) continue;
				for (size_t i = 0; i < mk.vtk.size(); i++) {
					int a = mk.nodeStart[i];
					int f = lookup(sortedKey(&mk.nodes[a], mk.nodeStart[i + 1] - a));
					if (f < 0) continue;
					int gh = s.face2element[2 * f + 1];
					g.push_back(gh);
					c.push_back(s.face2element[2 * f]);
					fl.push_back(f);
					mark[gh - nElem] = 1;
				}
			}
		};
		classify(BND_WALL, s.wallGhost, s.wallCell, s.wallFace, inWall);
		classify(BND_FARFIELD, s.farGhost, s.farCell, s.farFace, inFar);
		classify(BND_SYM, s.symGhost, s.symCell, s.symFace, inSym);
		for (auto &mk : m.markers) {
			std::string t = lower(mk.tag);
			if (!(t == "airfoil" || t == "wall" || t == "farfield" || t == "sym"))
				throw std::runtime_error("unsupported boundary tag '" + mk.tag +
				                         "' (the solver knows farfield, wall/airfoil and sym)");
		}
		// residual dispatch precedence of EulerSolver.cpp:331-357: wall, then interface or sym,
		// else farfield
		for (int k = 0; k < s.nGhost; k++) {
			if (inWall[k]) s.ghostKind[k] = BND_WALL;
			else if (inMpi[k]) s.ghostKind[k] = BND_MPI;
			else if (inSym[k]) s.ghostKind[k] = BND_SYM;
			else if (inFar[k]) s.ghostKind[k] = BND_FARFIELD;
			else throw std::runtime_error("boundary face " + std::to_string(s.nFaceInt + k) + " is on no marker");
		}
	}

	// ---- metrics (Metrics.cpp:9-309)
	const int nNode = m.nNode();
	std::vector<V3> P(nNode);
	for (int i = 0; i < nNode; i++) P[i] = {m.xyz[3 * i] / cref, m.xyz[3 * i + 1] / cref, m.xyz[3 * i + 2] / cref};

	const int nFace = s.nFace;
	s.faceNormal.resize(3 * (size_t) nFace);
	s.faceUnitNormal.resize(3 * (size_t) nFace);
	s.faceArea.resize(nFace);
	s.faceCenter.resize(3 * (size_t) nFace);
	for (int f = 0; f < nFace; f++) {
		int a = s.face2nodeStart[f];
		int nn = s.face2nodeStart[f + 1] - a;
		const int *fn = &s.face2node[a];
		V3 c{}, N{}, U{};
		double area = 0.0;
		if (nn == 3) {
			c = dv(add(add(P[fn[0]], P[fn[1]]), P[fn[2]]), 3.0);
			V3 AB = sub(P[fn[1]], P[fn[0]]);
			V3 AC = sub(P[fn[2]], P[fn[0]]);
			area = triArea(AB, AC);
			N = mul(cross(AB, AC), 0.5);
			U = dv(N, length(N));
		} else {
			V3 p0 = P[fn[0]], p1 = P[fn[1]], p2 = P[fn[2]], p3 = P[fn[3]];
			V3 AB = sub(p1, p0), AC = sub(p2, p0), AD = sub(p3, p0);
			double w123 = triArea(AB, AC);
			double w134 = triArea(AC, AD);
			V3 c123 = dv(add(add(p0, p1), p2), 3.0);
			V3 c134 = dv(add(add(p0, p2), p3), 3.0);
			c = dv(add(mul(c123, w123), mul(c134, w134)), w123 + w134);
			area = w134 + w123;
			double dXA = p3.x - p1.x, dXB = p2.x - p0.x;
			double dYA = p3.y - p1.y, dYB = p2.y - p0.y;
			double dZA = p3.z - p1.z, dZB = p2.z - p0.z;
			N = mul(V3{dZA * dYB - dYA * dZB, dXA * dZB - dZA * dXB, dYA * dXB - dXA * dYB}, 0.5);
			U = dv(N, area);
		}
		s.faceArea[f] = area;
		s.faceCenter[3 * (size_t) f] = c.x;
		s.faceCenter[3 * (size_t) f + 1] = c.y;
		s.faceCenter[3 * (size_t) f + 2] = c.z;
		s.faceNormal[3 * (size_t) f] = N.x;
		s.faceNormal[3 * (size_t) f + 1] = N.y;
		s.faceNormal[3 * (size_t) f + 2] = N.z;
		s.faceUnitNormal[3 * (size_t) f] = U.x;
		s.faceUnitNormal[3 * (size_t) f + 1] = U.y;
		s.faceUnitNormal[3 * (size_t) f + 2] = U.z;
	}

	s.cellVolume.resize(nElem);
	s.cellCentroid.resize(3 * (size_t) nElem);
	for (int e = 0; e < nElem; e++) {
		const int *en = &m.elemNodes[m.elemStart[e]];
		int nn = m.elemStart[e + 1] - m.elemStart[e];
		V3 c{};
		double vol = 0.0;
		if (nn == 8) {
			V3 AB = sub(P[en[1]], P[en[0]]), AC = sub(P[en[2]], P[en[0]]), AD = sub(P[en[3]], P[en[0]]);
			double area = triArea(AB, AC) + triArea(AC, AD);
			double dist = length(sub(P[en[0]], P[en[4]]));
			vol = area * dist;
			double sx = 0.0, sy = 0.0, sz = 0.0;
			for (int i = 0; i < 8; i++) {
				sx += P[en[i]].x;
				sy += P[en[i]].y;
				sz += P[en[i]].z;
			}
			c = {sx / 8.0, sy / 8.0, sz / 8.0};
		} else if (nn == 6) {
			V3 p[6];
			for (int i = 0; i < 6; i++) p[i] = P[en[i]];
			double v1 = tetVolume(p[0], p[1], p[2], p[3]);
			double v2 = tetVolume(p[1], p[4], p[2], p[3]);
			double v3 = tetVolume(p[3], p[5], p[4], p[2]);
			vol = v1 + v2 + v3;
			V3 c1 = tetCentroid(p[0], p[1], p[2], p[3]);
			V3 c2 = tetCentroid(p[1], p[2], p[4], p[3]);
			V3 c3 = tetCentroid(p[3], p[4], p[5], p[2]);
			c = dv(add(add(mul(c1, v1), mul(c2, v2)), mul(c3, v3)), vol);
		} else if (nn == 5) {
			V3 p[5];
			for (int i = 0; i < 5; i++) p[i] = P[en[i]];
			double v1 = tetVolume(p[0], p[2], p[3], p[4]);
			double v2 = tetVolume(p[0], p[1], p[2], p[4]);
			vol = v1 + v2;
			V3 c1 = tetCentroid(p[0], p[2], p[3], p[4]);
			V3 c2 = tetCentroid(p[0], p[1], p[2], p[4]);
			c = dv(add(mul(c1, v1), mul(c2, v2)), vol);
		} else {
			V3 sum{0.0, 0.0, 0.0};
			for (int i = 0; i < 4; i++) sum = add(sum, P[en[i]]);
			c = dv(sum, 4.0);
			vol = tetVolume(P[en[0]], P[en[1]], P[en[2]], P[en[3]]);
		}
		s.cellVolume[e] = vol;
		s.cellCentroid[3 * (size_t) e] = c.x;
		s.cellCentroid[3 * (size_t) e + 1] = c.y;
		s.cellCentroid[3 * (size_t) e + 2] = c.z;
	}

	// reorientation (Metrics.cpp:270-309): normals point from elem0 to elem1
	for (int f = 0; f < nFace; f++) {
		int e0 = s.face2element[2 * (size_t) f];
		int e1 = s.face2element[2 * (size_t) f + 1];
		V3 c0{s.faceCenter[3 * (size_t) f], s.faceCenter[3 * (size_t) f + 1], s.faceCenter[3 * (size_t) f + 2]};
		V3 N{s.faceNormal[3 * (size_t) f], s.faceNormal[3 * (size_t) f + 1], s.faceNormal[3 * (size_t) f + 2]};
		V3 k1{s.cellCentroid[3 * (size_t) e0], s.cellCentroid[3 * (size_t) e0 + 1], s.cellCentroid[3 * (size_t) e0 + 2]};
		double cos1 = dot(sub(k1, c0), N);
		bool flip;
		if (e1 < nElem) {
			V3 k2{s.cellCentroid[3 * (size_t) e1], s.cellCentroid[3 * (size_t) e1 + 1], s.cellCentroid[3 * (size_t) e1 + 2]};
			double cos2 = dot(sub(k2, c0), N);
			flip = cos1 > 0 && cos2 < 0;
		} else {
			flip = cos1 > 0;
		}
		if (flip) {
			for (int k = 0; k < 3; k++) {
				s.faceNormal[3 * (size_t) f + k] *= -1;
				s.faceUnitNormal[3 * (size_t) f + k] *= -1;
			}
		}
	}
	return s;
}

}// namespace e3d
