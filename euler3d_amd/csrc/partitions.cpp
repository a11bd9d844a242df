// Several partitions of one mesh as one resident problem, and the partition files of a mesh.
//
// The reference runs one MPI rank per partition (src/Solver.cpp:19-73). Each rank numbers its
// partition on its own (faces, ghosts, the per-cell order of accumulation), and the faces it
// shares with another partition are boundary faces whose ghost cell holds a copy of a cell of the
// other partition:
//   * after every update a rank copies each interface ghost's own adjacent cell into the ghost
//     (FlowField.cpp:129-138, 159-169);
//   * updateBC then swaps those ghost values pairwise between ranks (MPIHandler.cpp:84-267): the
//     values leave in the order of the requester list (sorted by the other partition's element,
//     MPIHandler.cpp:16-20) and land in the order of the sender list (sorted by this partition's
//     element, :21-25), both with std::sort, and p and H are recomputed from (rho, u, v, w, E)
//     with the expressions the update used, so they keep their bits;
//   * the flux of an interface face is Roe(own cell, ghost) and goes to the own cell only
//     (EulerSolver.cpp:343-346).
// On one device every partition is resident, so an interface ghost's state is read straight
// from the cell the exchange would deliver: mergePartitions points the face's elem1 at that cell.
// The pairing follows the two std::sort orders exactly (same keys, same input order, same
// algorithm), including when a cell has several faces toward the same partition and the two
// sorts order them differently.
//
// splitMesh writes what E3D_PRE writes for an element-to-partition assignment
// (Partition.cpp:123-460): elements in ascending global id, nodes numbered by first use, the
// physical markers each partition touches (kept in an unordered_map, whose iteration order the
// file follows), and the interface blocks. The reference computes the assignment with METIS,
// which is not part of its sources; partitionRcb is this library's own (recursive coordinate
// bisection), and any other assignment can be passed in.
#include "e3d_host.hpp"

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace e3d {

namespace {

struct IfaceEntry {
	int self, other, ghost; // this-partition element, other-partition element, this-partition ghost
};

}// namespace

SolverMesh mergePartitions(const std::vector<MeshFile> &files, const std::vector<SolverMesh> &parts,
                           std::vector<PartRange> &ranges, MeshFile &merged) {
	const int N = (int) parts.size();
	if (N == 0 || files.size() != parts.size()) throw std::invalid_argument("no partitions");
	ranges.assign(N, PartRange{});
	SolverMesh c;
	int nWall = 0;
	for (int r = 0; r < N; r++) {
		PartRange &R = ranges[r];
		R.elem0 = c.nElem, R.nElem = parts[r].nElem;
		R.node0 = r ? ranges[r - 1].node0 + ranges[r - 1].nNode : 0, R.nNode = files[r].nNode();
		R.faceInt0 = c.nFaceInt, R.nFaceInt = parts[r].nFaceInt;
		R.ghost0 = c.nGhost, R.nGhost = parts[r].nGhost;
		R.wall0 = nWall, R.nWall = (int) parts[r].wallGhost.size();
		c.nElem += R.nElem, c.nFaceInt += R.nFaceInt, c.nGhost += R.nGhost, nWall += R.nWall;
	}
	c.nFace = c.nFaceInt + c.nGhost;
	const int nE = c.nElem;
	auto cellG = [&](int r, int id) {
		const PartRange &R = ranges[r];
		return id < R.nElem ? R.elem0 + id : nE + R.ghost0 + (id - R.nElem);
	};
	auto faceG = [&](int r, int f) {
		const PartRange &R = ranges[r];
		return f < R.nFaceInt ? R.faceInt0 + f : c.nFaceInt + R.ghost0 + (f - R.nFaceInt);
	};

	// faces
	const size_t nF = c.nFace;
	c.face2element.resize(2 * nF);
	c.faceNormal.resize(3 * nF), c.faceUnitNormal.resize(3 * nF), c.faceArea.resize(nF), c.faceCenter.resize(3 * nF);
	std::vector<int> nodeCount(nF);
	for (int r = 0; r < N; r++) {
		const SolverMesh &p = parts[r];
		for (int f = 0; f < p.nFace; f++) {
			const size_t g = faceG(r, f);
			c.face2element[2 * g] = cellG(r, p.face2element[2 * (size_t) f]);
			c.face2element[2 * g + 1] = cellG(r, p.face2element[2 * (size_t) f + 1]);
			for (int k = 0; k < 3; k++) {
				c.faceNormal[3 * g + k] = p.faceNormal[3 * (size_t) f + k];
				c.faceUnitNormal[3 * g + k] = p.faceUnitNormal[3 * (size_t) f + k];
				c.faceCenter[3 * g + k] = p.faceCenter[3 * (size_t) f + k];
			}
			c.faceArea[g] = p.faceArea[f];
			nodeCount[g] = p.face2nodeStart[f + 1] - p.face2nodeStart[f];
		}
	}
	c.face2nodeStart.assign(nF + 1, 0);
	for (size_t g = 0; g < nF; g++) c.face2nodeStart[g + 1] = c.face2nodeStart[g] + nodeCount[g];
	c.face2node.resize(c.face2nodeStart[nF]);
	for (int r = 0; r < N; r++) {
		const SolverMesh &p = parts[r];
		for (int f = 0; f < p.nFace; f++) {
			int d = c.face2nodeStart[faceG(r, f)];
			for (int j = p.face2nodeStart[f]; j < p.face2nodeStart[f + 1]; j++) c.face2node[d++] = ranges[r].node0 + p.face2node[j];
		}
	}

	// cells
	c.element2faceStart.assign(nE + 1, 0);
	c.cellVolume.reserve(nE), c.cellCentroid.reserve(3 * (size_t) nE);
	c.partStart.assign(N + 1, 0);
	for (int r = 0; r < N; r++) {
		const SolverMesh &p = parts[r];
		for (int e = 0; e < p.nElem; e++) {
			const int g = ranges[r].elem0 + e;
			c.element2faceStart[g + 1] = c.element2faceStart[g] + (p.element2faceStart[e + 1] - p.element2faceStart[e]);
			for (int j = p.element2faceStart[e]; j < p.element2faceStart[e + 1]; j++) {
				c.element2face.push_back(faceG(r, p.element2face[j]));
				c.element2element.push_back(cellG(r, p.element2element[j]));
			}
			c.cellVolume.push_back(p.cellVolume[e]);
			for (int k = 0; k < 3; k++) c.cellCentroid.push_back(p.cellCentroid[3 * (size_t) e + k]);
		}
		c.partStart[r + 1] = ranges[r].elem0 + p.nElem;
	}

	// ghost lists, partition by partition in each partition's order
	for (int r = 0; r < N; r++) {
		const SolverMesh &p = parts[r];
		auto app = [&](const std::vector<int> &gh, const std::vector<int> &ce, const std::vector<int> &fa, std::vector<int> &G,
		               std::vector<int> &C, std::vector<int> &F) {
			for (size_t i = 0; i < gh.size(); i++) {
				G.push_back(cellG(r, gh[i]));
				C.push_back(cellG(r, ce[i]));
				F.push_back(faceG(r, fa[i]));
			}
		};
		app(p.wallGhost, p.wallCell, p.wallFace, c.wallGhost, c.wallCell, c.wallFace);
		app(p.farGhost, p.farCell, p.farFace, c.farGhost, c.farCell, c.farFace);
		app(p.symGhost, p.symCell, p.symFace, c.symGhost, c.symCell, c.symFace);
		c.ghostKind.insert(c.ghostKind.end(), p.ghostKind.begin(), p.ghostKind.end());
		for (size_t i = 0; i < p.mpiGhost.size(); i++) {
			c.mpiGhost.push_back(cellG(r, p.mpiGhost[i]));
			c.mpiCell.push_back(cellG(r, p.mpiCell[i]));
		}
	}

	// the ghost exchange (MPIHandler::sortInterface + updateFlowField)
	std::vector<std::vector<std::vector<IfaceEntry>>> req(N), snd(N);
	for (int r = 0; r < N; r++) {
		const MeshFile &m = files[r];
		for (size_t im = 0; im < m.interfaces.size(); im++) {
			const InterfaceMarker &mk = m.interfaces[im];
			std::vector<IfaceEntry> v(mk.thisElem.size());
			for (size_t i = 0; i < v.size(); i++) v[i] = {mk.thisElem[i], mk.otherElem[i], parts[r].interfaceGhost[im][i]};
			std::vector<IfaceEntry> a = v, b = v;
			std::sort(a.begin(), a.end(), [](const IfaceEntry &x, const IfaceEntry &y) { return x.other < y.other; });
			std::sort(b.begin(), b.end(), [](const IfaceEntry &x, const IfaceEntry &y) { return x.self < y.self; });
			req[r].push_back(std::move(a));
			snd[r].push_back(std::move(b));
		}
	}
	auto markerToward = [&](int r, int s) -> int {
		const auto &ifs = files[r].interfaces;
		for (size_t im = 0; im < ifs.size(); im++)
			if (ifs[im].partner == s) return (int) im;
		return -1;
	};
	for (int r = 0; r < N; r++) {
		const auto &ifs = files[r].interfaces;
		for (size_t im = 0; im < ifs.size(); im++) {
			const int s = ifs[im].partner;
			if (s < 0 || s >= N || s == r) throw std::runtime_error("interface toward partition " + std::to_string(s) + " of " + std::to_string(N));
			if (markerToward(r, s) != (int) im) continue; // only the first block toward s takes part
			const int jm = markerToward(s, r);
			if (jm < 0) throw std::runtime_error("partition " + std::to_string(s) + " has no interface toward " + std::to_string(r));
			const auto &mine = snd[r][im];
			const auto &theirs = req[s][jm];
			if (mine.size() != theirs.size())
				throw std::runtime_error("partitions " + std::to_string(r) + " and " + std::to_string(s) + " disagree on their interface");
			if (c.haloStart.empty()) c.haloStart.push_back(0);
			c.haloFrom.push_back(s), c.haloTo.push_back(r);
			for (size_t i = 0; i < mine.size(); i++) {
				const int g = mine[i].ghost;               // local ghost id in r
				const int f = parts[r].nFaceInt + (g - parts[r].nElem);
				c.face2element[2 * (size_t) faceG(r, f) + 1] = ranges[s].elem0 + theirs[i].self;
				c.haloSrc.push_back(ranges[s].elem0 + theirs[i].self);
				c.haloDst.push_back(cellG(r, g));
			}
			c.haloStart.push_back((int) c.haloSrc.size());
		}
	}
	// a ghost that no exchange reaches keeps the copy of its own cell: one message per partition,
	// from the partition to itself
	std::vector<std::vector<int>> selfSrc(N), selfDst(N);
	for (size_t i = 0; i < c.mpiGhost.size(); i++) {
		const int f = c.nFaceInt + (c.mpiGhost[i] - nE);
		if (c.face2element[2 * (size_t) f + 1] < nE) continue;
		c.face2element[2 * (size_t) f + 1] = c.mpiCell[i];
		const int r = (int) (std::upper_bound(c.partStart.begin(), c.partStart.end(), c.mpiCell[i]) - c.partStart.begin()) - 1;
		selfSrc[r].push_back(c.mpiCell[i]);
		selfDst[r].push_back(c.mpiGhost[i]);
	}
	for (int r = 0; r < N; r++) {
		if (selfSrc[r].empty()) continue;
		if (c.haloStart.empty()) c.haloStart.push_back(0);
		c.haloFrom.push_back(r), c.haloTo.push_back(r);
		c.haloSrc.insert(c.haloSrc.end(), selfSrc[r].begin(), selfSrc[r].end());
		c.haloDst.insert(c.haloDst.end(), selfDst[r].begin(), selfDst[r].end());
		c.haloStart.push_back((int) c.haloSrc.size());
	}

	// the merged file (node ids offset per partition)
	merged = MeshFile{};
	for (int r = 0; r < N; r++) {
		const MeshFile &m = files[r];
		merged.xyz.insert(merged.xyz.end(), m.xyz.begin(), m.xyz.end());
		for (int e = 0; e < m.nElem(); e++) {
			merged.elemVtk.push_back(m.elemVtk[e]);
			for (int j = m.elemStart[e]; j < m.elemStart[e + 1]; j++) merged.elemNodes.push_back(ranges[r].node0 + m.elemNodes[j]);
			if (merged.elemStart.empty()) merged.elemStart.push_back(0);
			merged.elemStart.push_back((int) merged.elemNodes.size());
		}
		merged.nElemDeclared += m.nElemDeclared;
	}
	if (merged.elemStart.empty()) merged.elemStart.push_back(0);
	return c;
}

std::vector<int> partitionRcb(const SolverMesh &sm, int nPart) {
	const int n = sm.nElem;
	if (nPart < 1 || nPart > n) throw std::invalid_argument("partition count must be in [1, number of cells]");
	std::vector<int> part(n, 0), ids(n);
	std::iota(ids.begin(), ids.end(), 0);
	struct Job {
		int lo, hi, p0, np;
	};
	std::vector<Job> stack{{0, n, 0, nPart}};
	while (!stack.empty()) {
		Job j = stack.back();
		stack.pop_back();
		if (j.np == 1) {
			for (int i = j.lo; i < j.hi; i++) part[ids[i]] = j.p0;
			continue;
		}
		double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
		for (int i = j.lo; i < j.hi; i++)
			for (int k = 0; k < 3; k++) {
				double x = sm.cellCentroid[3 * (size_t) ids[i] + k];
				lo[k] = std::min(lo[k], x), hi[k] = std::max(hi[k], x);
			}
		int ax = 0;
		for (int k = 1; k < 3; k++)
			if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
		const int nl = j.np / 2;
		const int cut = j.lo + (int) ((long long) (j.hi - j.lo) * nl / j.np);
		std::nth_element(ids.begin() + j.lo, ids.begin() + cut, ids.begin() + j.hi, [&](int a, int b) {
			double xa = sm.cellCentroid[3 * (size_t) a + ax], xb = sm.cellCentroid[3 * (size_t) b + ax];
			return xa < xb || (xa == xb && a < b);
		});
		stack.push_back({cut, j.hi, j.p0 + nl, j.np - nl});
		stack.push_back({j.lo, cut, j.p0, nl});
	}
	return part;
}

std::vector<MeshFile> splitMesh(const MeshFile &g, const SolverMesh &sm, const std::vector<int> &elem2part, int nPart) {
	const int nE = g.nElem(), nN = g.nNode();
	if ((int) elem2part.size() != nE) throw std::invalid_argument("one partition id per cell is required");
	for (int p : elem2part)
		if (p < 0 || p >= nPart) throw std::invalid_argument("partition id out of range");
	if (sm.nElem != nE) throw std::invalid_argument("the mesh is not preprocessed");

	// Partition::SolvePart2Elem (Partition.cpp:123-151): elements of each partition, ascending
	std::vector<std::vector<int>> part2elem(nPart);
	for (int e = 0; e < nE; e++) part2elem[elem2part[e]].push_back(e);
	// Partition::SolveElem2Node (:153-203): local nodes in order of first use
	std::vector<int> elemLocal(nE);
	std::vector<std::vector<std::pair<int, int>>> nodeMap(nPart); // (global, local), sorted by global
	std::vector<MeshFile> out(nPart);
	std::vector<int> seen(nN, -1);
	for (int p = 0; p < nPart; p++) {
		MeshFile &m = out[p];
		std::vector<int> l2g;
		m.elemStart.push_back(0);
		for (size_t i = 0; i < part2elem[p].size(); i++) {
			const int e = part2elem[p][i];
			elemLocal[e] = (int) i;
			m.elemVtk.push_back(g.elemVtk[e]);
			for (int j = g.elemStart[e]; j < g.elemStart[e + 1]; j++) {
				const int gn = g.elemNodes[j];
				if (seen[gn] != p) {
					seen[gn] = p;
					nodeMap[p].push_back({gn, (int) l2g.size()});
					l2g.push_back(gn);
				}
			}
		}
		std::sort(nodeMap[p].begin(), nodeMap[p].end());
		auto local = [&](int gn) { return std::lower_bound(nodeMap[p].begin(), nodeMap[p].end(), std::make_pair(gn, -1))->second; };
		for (int e : part2elem[p]) {
			for (int j = g.elemStart[e]; j < g.elemStart[e + 1]; j++) m.elemNodes.push_back(local(g.elemNodes[j]));
			m.elemStart.push_back((int) m.elemNodes.size());
		}
		m.xyz.resize(3 * l2g.size());
		for (size_t i = 0; i < l2g.size(); i++)
			for (int k = 0; k < 3; k++) m.xyz[3 * i + k] = g.xyz[3 * (size_t) l2g[i] + k];
		m.nElemDeclared = (long) part2elem[p].size();
	}
	auto localNode = [&](int p, int gn) {
		auto it = std::lower_bound(nodeMap[p].begin(), nodeMap[p].end(), std::make_pair(gn, -1));
		if (it == nodeMap[p].end() || it->first != gn) throw std::logic_error("node not in partition");
		return it->second;
	};

	// Partition::SolveBorder (:220-286): per partition pair (i < j), in i's element order and
	// each element's neighbour order, the first face of the element shared with the neighbour
	std::vector<std::vector<InterfaceMarker>> iface(nPart, std::vector<InterfaceMarker>(nPart));
	for (int p = 0; p < nPart; p++) {
		for (int e : part2elem[p]) {
			for (int j = sm.element2faceStart[e]; j < sm.element2faceStart[e + 1]; j++) {
				const int nb = sm.element2element[j];
				if (nb >= nE) continue;
				const int q = elem2part[nb];
				if (q <= p) continue;
				int face = -1;
				for (int k = sm.element2faceStart[e]; k < sm.element2faceStart[e + 1] && face < 0; k++) {
					const int f = sm.element2face[k];
					if (sm.face2element[2 * (size_t) f] == nb || sm.face2element[2 * (size_t) f + 1] == nb) face = f;
				}
				const int a = sm.face2nodeStart[face], b = sm.face2nodeStart[face + 1];
				const int vtk = (b - a) == 4 ? VTK_QUAD : (b - a) == 3 ? VTK_TRI : 7;
				InterfaceMarker &mp = iface[p][q], &mq = iface[q][p];
				mp.partner = q, mq.partner = p;
				mp.thisElem.push_back(elemLocal[e]), mp.otherElem.push_back(elemLocal[nb]), mp.vtk.push_back(vtk);
				mq.thisElem.push_back(elemLocal[nb]), mq.otherElem.push_back(elemLocal[e]), mq.vtk.push_back(vtk);
				for (int k = a; k < b; k++) {
					mp.nodes.push_back(localNode(p, sm.face2node[k]));
					mq.nodes.push_back(localNode(q, sm.face2node[k]));
				}
				mp.nodeStart.push_back((int) mp.nodes.size());
				mq.nodeStart.push_back((int) mq.nodes.size());
			}
		}
	}
	for (int p = 0; p < nPart; p++) {
		for (int q = 0; q < nPart; q++)
			if (!iface[p][q].thisElem.empty()) out[p].interfaces.push_back(std::move(iface[p][q]));
		out[p].nInternalMarker = (int) out[p].interfaces.size();
	}

	// Partition::PhysicalPartitionSolve (:288-336): each marker element goes to the partition of
	// the cell that owns its face; a partition's markers are kept in an unordered_map keyed by tag
	std::vector<std::array<int, 4>> bkey(sm.nGhost);
	std::vector<std::pair<std::array<int, 4>, int>> keys;
	keys.reserve(sm.nGhost);
	for (int k = 0; k < sm.nGhost; k++) {
		const int f = sm.nFaceInt + k;
		std::array<int, 4> key{-1, -1, -1, -1};
		const int a = sm.face2nodeStart[f], n = sm.face2nodeStart[f + 1] - a;
		for (int i = 0; i < n; i++) key[i] = sm.face2node[a + i];
		std::sort(key.begin(), key.begin() + n);
		keys.push_back({key, f});
	}
	std::sort(keys.begin(), keys.end());
	std::vector<std::unordered_map<std::string, size_t>> tagIndex(nPart);
	std::vector<std::vector<Marker>> markers(nPart);
	for (const Marker &mk : g.markers) {
		for (size_t i = 0; i < mk.vtk.size(); i++) {
			const int a = mk.nodeStart[i], n = mk.nodeStart[i + 1] - a;
			std::array<int, 4> key{-1, -1, -1, -1};
			for (int t = 0; t < n; t++) key[t] = mk.nodes[a + t];
			std::sort(key.begin(), key.begin() + n);
			auto it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(key, -1));
			if (it == keys.end() || it->first != key) throw std::runtime_error("marker element of '" + mk.tag + "' is on no boundary face");
			const int owner = sm.face2element[2 * (size_t) it->second];
			const int p = elem2part[owner];
			auto ins = tagIndex[p].emplace(mk.tag, markers[p].size());
			if (ins.second) {
				Marker nm;
				nm.tag = mk.tag;
				nm.nodeStart.push_back(0);
				markers[p].push_back(std::move(nm));
			}
			Marker &dst = markers[p][ins.first->second];
			dst.vtk.push_back(mk.vtk[i]);
			for (int t = 0; t < n; t++) dst.nodes.push_back(localNode(p, mk.nodes[a + t]));
			dst.nodeStart.push_back((int) dst.nodes.size());
		}
	}
	for (int p = 0; p < nPart; p++)
		for (auto &kv : tagIndex[p]) out[p].markers.push_back(std::move(markers[p][kv.second]));
	return out;
}

}// namespace e3d
