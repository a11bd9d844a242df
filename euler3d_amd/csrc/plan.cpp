// Device data plan: locality order of the cells, fused-stage blocks and device-numbered arrays.
//
// The device keeps cells in a Morton (Z-curve) order of their centroids, so kCellsPerBlock
// consecutive cells form a compact cluster whose faces are mostly shared inside the cluster.
// The reference numbering is kept where it determines the FP64 bits (face orientation and the
// per-cell order of accumulation, ascending reference face id, EulerSolver.cpp:306-358), so the
// order changes memory locality only.
#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <functional>
#include <thread>
#include <cmath>
#include <numeric>
#include <type_traits>
#include <stdexcept>

namespace e3d {

namespace {

uint64_t spread3(uint64_t x) {
	x &= 0x1fffff;
	x = (x | x << 32) & 0x1f00000000ffffULL;
	x = (x | x << 16) & 0x1f0000ff0000ffULL;
	x = (x | x << 8) & 0x100f00f00f00f00fULL;
	x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
	x = (x | x << 2) & 0x1249249249249249ULL;
	return x;
}

std::vector<int> mortonOrder(const SolverMesh &sm) {
	const int n = sm.nElem;
	double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
	for (int e = 0; e < n; e++)
		for (int k = 0; k < 3; k++) {
			double x = sm.cellCentroid[3 * (size_t) e + k];
			lo[k] = std::min(lo[k], x);
			hi[k] = std::max(hi[k], x);
		}
	double ext = 0;
	for (int k = 0; k < 3; k++) ext = std::max(ext, hi[k] - lo[k]);
	double scale = ext > 0 ? 2097151.0 / ext : 0.0;
	std::vector<uint64_t> key(n);
	for (int e = 0; e < n; e++) {
		uint64_t q[3];
		for (int k = 0; k < 3; k++) {
			double t = (sm.cellCentroid[3 * (size_t) e + k] - lo[k]) * scale;
			q[k] = (uint64_t) std::min(2097151.0, std::max(0.0, t));
		}
		key[e] = spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2;
	}
	// the cells of a composite mesh stay grouped by partition (Morton order inside each), so that
	// per-partition reductions run over contiguous ranges
	std::vector<int> part(n, 0);
	for (size_t r = 0; r + 1 < sm.partStart.size(); r++)
		for (int e = sm.partStart[r]; e < sm.partStart[r + 1]; e++) part[e] = (int) r;
	std::vector<int> order(n);
	std::iota(order.begin(), order.end(), 0);
	std::sort(order.begin(), order.end(), [&](int a, int b) {
		if (part[a] != part[b]) return part[a] < part[b];
		return key[a] < key[b] || (key[a] == key[b] && a < b);
	});
	return order;
}

// Blocks by recursive coordinate bisection of the cell centroids: a range of cells is split so
// that the lower part holds half of the range's blocks (a whole number of cellsPerBlock cells);
// leaves are single blocks, in Morton order inside. Each split goes across the longest extent,
// unless the median plane of another axis cuts fewer than 1/1.5 as many faces (counted exactly for
// all three axes in one pass): stretched cells (a wall-resolved O-grid) make the longest extent a
// poor choice there (wing: 454 -> 291 halo cells per block), while on near-isotropic cells it
// keeps the halos smallest. Box-shaped blocks cut fewer
// faces than Morton runs (bump channel: 2.38 instead of 2.45 face slots per cell), so fewer fluxes
// are computed twice and fewer halo cells are read. The cells of a composite mesh stay grouped by
// partition. Subtrees run on up to 16 threads; the result does not depend on their number.
std::vector<int> bisectOrder(const SolverMesh &sm, int cellsPerBlock) {
	// the bisection runs on Morton indices, over Morton-ordered copies of the centroids and of the
	// interior neighbours, so that its memory accesses stay local
	const std::vector<int> morton = mortonOrder(sm);
	const int n = sm.nElem;
	std::vector<int> rank(n);
	for (int i = 0; i < n; i++) rank[morton[i]] = i;
	std::vector<double> xyz(3 * (size_t) n);
	std::vector<int> nstart(n + 1, 0), nbr;
	nbr.reserve(sm.element2face.size());
	for (int i = 0; i < n; i++) {
		const int c = morton[i];
		for (int k = 0; k < 3; k++) xyz[3 * (size_t) i + k] = sm.cellCentroid[3 * (size_t) c + k];
		for (int q = sm.element2faceStart[c]; q < sm.element2faceStart[c + 1]; q++) {
			const int f = sm.element2face[q];
			const int a = sm.face2element[2 * (size_t) f], o = a == c ? sm.face2element[2 * (size_t) f + 1] : a;
			if (o < n) nbr.push_back(rank[o]);
		}
		nstart[i + 1] = (int) nbr.size();
	}
	std::vector<int> ids(n);
	std::iota(ids.begin(), ids.end(), 0);
	// Morton order keeps the partitions of a composite mesh in their ranges
	std::vector<int> bounds = sm.partStart.size() > 1 ? sm.partStart : std::vector<int>{0, n};
	// job id of each cell; a cell is read by the jobs of its neighbours' threads, so relaxed atomics
	std::vector<std::atomic<int>> stamp(n);
	for (auto &x : stamp) x.store(0, std::memory_order_relaxed);
	std::atomic<int> jobs{0};
	// (coordinate, index) order along an axis
	auto before = [&](int ax, int x, int y) {
		const double a = xyz[3 * (size_t) x + ax], b = xyz[3 * (size_t) y + ax];
		return a < b || (a == b && x < y);
	};
	// f(i0, i1) over [lo, hi), on several threads when the range is large
	auto chunks = [](int lo, int hi, const std::function<void(int, int)> &f) {
		const int m = hi - lo, T = m > (1 << 20) ? 16 : 1;
		if (T == 1) return f(lo, hi);
		std::vector<std::thread> th;
		for (int t = 0; t < T; t++) th.emplace_back(f, lo + (int) ((long) m * t / T), lo + (int) ((long) m * (t + 1) / T));
		for (auto &x : th) x.join();
	};
	std::function<void(int, int, int)> run = [&](int lo, int hi, int depth) {
		const int nb = (hi - lo + cellsPerBlock - 1) / cellsPerBlock;
		if (nb <= 1) {
			std::sort(ids.begin() + lo, ids.begin() + hi);
			return;
		}
		const int cut = lo + (nb / 2) * cellsPerBlock, half = cut - lo;
		const int id = ++jobs;
		chunks(lo, hi, [&](int a, int b) {
			for (int i = a; i < b; i++) stamp[ids[i]].store(id, std::memory_order_relaxed);
		});
		// the median cell of each axis, selected on (coordinate, index) pairs held contiguously
		struct KV {
			double x;
			int id;
		};
		auto less = [](const KV &a, const KV &b) { return a.x < b.x || (a.x == b.x && a.id < b.id); };
		std::vector<KV> kv(hi - lo);
		auto select = [&](int ax) {
			for (int i = lo; i < hi; i++) kv[i - lo] = {xyz[3 * (size_t) ids[i] + ax], ids[i]};
			std::nth_element(kv.begin(), kv.begin() + half, kv.end(), less);
		};
		int pivot[3];
		for (int ax = 0; ax < 3; ax++) {
			select(ax);
			pivot[ax] = kv[half].id;
		}
		// faces cut by each axis's split (each cut face counted from both sides)
		std::atomic<long> cuts[3];
		for (auto &x : cuts) x.store(0);
		chunks(lo, hi, [&](int a, int b) {
			long cc[3] = {0, 0, 0};
			for (int i = a; i < b; i++) {
				const int c = ids[i];
				bool sc[3];
				for (int ax = 0; ax < 3; ax++) sc[ax] = before(ax, c, pivot[ax]);
				for (int k = nstart[c]; k < nstart[c + 1]; k++) {
					const int o = nbr[k];
					if (stamp[o].load(std::memory_order_relaxed) != id) continue;
					for (int ax = 0; ax < 3; ax++) cc[ax] += before(ax, o, pivot[ax]) != sc[ax];
				}
			}
			for (int ax = 0; ax < 3; ax++) cuts[ax] += cc[ax];
		});
		// the longest extent unless another axis cuts far fewer faces
		double ext[3];
		for (int ax = 0; ax < 3; ax++) {
			double a = 1e300, b = -1e300;
			for (int i = lo; i < hi; i++) a = std::min(a, xyz[3 * (size_t) ids[i] + ax]), b = std::max(b, xyz[3 * (size_t) ids[i] + ax]);
			ext[ax] = b - a;
		}
		int longest = 0, best = 0;
		for (int ax = 1; ax < 3; ax++) {
			if (ext[ax] > ext[longest]) longest = ax;
			if (cuts[ax].load() < cuts[best].load()) best = ax;
		}
		if ((double) cuts[longest].load() <= 1.5 * (double) cuts[best].load()) best = longest;
		select(best);
		for (int i = lo; i < hi; i++) ids[i] = kv[i - lo].id;
		std::vector<KV>().swap(kv);
		if (depth < 4 && hi - lo > (1 << 16)) {
			std::thread left([&, lo, cut, depth] { run(lo, cut, depth + 1); });
			run(cut, hi, depth + 1);
			left.join();
		} else {
			run(lo, cut, depth + 1);
			run(cut, hi, depth + 1);
		}
	};
	std::vector<std::thread> parts;
	for (size_t r = 0; r + 1 < bounds.size(); r++) parts.emplace_back([&, r] { run(bounds[r], bounds[r + 1], 0); });
	for (auto &t : parts) t.join();
	for (auto &x : ids) x = morton[x];
	return ids;
}

int8_t slotKind(const SolverMesh &sm, int f, bool exchange) {
	if (f < sm.nFaceInt) return 0;
	switch (sm.ghostKind[f - sm.nFaceInt]) {
		case BND_WALL: return 1;
		case BND_SYM: return 2;
		// the other partition's cell is elem1 (a one-sided interior face), or with the exchange
		// the ghost it fills
		case BND_MPI: return exchange ? 4 : 0;
		default: return 3;
	}
}

}// namespace

HostPlan buildPlan(const SolverMesh &sm, int order, int cellsPerBlock, bool exchange) {
	HostPlan P;
	const int nE = sm.nElem, nF = sm.nFace;
	P.nElem = nE;
	P.nFace = nF;
	P.nFaceInt = sm.nFaceInt;
	P.nGhost = sm.nGhost;
	if (order == ORDER_MORTON) {
		P.new2old = mortonOrder(sm);
	} else if (order == ORDER_BISECT) {
		P.new2old = bisectOrder(sm, cellsPerBlock);
	} else {
		P.new2old.resize(nE);
		std::iota(P.new2old.begin(), P.new2old.end(), 0);
	}
	P.old2new.assign(nE, -1);
	for (int i = 0; i < nE; i++) P.old2new[P.new2old[i]] = i;
	auto dev = [&](int cell) { return cell < nE ? P.old2new[cell] : cell; };

	// faces in reference order
	P.fe0.resize(nF), P.fe1.resize(nF);
	P.fnx.resize(nF), P.fny.resize(nF), P.fnz.resize(nF), P.fS.resize(nF);
	for (int f = 0; f < nF; f++) {
		P.fe0[f] = dev(sm.face2element[2 * (size_t) f]);
		P.fe1[f] = dev(sm.face2element[2 * (size_t) f + 1]);
		if (exchange && f >= sm.nFaceInt && sm.ghostKind[f - sm.nFaceInt] == BND_MPI) P.fe1[f] = nE + (f - sm.nFaceInt);
		P.fnx[f] = sm.faceUnitNormal[3 * (size_t) f];
		P.fny[f] = sm.faceUnitNormal[3 * (size_t) f + 1];
		P.fnz[f] = sm.faceUnitNormal[3 * (size_t) f + 2];
		P.fS[f] = sm.faceArea[f];
	}
	P.gkind.resize(sm.nGhost);
	for (int k = 0; k < sm.nGhost; k++)
		P.gkind[k] = (int8_t) (sm.ghostKind[k] == BND_WALL ? 1 : sm.ghostKind[k] == BND_SYM ? 2 : sm.ghostKind[k] == BND_MPI ? 3 : 0);

	// cells in device order: static time-step sums (local-face order, TimeIntegration.cpp:20-26)
	// and the face list in ascending reference face id
	P.vol.resize(nE), P.sx.resize(nE), P.sy.resize(nE), P.sz.resize(nE);
	P.fstart.assign(nE + 1, 0);
	int maxFaces = 0;
	for (int c = 0; c < nE; c++) {
		int old = P.new2old[c];
		int b = sm.element2faceStart[old], en = sm.element2faceStart[old + 1];
		maxFaces = std::max(maxFaces, en - b);
		P.fstart[c + 1] = P.fstart[c] + (en - b);
	}
	P.flist.resize(P.fstart[nE]);
	for (int c = 0; c < nE; c++) {
		int old = P.new2old[c];
		int b = sm.element2faceStart[old], en = sm.element2faceStart[old + 1];
		double ax = 0, ay = 0, az = 0;
		for (int j = b; j < en; j++) {
			int f = sm.element2face[j];
			ax += std::fabs(sm.faceNormal[3 * (size_t) f]);
			ay += std::fabs(sm.faceNormal[3 * (size_t) f + 1]);
			az += std::fabs(sm.faceNormal[3 * (size_t) f + 2]);
		}
		P.vol[c] = sm.cellVolume[old];
		P.sx[c] = ax, P.sy[c] = ay, P.sz[c] = az;
		int fl[8];
		for (int j = b; j < en; j++) fl[j - b] = sm.element2face[j];
		std::sort(fl, fl + (en - b));
		for (int j = 0; j < en - b; j++) {
			int f = fl[j];
			// the second cell of an interior face subtracts; boundary (and partition interface)
			// faces only ever add to their own cell
			bool second = f < sm.nFaceInt && sm.face2element[2 * (size_t) f + 1] == old;
			P.flist[P.fstart[c] + j] = second ? (int) ((unsigned) f | 0x80000000u) : f;
		}
	}

	// fused-stage blocks: every face touching the block gets one slot in it
	P.cellsPerBlock = cellsPerBlock;
	P.nBlocks = (nE + cellsPerBlock - 1) / cellsPerBlock;
	P.csrStride = maxFaces;
	P.csr.assign((size_t) nE * maxFaces, 0xFFFF);
	P.blockSlot.assign(P.nBlocks + 1, 0);
	std::vector<int> slotOf(nF, -1), slotBlock(nF, -1);
	size_t reserve = (size_t) nF + nF / 2;
	P.se0.reserve(reserve), P.se1.reserve(reserve), P.skind.reserve(reserve);
	P.snx.reserve(reserve), P.sny.reserve(reserve), P.snz.reserve(reserve), P.sS.reserve(reserve);
	for (int blk = 0; blk < P.nBlocks; blk++) {
		int c0 = blk * cellsPerBlock, c1 = std::min(nE, c0 + cellsPerBlock);
		int ns = 0;
		for (int c = c0; c < c1; c++) {
			for (int j = P.fstart[c]; j < P.fstart[c + 1]; j++) {
				unsigned code = (unsigned) P.flist[j];
				int f = (int) (code & 0x7fffffffu);
				int slot;
				if (slotBlock[f] == blk) {
					slot = slotOf[f];
				} else {
					slot = ns++;
					slotBlock[f] = blk;
					slotOf[f] = slot;
					P.se0.push_back(P.fe0[f]);
					P.se1.push_back(P.fe1[f]);
					P.skind.push_back(slotKind(sm, f, exchange));
					P.snx.push_back(P.fnx[f]);
					P.sny.push_back(P.fny[f]);
					P.snz.push_back(P.fnz[f]);
					P.sS.push_back(P.fS[f]);
				}
				if (slot > 0x7FFF) throw std::runtime_error("fused-stage block has more than 32768 face slots");
				P.csr[(size_t) c * maxFaces + (j - P.fstart[c])] = (uint16_t) (slot | ((code >> 31) << 15));
			}
		}
		P.blockSlot[blk + 1] = P.blockSlot[blk] + ns;
		P.maxSlots = std::max(P.maxSlots, ns);
	}

	// Within each block, interior slots first, then the boundary slots grouped by kind (wall,
	// symmetry, farfield), each group in its order of first appearance. The flux loop hands 64
	// consecutive slots to a wave, so a boundary face's ghost and flux code then runs in the few
	// waves that hold boundary slots instead of diverging the waves of every block at the wall.
	// The gather order is the CSR order (ascending reference face id), which the renumbering
	// keeps, so results do not change.
	{
		std::vector<int> newIdx;
		for (int blk = 0; blk < P.nBlocks; blk++) {
			const int s0 = P.blockSlot[blk], s1 = P.blockSlot[blk + 1], ns = s1 - s0;
			int count[6] = {0, 0, 0, 0, 0, 0};
			for (int s = s0; s < s1; s++) count[P.skind[s] + 1]++;
			for (int k = 1; k < 6; k++) count[k] += count[k - 1];
			newIdx.assign(ns, 0);
			for (int j = 0; j < ns; j++) newIdx[j] = count[P.skind[s0 + j]]++;
			auto permute = [&](auto &v) {
				using T = typename std::decay_t<decltype(v)>::value_type;
				std::vector<T> tmp(v.begin() + s0, v.begin() + s1);
				for (int j = 0; j < ns; j++) v[s0 + newIdx[j]] = tmp[j];
			};
			permute(P.se0), permute(P.se1), permute(P.skind), permute(P.snx), permute(P.sny), permute(P.snz);
			permute(P.sS);
			const int c0 = blk * cellsPerBlock, c1 = std::min(nE, c0 + cellsPerBlock);
			for (size_t i = (size_t) c0 * maxFaces; i < (size_t) c1 * maxFaces; i++) {
				const uint16_t code = P.csr[i];
				if (code == 0xFFFF) continue;
				P.csr[i] = (uint16_t) (newIdx[code & 0x7FFF] | (code & 0x8000));
			}
		}
	}

	// halo lists and block-local slot indices
	P.blockHalo.assign(P.nBlocks + 1, 0);
	P.spack.resize(P.se0.size());
	std::vector<int> halo;
	for (int blk = 0; blk < P.nBlocks; blk++) {
		int c0 = blk * cellsPerBlock, c1 = std::min(nE, c0 + cellsPerBlock);
		int s0 = P.blockSlot[blk], s1 = P.blockSlot[blk + 1];
		halo.clear();
		for (int s = s0; s < s1; s++)
			for (int cell : {P.se0[s], P.se1[s]})
				if (cell < nE && (cell < c0 || cell >= c1)) halo.push_back(cell);
		std::sort(halo.begin(), halo.end());
		halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
		if (cellsPerBlock + (int) halo.size() > 0x7FF) throw std::runtime_error("fused-stage block halo too large");
		auto local = [&](int cell) -> uint32_t {
			if (cell >= nE) return 0x7FF;
			if (cell >= c0 && cell < c1) return (uint32_t) (cell - c0);
			return (uint32_t) (cellsPerBlock + (std::lower_bound(halo.begin(), halo.end(), cell) - halo.begin()));
		};
		for (int s = s0; s < s1; s++)
			P.spack[s] = local(P.se0[s]) | local(P.se1[s]) << 11 | (uint32_t) P.skind[s] << 22;
		P.haloCell.insert(P.haloCell.end(), halo.begin(), halo.end());
		P.blockHalo[blk + 1] = (int) P.haloCell.size();
		P.maxHalo = std::max(P.maxHalo, (int) halo.size());
	}

	// the interface exchange: messages keep their entries contiguous in the buffer, as
	// MPIHandler::updateFlowField sends them (rho's, then u's, v's, w's, E's)
	if (exchange) {
		for (size_t k = 0; k + 1 < sm.haloStart.size(); k++) {
			const int a = sm.haloStart[k], b = sm.haloStart[k + 1], len = b - a;
			for (int i = a; i < b; i++) {
				P.xSrc.push_back(dev(sm.haloSrc[i]));
				P.xDst.push_back(sm.haloDst[i]);
				P.xSlot.push_back(5 * a + (i - a));
				P.xSlot.push_back(len);
			}
		}
		P.xBufLen = 5 * (int) P.xSrc.size();
	}

	// wall list for the force integration, in the reference's order (Mesh.hpp:594-627)
	size_t nw = sm.wallGhost.size();
	P.wallCell.resize(nw);
	for (int k = 0; k < 3; k++) P.wallN[k].resize(nw), P.wallC[k].resize(nw);
	for (size_t i = 0; i < nw; i++) {
		P.wallCell[i] = P.old2new[sm.wallCell[i]];
		int f = sm.wallFace[i];
		for (int k = 0; k < 3; k++) {
			P.wallN[k][i] = sm.faceNormal[3 * (size_t) f + k];
			P.wallC[k][i] = sm.faceCenter[3 * (size_t) f + k];
		}
	}
	return P;
}

}// namespace e3d
