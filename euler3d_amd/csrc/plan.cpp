// Device data plan: locality order of the cells, fused-stage blocks and device-numbered arrays.
//
// The device keeps cells in a Morton (Z-curve) order of their centroids, so kCellsPerBlock
// consecutive cells form a compact cluster whose faces are mostly shared inside the cluster.
// The reference numbering is kept where it determines the FP64 bits (face orientation and the
// per-cell order of accumulation, ascending reference face id, EulerSolver.cpp:306-358), so the
// order changes memory locality only.
//
// Planning is part of admitting a case to the batch engine, so it runs on up to 16 host threads:
// every step is either a map over faces, cells or blocks into disjoint outputs, or a per-block
// pass whose results are stitched by prefix sums. The plan does not depend on the thread count.
#include "plan.hpp"
#include "threads.hpp"

#include <algorithm>
#include <atomic>
#include <functional>
#include <thread>
#include <tuple>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <type_traits>
#include <stdexcept>

namespace e3d {

namespace {

// E3D_HALO_GROUPING=0 keeps the Morton order inside blocks
bool haloGrouping() {
	const char *v = std::getenv("E3D_HALO_GROUPING");
	return !(v && std::atoi(v) == 0);
}

int planThreads() {
	static const int n = [] {
		const unsigned h = std::thread::hardware_concurrency();
		return (int) std::max(1u, std::min(16u, h ? h : 1u));
	}();
	return n;
}

// f(lo, hi) over [0, n) in contiguous chunks of at least `grain` items, on up to maxT (default
// planThreads()) threads. A chunk that throws is reported after every chunk has ended.
template<typename F>
void parallelFor(size_t n, size_t grain, F &&f, int maxT = 0) {
	const size_t T = std::min<size_t>((size_t) (maxT > 0 ? maxT : planThreads()),
	                                  std::max<size_t>(1, n / std::max<size_t>(grain, 1)));
	if (T <= 1) {
		f((size_t) 0, n);
		return;
	}
	std::vector<std::function<void()>> tasks;
	tasks.reserve(T);
	for (size_t t = 0; t < T; t++) tasks.emplace_back([&, t] { f(n * t / T, n * (t + 1) / T); });
	runAll(tasks);
}

// sort under a strict total order: sorted runs on several threads, then pairwise merges
template<typename T, typename Less>
void parallelSort(std::vector<T> &v, Less less) {
	const size_t n = v.size();
	size_t P = 1;
	while (P * 2 <= (size_t) planThreads() && n / (P * 2) >= 32768) P *= 2;
	if (P == 1) {
		std::sort(v.begin(), v.end(), less);
		return;
	}
	std::vector<size_t> cut(P + 1);
	for (size_t i = 0; i <= P; i++) cut[i] = n * i / P;
	parallelFor(P, 1, [&](size_t a, size_t b) {
		for (size_t i = a; i < b; i++) std::sort(v.begin() + (long) cut[i], v.begin() + (long) cut[i + 1], less);
	});
	std::vector<T> tmp(n);
	for (size_t w = 1; w < P; w *= 2) {
		parallelFor(P / (2 * w), 1, [&](size_t a, size_t b) {
			for (size_t k = a; k < b; k++) {
				const size_t i = 2 * w * k;
				std::merge(v.begin() + (long) cut[i], v.begin() + (long) cut[i + w], v.begin() + (long) cut[i + w],
				           v.begin() + (long) cut[i + 2 * w], tmp.begin() + (long) cut[i], less);
			}
		});
		v.swap(tmp);
	}
}

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
	// the cells of a composite mesh stay grouped by partition (Morton order inside each), so that
	// per-partition reductions run over contiguous ranges
	struct Key {
		int part;
		int idx;
		uint64_t key;
	};
	std::vector<Key> keys(n);
	std::vector<int> part(n, 0);
	for (size_t r = 0; r + 1 < sm.partStart.size(); r++)
		for (int e = sm.partStart[r]; e < sm.partStart[r + 1]; e++) part[e] = (int) r;
	parallelFor((size_t) n, 1 << 14, [&](size_t a, size_t b) {
		for (size_t e = a; e < b; e++) {
			uint64_t q[3];
			for (int k = 0; k < 3; k++) {
				double t = (sm.cellCentroid[3 * e + k] - lo[k]) * scale;
				q[k] = (uint64_t) std::min(2097151.0, std::max(0.0, t));
			}
			keys[e] = {part[e], (int) e, spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2};
		}
	});
	parallelSort(keys, [](const Key &a, const Key &b) {
		if (a.part != b.part) return a.part < b.part;
		return a.key < b.key || (a.key == b.key && a.idx < b.idx);
	});
	std::vector<int> order(n);
	for (int i = 0; i < n; i++) order[i] = keys[i].idx;
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
// partition. Subtrees, the three axis selections of a large range and its passes over cells run
// on several threads; the result does not depend on their number.
std::vector<int> bisectOrder(const SolverMesh &sm, int cellsPerBlock) {
	// the bisection runs on Morton indices, over Morton-ordered copies of the centroids and of the
	// interior neighbours, so that its memory accesses stay local
	const std::vector<int> morton = mortonOrder(sm);
	const int n = sm.nElem;
	std::vector<int> rank(n);
	parallelFor((size_t) n, 1 << 15, [&](size_t a, size_t b) {
		for (size_t i = a; i < b; i++) rank[morton[i]] = (int) i;
	});
	std::vector<double> xyz(3 * (size_t) n);
	std::vector<int> nstart(n + 1, 0), nbr;
	auto eachNbr = [&](int i, auto &&g) {
		const int c = morton[i];
		for (int q = sm.element2faceStart[c]; q < sm.element2faceStart[c + 1]; q++) {
			const int f = sm.element2face[q];
			const int a = sm.face2element[2 * (size_t) f], o = a == c ? sm.face2element[2 * (size_t) f + 1] : a;
			if (o < n) g(rank[o]);
		}
	};
	parallelFor((size_t) n, 1 << 14, [&](size_t a, size_t b) {
		for (size_t i = a; i < b; i++) {
			const int c = morton[i];
			for (int k = 0; k < 3; k++) xyz[3 * i + k] = sm.cellCentroid[3 * (size_t) c + k];
			int cnt = 0;
			eachNbr((int) i, [&](int) { cnt++; });
			nstart[i + 1] = cnt;
		}
	});
	for (int i = 0; i < n; i++) nstart[i + 1] += nstart[i];
	nbr.resize(nstart[n]);
	parallelFor((size_t) n, 1 << 14, [&](size_t a, size_t b) {
		for (size_t i = a; i < b; i++) {
			int k = nstart[i];
			eachNbr((int) i, [&](int o) { nbr[k++] = o; });
		}
	});
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
	// f(i0, i1) over [lo, hi), on several threads when the range is large; subtrees at depth d run
	// side by side with 2^d - 1 others, so each takes a 2^-d share of the threads
	auto chunks = [](int lo, int hi, int depth, const std::function<void(int, int)> &f) {
		const int m = hi - lo, T = planThreads() >> std::min(depth, 30);
		if (m <= (1 << 16) || T <= 1) return f(lo, hi);
		parallelFor((size_t) m, 1 << 15, [&](size_t a, size_t b) { f(lo + (int) a, lo + (int) b); }, T);
	};
	struct KV {
		double x;
		int id;
	};
	std::function<void(int, int, int)> run = [&](int lo, int hi, int depth) {
		const int nb = (hi - lo + cellsPerBlock - 1) / cellsPerBlock;
		if (nb <= 1) {
			std::sort(ids.begin() + lo, ids.begin() + hi);
			return;
		}
		const int cut = lo + (nb / 2) * cellsPerBlock, half = cut - lo;
		const int id = ++jobs;
		chunks(lo, hi, depth, [&](int a, int b) {
			for (int i = a; i < b; i++) stamp[ids[i]].store(id, std::memory_order_relaxed);
		});
		// the median cell of each axis, selected on (coordinate, index) pairs held contiguously;
		// the chosen axis's selection is also the split (nth_element is deterministic)
		auto less = [](const KV &a, const KV &b) { return a.x < b.x || (a.x == b.x && a.id < b.id); };
		std::vector<KV> kv[3];
		double ext[3];
		int pivot[3];
		auto select = [&](int ax) {
			kv[ax].resize(hi - lo);
			double a = 1e300, b = -1e300;
			for (int i = lo; i < hi; i++) {
				const double x = xyz[3 * (size_t) ids[i] + ax];
				kv[ax][i - lo] = {x, ids[i]};
				a = std::min(a, x), b = std::max(b, x);
			}
			ext[ax] = b - a;
			std::nth_element(kv[ax].begin(), kv[ax].begin() + half, kv[ax].end(), less);
			pivot[ax] = kv[ax][half].id;
		};
		if (hi - lo > (1 << 15) && depth < 3) {
			runAll({[&] { select(0); }, [&] { select(1); }, [&] { select(2); }});
		} else {
			for (int ax = 0; ax < 3; ax++) select(ax);
		}
		// faces cut by each axis's split (each cut face counted from both sides)
		std::atomic<long> cuts[3];
		for (auto &x : cuts) x.store(0);
		chunks(lo, hi, depth, [&](int a, int b) {
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
		int longest = 0, best = 0;
		for (int ax = 1; ax < 3; ax++) {
			if (ext[ax] > ext[longest]) longest = ax;
			if (cuts[ax].load() < cuts[best].load()) best = ax;
		}
		if ((double) cuts[longest].load() <= 1.5 * (double) cuts[best].load()) best = longest;
		for (int i = lo; i < hi; i++) ids[i] = kv[best][i - lo].id;
		for (auto &v : kv) std::vector<KV>().swap(v);
		if (depth < 6 && hi - lo > (1 << 14)) {
			runAll({[&, cut, hi, depth] { run(cut, hi, depth + 1); }, [&, lo, cut, depth] { run(lo, cut, depth + 1); }});
		} else {
			run(lo, cut, depth + 1);
			run(cut, hi, depth + 1);
		}
	};
	std::vector<std::function<void()>> parts;
	for (size_t r = 0; r + 1 < bounds.size(); r++) parts.emplace_back([&, r] { run(bounds[r], bounds[r + 1], 0); });
	runAll(parts);
	parallelFor((size_t) n, 1 << 15, [&](size_t a, size_t b) {
		for (size_t i = a; i < b; i++) ids[i] = morton[ids[i]];
	});
	return ids;
}

// Within each block, the cells that no other block reads come first, then the cells that other
// blocks read as halo, grouped by the lowest-numbered such block (in their previous order inside
// each group). A block's halo then arrives in runs of consecutive cells of each neighbouring block,
// so that its halo loads touch fewer lines. Only the order inside blocks changes, and inside the
// part of a block that each partition of a composite mesh holds: a block may straddle two
// partitions, whose cells must keep their contiguous device ranges (the per-partition smoothing
// decisions and coefficient sums read those ranges).
void groupHaloSources(const SolverMesh &sm, int cellsPerBlock, std::vector<int> &new2old) {
	const int n = sm.nElem;
	std::vector<int> old2new(n);
	for (int i = 0; i < n; i++) old2new[new2old[i]] = i;
	const int nb = (n + cellsPerBlock - 1) / cellsPerBlock;
	// partition of each device position (the bisection keeps partitions in their ranges)
	auto partOf = [&](int i) -> int {
		if (sm.partStart.size() <= 2) return 0;
		return (int) (std::upper_bound(sm.partStart.begin(), sm.partStart.end(), i) - sm.partStart.begin()) - 1;
	};
	parallelFor((size_t) nb, 16, [&](size_t ba, size_t bb) {
		std::vector<std::tuple<int, int, int>> key;
		for (size_t blk = ba; blk < bb; blk++) {
			const int c0 = (int) blk * cellsPerBlock, c1 = std::min(n, c0 + cellsPerBlock);
			key.clear();
			for (int i = c0; i < c1; i++) {
				const int c = new2old[i];
				int sig = -1;
				for (int q = sm.element2faceStart[c]; q < sm.element2faceStart[c + 1]; q++) {
					const int f = sm.element2face[q];
					const int a = sm.face2element[2 * (size_t) f], o = a == c ? sm.face2element[2 * (size_t) f + 1] : a;
					if (o >= n) continue;
					const int ob = old2new[o] / cellsPerBlock;
					if (ob != (int) blk && (sig < 0 || ob < sig)) sig = ob;
				}
				key.emplace_back(partOf(i), sig, i);
			}
			std::sort(key.begin(), key.end());
			std::vector<int> cells(key.size());
			for (size_t k = 0; k < key.size(); k++) cells[k] = new2old[std::get<2>(key[k])];
			std::copy(cells.begin(), cells.end(), new2old.begin() + c0);
		}
	});
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

std::vector<int> cellOrder(const SolverMesh &sm, int order, int cellsPerBlock) {
	std::vector<int> new2old;
	if (order == ORDER_MORTON) {
		new2old = mortonOrder(sm);
	} else if (order == ORDER_BISECT) {
		new2old = bisectOrder(sm, cellsPerBlock);
		if (haloGrouping()) groupHaloSources(sm, cellsPerBlock, new2old);
	} else {
		new2old.resize(sm.nElem);
		std::iota(new2old.begin(), new2old.end(), 0);
	}
	return new2old;
}

HostPlan buildPlan(const SolverMesh &sm, int order, int cellsPerBlock, bool exchange) {
	HostPlan P;
	const int nE = sm.nElem, nF = sm.nFace;
	P.nElem = nE;
	P.nFace = nF;
	P.nFaceInt = sm.nFaceInt;
	P.nGhost = sm.nGhost;
	P.new2old = cellOrder(sm, order, cellsPerBlock);
	P.old2new.assign(nE, -1);
	for (int i = 0; i < nE; i++) P.old2new[P.new2old[i]] = i;
	auto dev = [&](int cell) { return cell < nE ? P.old2new[cell] : cell; };

	// faces in reference order
	P.fe0.resize(nF), P.fe1.resize(nF);
	P.fnx.resize(nF), P.fny.resize(nF), P.fnz.resize(nF), P.fS.resize(nF);
	parallelFor((size_t) nF, 1 << 15, [&](size_t a, size_t b) {
		for (size_t f = a; f < b; f++) {
			P.fe0[f] = dev(sm.face2element[2 * f]);
			P.fe1[f] = dev(sm.face2element[2 * f + 1]);
			if (exchange && (int) f >= sm.nFaceInt && sm.ghostKind[f - sm.nFaceInt] == BND_MPI)
				P.fe1[f] = nE + ((int) f - sm.nFaceInt);
			P.fnx[f] = sm.faceUnitNormal[3 * f];
			P.fny[f] = sm.faceUnitNormal[3 * f + 1];
			P.fnz[f] = sm.faceUnitNormal[3 * f + 2];
			P.fS[f] = sm.faceArea[f];
		}
	});
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
	parallelFor((size_t) nE, 1 << 14, [&](size_t ca, size_t cb) {
		for (size_t c = ca; c < cb; c++) {
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
	});

	// Fused-stage blocks: every face touching the block gets one slot in it, numbered in order of
	// first appearance over the block's cells and their faces. Within each block the interior
	// slots then come first, then the boundary slots grouped by kind (wall, symmetry, farfield,
	// partition interface), each group in its order of first appearance (a stable counting sort
	// by kind). The flux loop hands 64 consecutive slots to a wave, so a boundary face's ghost and
	// flux code then runs in the few waves that hold boundary slots instead of diverging the waves
	// of every block at the wall. The gather order is the CSR order (ascending reference face
	// id), which the renumbering keeps, so results do not change.
	P.cellsPerBlock = cellsPerBlock;
	P.nBlocks = (nE + cellsPerBlock - 1) / cellsPerBlock;
	P.csrStride = maxFaces;
	P.csr.assign((size_t) nE * maxFaces, 0xFFFF);
	P.blockSlot.assign(P.nBlocks + 1, 0);
	std::vector<std::vector<int>> bfaces(P.nBlocks); // per block: its slots' reference face ids
	std::atomic<bool> tooMany{false};
	parallelFor((size_t) P.nBlocks, 16, [&](size_t ba, size_t bb) {
		// face id -> slot of the current block: open addressing, entries of earlier blocks are
		// told apart by their block stamp
		size_t cap = 16;
		while (cap < 4 * (size_t) cellsPerBlock * std::max(1, maxFaces)) cap *= 2;
		std::vector<int> hkey(cap), hval(cap), hblk(cap, -1);
		std::vector<int> kindOf, newIdx;
		std::vector<int> faces;
		for (size_t blk = ba; blk < bb; blk++) {
			const int c0 = (int) blk * cellsPerBlock, c1 = std::min(nE, c0 + cellsPerBlock);
			faces.clear();
			for (int c = c0; c < c1; c++) {
				for (int j = P.fstart[c]; j < P.fstart[c + 1]; j++) {
					unsigned code = (unsigned) P.flist[j];
					int f = (int) (code & 0x7fffffffu);
					size_t h = ((uint32_t) f * 2654435761u) & (cap - 1);
					while (hblk[h] == (int) blk && hkey[h] != f) h = (h + 1) & (cap - 1);
					int slot;
					if (hblk[h] == (int) blk) {
						slot = hval[h];
					} else {
						slot = (int) faces.size();
						faces.push_back(f);
						hblk[h] = (int) blk, hkey[h] = f, hval[h] = slot;
					}
					if (slot > 0x7FFF) tooMany = true;
					P.csr[(size_t) c * maxFaces + (j - P.fstart[c])] = (uint16_t) (slot | ((code >> 31) << 15));
				}
			}
			const int ns = (int) faces.size();
			int count[6] = {0, 0, 0, 0, 0, 0};
			kindOf.resize(ns);
			for (int j = 0; j < ns; j++) count[(kindOf[j] = slotKind(sm, faces[j], exchange)) + 1]++;
			for (int k = 1; k < 6; k++) count[k] += count[k - 1];
			newIdx.resize(ns);
			for (int j = 0; j < ns; j++) newIdx[j] = count[kindOf[j]]++;
			std::vector<int> &out = bfaces[blk];
			out.resize(ns);
			for (int j = 0; j < ns; j++) out[newIdx[j]] = faces[j];
			for (size_t i = (size_t) c0 * maxFaces; i < (size_t) c1 * maxFaces; i++) {
				const uint16_t code = P.csr[i];
				if (code == 0xFFFF) continue;
				P.csr[i] = (uint16_t) (newIdx[code & 0x7FFF] | (code & 0x8000));
			}
		}
	});
	if (tooMany) throw std::runtime_error("fused-stage block has more than 32768 face slots");
	for (int blk = 0; blk < P.nBlocks; blk++) {
		const int ns = (int) bfaces[blk].size();
		P.blockSlot[blk + 1] = P.blockSlot[blk] + ns;
		P.maxSlots = std::max(P.maxSlots, ns);
	}
	const size_t nSlots = (size_t) P.blockSlot[P.nBlocks];
	P.se0.resize(nSlots), P.se1.resize(nSlots), P.skind.resize(nSlots);
	P.snx.resize(nSlots), P.sny.resize(nSlots), P.snz.resize(nSlots), P.sS.resize(nSlots);
	P.spack.resize(nSlots);

	// halo lists and block-local slot indices
	P.blockHalo.assign(P.nBlocks + 1, 0);
	std::vector<std::vector<int>> bhalo(P.nBlocks);
	std::atomic<bool> haloTooLarge{false};
	parallelFor((size_t) P.nBlocks, 16, [&](size_t ba, size_t bb) {
		for (size_t blk = ba; blk < bb; blk++) {
			const int c0 = (int) blk * cellsPerBlock, c1 = std::min(nE, c0 + cellsPerBlock);
			const int s0 = P.blockSlot[blk];
			const std::vector<int> &faces = bfaces[blk];
			const int ns = (int) faces.size();
			std::vector<int> &halo = bhalo[blk];
			for (int j = 0; j < ns; j++) {
				const int f = faces[j], s = s0 + j;
				P.se0[s] = P.fe0[f];
				P.se1[s] = P.fe1[f];
				P.skind[s] = slotKind(sm, f, exchange);
				P.snx[s] = P.fnx[f];
				P.sny[s] = P.fny[f];
				P.snz[s] = P.fnz[f];
				P.sS[s] = P.fS[f];
				for (int cell : {P.se0[s], P.se1[s]})
					if (cell < nE && (cell < c0 || cell >= c1)) halo.push_back(cell);
			}
			std::sort(halo.begin(), halo.end());
			halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
			if (cellsPerBlock + (int) halo.size() > 0x7FF) {
				haloTooLarge = true;
				continue;
			}
			auto local = [&](int cell) -> uint32_t {
				if (cell >= nE) return 0x7FF;
				if (cell >= c0 && cell < c1) return (uint32_t) (cell - c0);
				return (uint32_t) (cellsPerBlock + (std::lower_bound(halo.begin(), halo.end(), cell) - halo.begin()));
			};
			for (int s = s0; s < s0 + ns; s++)
				P.spack[s] = local(P.se0[s]) | local(P.se1[s]) << 11 | (uint32_t) P.skind[s] << 22;
		}
	});
	if (haloTooLarge) throw std::runtime_error("fused-stage block halo too large");
	std::vector<std::vector<int>>().swap(bfaces);
	for (int blk = 0; blk < P.nBlocks; blk++) {
		P.blockHalo[blk + 1] = P.blockHalo[blk] + (int) bhalo[blk].size();
		P.maxHalo = std::max(P.maxHalo, (int) bhalo[blk].size());
	}
	P.haloCell.resize(P.blockHalo[P.nBlocks]);
	parallelFor((size_t) P.nBlocks, 64, [&](size_t ba, size_t bb) {
		for (size_t blk = ba; blk < bb; blk++) std::copy(bhalo[blk].begin(), bhalo[blk].end(), P.haloCell.begin() + P.blockHalo[blk]);
	});

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
