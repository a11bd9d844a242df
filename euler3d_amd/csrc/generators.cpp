// Synthetic mesh generators for the solver's benchmark and parity workloads. All randomness comes
// from splitmix64, so a (generator, arguments, seed) triple always produces the same file.
//
//   genBox       Cartesian hex box, one tag per side (the reference's StructuredBlock_8 fixture
//                family, test/mesh/StructuredBlock_8.su2).
//   genNacaOgrid NACA 4-digit section, periodic O-grid of hexes extruded in z, tags wall /
//                farfield / sym (the naca0012_129x129x2 example, examples/naca0012_129x129x2_transsonic_125).
//   genBumpTets  3L x L x L channel with a 10%-thick circular-arc bump, each hex split into 6
//                tets, seeded interior-node jitter and random cell/node numbering; inlet and
//                outlet tagged farfield, all four walls tagged wall.
//   genWing      swept, tapered wing of symmetric 4-digit section in a box, H-H hex grid with the
//                wing as a doubled internal sheet; wing surface wall, root plane sym, outer box
//                farfield.
#include "e3d_host.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <map>
#include <stdexcept>

namespace e3d {

namespace {

// Appends boundary markers to a mesh whose volume cells are complete: every (cell, local face)
// slot that no other cell shares becomes a marker element, tagged by classify(face nodes) which
// returns an index into tags. Marker elements keep the owner's local face node order.
void addBoundaryMarkers(MeshFile &m, const std::vector<std::string> &tags,
                        const std::function<int(const int *nodes, int n)> &classify) {
	struct Slot {
		std::array<int, 4> k;
		int elem, face;
	};
	std::vector<Slot> slots;
	slots.reserve((size_t) m.nElem() * 6);
	for (int e = 0; e < m.nElem(); e++) {
		int vtk = m.elemVtk[e];
		const int *en = &m.elemNodes[m.elemStart[e]];
		for (int f = 0; f < localFaceCount(vtk); f++) {
			int ln[4];
			int n = localFaceNodes(vtk, f, ln);
			std::array<int, 4> k{-1, -1, -1, -1};
			for (int i = 0; i < n; i++) k[i] = en[ln[i]];
			std::sort(k.begin(), k.begin() + n);
			slots.push_back({k, e, f});
		}
	}
	std::vector<uint32_t> order(slots.size());
	for (size_t i = 0; i < order.size(); i++) order[i] = (uint32_t) i;
	std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
		return slots[a].k < slots[b].k || (slots[a].k == slots[b].k && a < b);
	});
	std::vector<uint8_t> boundary(slots.size(), 0);
	for (size_t i = 0; i < order.size();) {
		size_t j = i + 1;
		while (j < order.size() && slots[order[j]].k == slots[order[i]].k) j++;
		if (j - i == 1) boundary[order[i]] = 1;
		i = j;
	}
	m.markers.assign(tags.size(), Marker{});
	for (size_t t = 0; t < tags.size(); t++) {
		m.markers[t].tag = tags[t];
		m.markers[t].nodeStart.push_back(0);
	}
	for (size_t s = 0; s < slots.size(); s++) {
		if (!boundary[s]) continue;
		int e = slots[s].elem;
		int vtk = m.elemVtk[e];
		int ln[4];
		int n = localFaceNodes(vtk, slots[s].face, ln);
		int nodes[4];
		for (int i = 0; i < n; i++) nodes[i] = m.elemNodes[m.elemStart[e] + ln[i]];
		int t = classify(nodes, n);
		if (t < 0 || t >= (int) tags.size()) throw std::runtime_error("generator: unclassified boundary face");
		Marker &mk = m.markers[t];
		mk.vtk.push_back(n == 3 ? VTK_TRI : VTK_QUAD);
		for (int i = 0; i < n; i++) mk.nodes.push_back(nodes[i]);
		mk.nodeStart.push_back((int) mk.nodes.size());
	}
	// drop empty markers
	std::vector<Marker> kept;
	for (auto &mk : m.markers)
		if (!mk.vtk.empty()) kept.push_back(std::move(mk));
	m.markers = std::move(kept);
}

void pushElem(MeshFile &m, int vtk, std::initializer_list<int> nodes) {
	m.elemVtk.push_back(vtk);
	for (int n : nodes) m.elemNodes.push_back(n);
	m.elemStart.push_back((int) m.elemNodes.size());
}

double nacaThickness(double x, double t) {
	// closed trailing edge 4-digit thickness (coefficient -0.1036), t = thickness / chord
	double xs = x < 0 ? 0 : x;
	return 5.0 * t * (0.2969 * std::sqrt(xs) - 0.1260 * xs - 0.3516 * xs * xs + 0.2843 * xs * xs * xs - 0.1036 * xs * xs * xs * xs);
}

double nacaCamber(double x, double m, double p) {
	if (m <= 0.0 || p <= 0.0) return 0.0;
	if (x < p) return m / (p * p) * (2 * p * x - x * x);
	return m / ((1 - p) * (1 - p)) * ((1 - 2 * p) + 2 * p * x - x * x);
}

// ratio r of a geometric progression with first step ds whose n steps sum to L
double geometricRatio(double ds, double L, int n) {
	if (ds * n >= L) return 1.0;
	double lo = 1.0 + 1e-12, hi = 2.0;
	while (ds * (std::pow(hi, n) - 1) / (hi - 1) < L) hi *= 2;
	for (int it = 0; it < 200; it++) {
		double mid = 0.5 * (lo + hi);
		double s = ds * (std::pow(mid, n) - 1) / (mid - 1);
		if (s < L) lo = mid;
		else hi = mid;
	}
	return 0.5 * (lo + hi);
}

// normalised geometric distribution eta_0 = 0 .. eta_n = 1 with first step ds / L
std::vector<double> geometricEta(double ds, double L, int n) {
	std::vector<double> eta(n + 1);
	double r = geometricRatio(ds, L, n);
	if (r == 1.0) {
		for (int j = 0; j <= n; j++) eta[j] = (double) j / n;
		return eta;
	}
	double total = (std::pow(r, n) - 1) / (r - 1);
	for (int j = 0; j <= n; j++) eta[j] = ((std::pow(r, j) - 1) / (r - 1)) / total;
	eta[n] = 1.0;
	return eta;
}

}// namespace

MeshFile genBox(int nx, int ny, int nz, double lx, double ly, double lz, const char *tags6[6]) {
	if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("genBox: sizes must be >= 1");
	MeshFile m;
	auto id = [&](int i, int j, int k) { return (k * (ny + 1) + j) * (nx + 1) + i; };
	m.xyz.resize(3 * (size_t) (nx + 1) * (ny + 1) * (nz + 1));
	for (int k = 0; k <= nz; k++)
		for (int j = 0; j <= ny; j++)
			for (int i = 0; i <= nx; i++) {
				size_t n = id(i, j, k);
				m.xyz[3 * n] = lx * i / nx;
				m.xyz[3 * n + 1] = ly * j / ny;
				m.xyz[3 * n + 2] = lz * k / nz;
			}
	m.elemStart.push_back(0);
	for (int k = 0; k < nz; k++)
		for (int j = 0; j < ny; j++)
			for (int i = 0; i < nx; i++)
				pushElem(m, VTK_HEX, {id(i, j, k), id(i + 1, j, k), id(i + 1, j + 1, k), id(i, j + 1, k), id(i, j, k + 1),
				                      id(i + 1, j, k + 1), id(i + 1, j + 1, k + 1), id(i, j + 1, k + 1)});
	// merge repeated tag names into one marker, in first-appearance order
	std::vector<std::string> tags;
	int side2tag[6];
	for (int s = 0; s < 6; s++) {
		auto it = std::find(tags.begin(), tags.end(), std::string(tags6[s]));
		if (it == tags.end()) {
			side2tag[s] = (int) tags.size();
			tags.push_back(tags6[s]);
		} else {
			side2tag[s] = (int) (it - tags.begin());
		}
	}
	addBoundaryMarkers(m, tags, [&](const int *nodes, int n) {
		int ii[4], jj[4], kk[4];
		for (int q = 0; q < n; q++) {
			ii[q] = nodes[q] % (nx + 1);
			jj[q] = (nodes[q] / (nx + 1)) % (ny + 1);
			kk[q] = nodes[q] / ((nx + 1) * (ny + 1));
		}
		auto all = [&](int *a, int v) {
			for (int q = 0; q < n; q++)
				if (a[q] != v) return false;
			return true;
		};
		if (all(ii, 0)) return side2tag[0];
		if (all(ii, nx)) return side2tag[1];
		if (all(jj, 0)) return side2tag[2];
		if (all(jj, ny)) return side2tag[3];
		if (all(kk, 0)) return side2tag[4];
		if (all(kk, nz)) return side2tag[5];
		return -1;
	});
	return m;
}

MeshFile genNacaOgrid(int ni, int nj, int nk, double t, double camber, double camberPos, double wallSpacing,
                      double farR, double dz, uint64_t) {
	// ni nodes around including the repeated cut node (ni - 1 unique), nj nodes wall -> farfield,
	// nk node layers in z.
	if (ni < 5 || nj < 2 || nk < 2) throw std::invalid_argument("genNacaOgrid: sizes too small");
	const int nu = ni - 1;
	MeshFile m;
	auto id = [&](int i, int j, int k) { return (k * nj + j) * nu + (i % nu); };
	std::vector<double> sx(nu), sy(nu), fx(nu), fy(nu);
	const double pi = 3.141592653589793238;
	for (int i = 0; i < nu; i++) {
		double th = 2.0 * pi * i / nu;
		double x = 0.5 * (1.0 + std::cos(th));
		double yc = nacaCamber(x, camber, camberPos);
		double yt = nacaThickness(x, t);
		sx[i] = x;
		sy[i] = (i == 0) ? yc : (th <= pi ? yc - yt : yc + yt);
		fx[i] = 0.5 + farR * std::cos(th);
		fy[i] = -farR * std::sin(th);
	}
	std::vector<double> eta = geometricEta(wallSpacing, farR, nj - 1);
	m.xyz.resize(3 * (size_t) nu * nj * nk);
	for (int k = 0; k < nk; k++)
		for (int j = 0; j < nj; j++)
			for (int i = 0; i < nu; i++) {
				size_t n = id(i, j, k);
				m.xyz[3 * n] = sx[i] + eta[j] * (fx[i] - sx[i]);
				m.xyz[3 * n + 1] = sy[i] + eta[j] * (fy[i] - sy[i]);
				m.xyz[3 * n + 2] = dz * k;
			}
	m.elemStart.push_back(0);
	for (int k = 0; k < nk - 1; k++)
		for (int j = 0; j < nj - 1; j++)
			for (int i = 0; i < nu; i++)
				pushElem(m, VTK_HEX, {id(i, j, k), id(i + 1, j, k), id(i + 1, j + 1, k), id(i, j + 1, k), id(i, j, k + 1),
				                      id(i + 1, j, k + 1), id(i + 1, j + 1, k + 1), id(i, j + 1, k + 1)});
	addBoundaryMarkers(m, {"farfield", "wall", "sym"}, [&](const int *nodes, int n) {
		bool allJ0 = true, allJN = true, allK0 = true, allKN = true;
		for (int q = 0; q < n; q++) {
			int j = (nodes[q] / nu) % nj;
			int k = nodes[q] / (nu * nj);
			allJ0 &= j == 0;
			allJN &= j == nj - 1;
			allK0 &= k == 0;
			allKN &= k == nk - 1;
		}
		if (allJN) return 0;
		if (allJ0) return 1;
		if (allK0 || allKN) return 2;
		return -1;
	});
	return m;
}

MeshFile genBumpTets(int nx, int ny, int nz, double jitter, uint64_t seed, bool permute) {
	if (nx < 3 || ny < 1 || nz < 1) throw std::invalid_argument("genBumpTets: sizes too small");
	const double L = 1.0, Lx = 3.0 * L;
	// circular arc through (L,0), (2L,0), (1.5L, 0.1L)
	const double h = 0.1 * L, half = 0.5 * L;
	const double R = (half * half + h * h) / (2 * h);
	auto yb = [&](double x) {
		if (x <= L || x >= 2 * L) return 0.0;
		double d = x - 1.5 * L;
		return std::sqrt(R * R - d * d) - (R - h);
	};
	MeshFile m;
	const int NX = nx + 1, NY = ny + 1, NZ = nz + 1;
	auto id = [&](int i, int j, int k) { return (k * NY + j) * NX + i; };
	const size_t nNode = (size_t) NX * NY * NZ;
	m.xyz.resize(3 * nNode);
	SplitMix64 rng(seed);
	for (int k = 0; k < NZ; k++)
		for (int j = 0; j < NY; j++)
			for (int i = 0; i < NX; i++) {
				double x = Lx * i / nx;
				double y0 = yb(x);
				double y = y0 + (L - y0) * j / ny;
				double z = L * k / nz;
				bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
				if (interior && jitter > 0) {
					double dx = Lx / nx, dy = (L - y0) / ny, dzz = L / nz;
					x += jitter * dx * (2.0 * rng.uniform() - 1.0);
					y += jitter * dy * (2.0 * rng.uniform() - 1.0);
					z += jitter * dzz * (2.0 * rng.uniform() - 1.0);
				}
				size_t n = id(i, j, k);
				m.xyz[3 * n] = x;
				m.xyz[3 * n + 1] = y;
				m.xyz[3 * n + 2] = z;
			}
	// Kuhn split of each hex into 6 tets along the (0,0,0)-(1,1,1) diagonal; conforming between
	// neighbouring hexes because every hex uses the same diagonal direction.
	static const int kuhn[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
	m.elemStart.reserve((size_t) nx * ny * nz * 6 + 1);
	m.elemStart.push_back(0);
	m.elemVtk.reserve((size_t) nx * ny * nz * 6);
	m.elemNodes.reserve((size_t) nx * ny * nz * 24);
	for (int k = 0; k < nz; k++)
		for (int j = 0; j < ny; j++)
			for (int i = 0; i < nx; i++) {
				int c[8] = {id(i, j, k), id(i + 1, j, k), id(i + 1, j + 1, k), id(i, j + 1, k),
				            id(i, j, k + 1), id(i + 1, j, k + 1), id(i + 1, j + 1, k + 1), id(i, j + 1, k + 1)};
				for (auto &t : kuhn) pushElem(m, VTK_TET, {c[t[0]], c[t[1]], c[t[2]], c[t[3]]});
			}
	addBoundaryMarkers(m, {"farfield", "wall"}, [&](const int *nodes, int n) {
		bool allI0 = true, allIN = true;
		for (int q = 0; q < n; q++) {
			int i = nodes[q] % NX;
			allI0 &= i == 0;
			allIN &= i == nx;
		}
		if (allI0 || allIN) return 0;
		return 1;
	});
	if (permute) {
		// random node numbering and random cell order (Fisher-Yates on splitmix64)
		SplitMix64 prng(seed ^ 0x5bd1e995ULL);
		std::vector<int> nodePerm(nNode);
		for (size_t i = 0; i < nNode; i++) nodePerm[i] = (int) i;
		for (size_t i = nNode - 1; i > 0; i--) std::swap(nodePerm[i], nodePerm[prng.next() % (i + 1)]);
		std::vector<double> xyz(3 * nNode);
		for (size_t i = 0; i < nNode; i++)
			for (int c = 0; c < 3; c++) xyz[3 * (size_t) nodePerm[i] + c] = m.xyz[3 * i + c];
		m.xyz.swap(xyz);
		for (auto &n : m.elemNodes) n = nodePerm[n];
		for (auto &mk : m.markers)
			for (auto &n : mk.nodes) n = nodePerm[n];
		const size_t nE = m.elemVtk.size();
		std::vector<int> cellOrder(nE);
		for (size_t i = 0; i < nE; i++) cellOrder[i] = (int) i;
		for (size_t i = nE - 1; i > 0; i--) std::swap(cellOrder[i], cellOrder[prng.next() % (i + 1)]);
		std::vector<int> vtk(nE), start(nE + 1), nodes;
		nodes.reserve(m.elemNodes.size());
		start[0] = 0;
		for (size_t i = 0; i < nE; i++) {
			int e = cellOrder[i];
			vtk[i] = m.elemVtk[e];
			for (int q = m.elemStart[e]; q < m.elemStart[e + 1]; q++) nodes.push_back(m.elemNodes[q]);
			start[i + 1] = (int) nodes.size();
		}
		m.elemVtk.swap(vtk);
		m.elemStart.swap(start);
		m.elemNodes.swap(nodes);
	}
	return m;
}

MeshFile genWing(int ni, int nj, int nk, double t, double sweepDeg, double taper, double semispan, double farfield) {
	// H-H grid: i streamwise, j normal (y), k spanwise (z >= 0, root plane z = 0).
	// The wing is the sheet j = jw over the planform cells; its nodes are doubled there (a lower
	// copy used by cells below, an upper copy by cells above) and displaced by the section
	// thickness. Planform edges (leading edge, trailing edge, tip) have zero thickness, so the
	// two copies coincide there and single nodes are used.
	if (ni < 9 || nj < 5 || nk < 5) throw std::invalid_argument("genWing: sizes too small");
	const double pi = 3.141592653589793238;
	const double tanS = std::tan(sweepDeg * pi / 180.0);
	const int jw = (nj - 1) / 2;
	// chordwise node range on the wing and spanwise extent
	const int iLE = (ni - 1) / 4, iTE = iLE + (ni - 1) / 2;
	const int kTip = (nk - 1) * 2 / 3;
	const int nChord = iTE - iLE;
	auto chord = [&](double z) { return 1.0 - (1.0 - taper) * std::min(z, semispan) / semispan; };
	auto xle = [&](double z) { return std::min(z, semispan) * tanS; };
	// spanwise stations: clustered to the tip, stretched to the outer box beyond
	std::vector<double> zk(nk);
	for (int k = 0; k <= kTip; k++) zk[k] = semispan * (1.0 - std::cos(0.5 * pi * k / kTip));
	{
		double dz0 = zk[kTip] - zk[kTip - 1];
		std::vector<double> eta = geometricEta(dz0, farfield - semispan, nk - 1 - kTip);
		for (int k = kTip; k < nk; k++) zk[k] = semispan + eta[k - kTip] * (farfield - semispan);
	}
	// chordwise coordinate xi in [0,1] on the wing (cosine clustering at LE and TE)
	std::vector<double> xi(nChord + 1);
	for (int q = 0; q <= nChord; q++) xi[q] = 0.5 * (1.0 - std::cos(pi * q / nChord));
	// vertical distribution from the sheet to +-farfield: jw cells below, nj - 1 - jw above (one
	// more when nj is even)
	std::vector<double> etaLo = geometricEta(0.02 * 1.0 / 1.0 * 0.25, farfield, jw);
	std::vector<double> etaHi = geometricEta(0.02 * 1.0 / 1.0 * 0.25, farfield, nj - 1 - jw);
	auto xAt = [&](int i, double z) {
		double x0 = xle(z), c = chord(z);
		if (i >= iLE && i <= iTE) return x0 + c * xi[i - iLE];
		if (i < iLE) {
			double ds = c * (xi[1] - xi[0]);
			std::vector<double> e = geometricEta(ds, x0 + farfield, iLE);
			return x0 - e[iLE - i] * (x0 + farfield);
		}
		double xte = x0 + c;
		double ds = c * (xi[nChord] - xi[nChord - 1]);
		int n = ni - 1 - iTE;
		std::vector<double> e = geometricEta(ds, farfield + semispan * tanS + 1.0 - xte, n);
		return xte + e[i - iTE] * (farfield + semispan * tanS + 1.0 - xte);
	};
	auto onWing = [&](int i, int k) { return i >= iLE && i <= iTE && k <= kTip; };
	// thickness of the sheet at node (i,k): zero on the planform edges
	auto halfThick = [&](int i, int k) {
		if (!onWing(i, k) || i == iLE || i == iTE || k == kTip) return 0.0;
		double z = zk[k];
		double c = chord(z);
		double taperTip = std::sqrt(std::max(0.0, 1.0 - std::pow(z / semispan, 8.0)));
		return c * nacaThickness(xi[i - iLE], t) * taperTip;
	};
	MeshFile m;
	// base node ids, plus an upper copy for interior planform nodes of the sheet
	auto base = [&](int i, int j, int k) { return (k * nj + j) * ni + i; };
	const int nBase = ni * nj * nk;
	std::vector<int> upperId((size_t) ni * nk, -1);
	std::vector<int> upperK; // k of each upper copy, indexed by id - nBase
	int next = nBase;
	for (int k = 0; k < nk; k++)
		for (int i = 0; i < ni; i++)
			if (onWing(i, k) && i != iLE && i != iTE && k != kTip) {
				upperId[k * ni + i] = next++;
				upperK.push_back(k);
			}
	m.xyz.resize(3 * (size_t) next);
	for (int k = 0; k < nk; k++) {
		double z = zk[k];
		for (int i = 0; i < ni; i++) {
			double x = xAt(i, z);
			double ht = halfThick(i, k);
			for (int j = 0; j < nj; j++) {
				double y;
				if (j < jw) y = -ht - etaLo[jw - j] * (farfield - ht);
				else if (j > jw) y = ht + etaHi[j - jw] * (farfield - ht);
				else y = -ht; // lower copy (or the single sheet node)
				size_t n = base(i, j, k);
				m.xyz[3 * n] = x;
				m.xyz[3 * n + 1] = y;
				m.xyz[3 * n + 2] = z;
			}
			int u = upperId[k * ni + i];
			if (u >= 0) {
				m.xyz[3 * (size_t) u] = x;
				m.xyz[3 * (size_t) u + 1] = ht;
				m.xyz[3 * (size_t) u + 2] = z;
			}
		}
	}
	auto node = [&](int i, int j, int k, bool above) {
		if (j == jw && above) {
			int u = upperId[k * ni + i];
			if (u >= 0) return u;
		}
		return base(i, j, k);
	};
	m.elemStart.push_back(0);
	for (int k = 0; k < nk - 1; k++)
		for (int j = 0; j < nj - 1; j++)
			for (int i = 0; i < ni - 1; i++) {
				// a cell just above the sheet uses the upper copies of its j == jw nodes
				bool above = (j == jw);
				pushElem(m, VTK_HEX, {node(i, j, k, above), node(i + 1, j, k, above), node(i + 1, j + 1, k, above),
				                      node(i, j + 1, k, above), node(i, j, k + 1, above), node(i + 1, j, k + 1, above),
				                      node(i + 1, j + 1, k + 1, above), node(i, j + 1, k + 1, above)});
			}
	addBoundaryMarkers(m, {"wall", "sym", "farfield"}, [&](const int *nodes, int n) {
		// wing sheet faces: every node on the sheet row j == jw (base or upper copy)
		bool allSheet = true, allK0 = true;
		for (int q = 0; q < n; q++) {
			int id = nodes[q];
			int j, k;
			if (id >= nBase) {
				j = jw;
				k = upperK[id - nBase];
			} else {
				j = (id / ni) % nj;
				k = id / (ni * nj);
			}
			allSheet &= j == jw;
			allK0 &= k == 0;
		}
		if (allK0) return 1;
		if (allSheet) return 0;
		return 2;
	});
	return m;
}

}// namespace e3d
