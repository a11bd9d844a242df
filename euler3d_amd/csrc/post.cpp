// Tecplot ASCII writer, the E3D_POST post-processor (src/post/post.cpp:18-133,
// src/parser/SolutionPost.cpp:12-59, src/postprocessor.cpp:17-40).
//
// File layout, as post.cpp:49-133 writes it:
//   VARIABLES="X","Y","Z","Rho","U","V","W","P","E"
//   per partition: a FEBRICK zone header with the node count and the file's NELEM count, the node
//   X, then Y, then Z (one "%.12e" per line), the cell-centred rho, u, v, w, p, E from the
//   partition's .sol file (same format), then one line of eight 1-based node ids per volume
//   element, tab-separated. Tets repeat nodes as (0,1,2,2,3,3,3,3), pyramids (0,1,2,3,4,4,4,4),
//   prisms (0,1,1,2,3,4,4,5); hexes are written as they are, any other kind as its own node list
//   with a tab after each id.
//
// .sol reading follows SolutionPost::parseFile: the header line is skipped and NELEM lines are
// read with six numbers each. A number missing from a short or absent line keeps the value read
// before it (the reference extracts into variables that persist from line to line); a number that
// does not parse reads as 0 and ends that line.
//
// Formatting is the same printf conversion the reference uses, done by a pool of threads over
// blocks of lines and written in order, so the bytes are identical and multi-million-cell
// partitions write at disk speed.
#include "e3d_host.hpp"
#include "threads.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace e3d {

namespace {

int workerCount() {
	unsigned h = std::thread::hardware_concurrency();
	return (int) std::max(1u, std::min(h, 16u));
}

// Run f(i) for i in [0, n) on the worker pool (contiguous ranges per thread).
void parallelFor(size_t n, const std::function<void(size_t, size_t)> &f) {
	int t = workerCount();
	if (n < 4096 || t == 1) {
		f(0, n);
		return;
	}
	std::vector<std::function<void()>> tasks;
	size_t per = (n + t - 1) / t;
	for (int k = 0; k < t; k++) {
		size_t a = k * per, b = std::min(n, a + per);
		if (a >= b) break;
		tasks.emplace_back([&f, a, b] { f(a, b); });
	}
	runAll(tasks);
}

struct Out {
	std::FILE *f;
	std::string path;

	void put(const char *p, size_t n) {
		if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("write failed: " + path);
	}
};

// Lines [0, n) formatted by line(i, buf) -> chars written (at most 160), in blocks of lines
// formatted concurrently and written in order.
void writeLines(Out &o, size_t n, const std::function<int(size_t, char *)> &line) {
	constexpr size_t kBlock = 1 << 15;
	const int t = workerCount();
	std::vector<std::string> blk(t);
	for (size_t base = 0; base < n; base += kBlock * t) {
		auto fmt = [&](int k) {
			size_t a = base + k * kBlock, b = std::min(n, a + kBlock);
			std::string &s = blk[k];
			s.clear();
			if (a >= b) return;
			s.resize((b - a) * 160);
			char *p = s.data();
			for (size_t i = a; i < b; i++) p += line(i, p);
			s.resize(p - s.data());
		};
		int used = (int) std::min<size_t>(t, (n - base + kBlock - 1) / kBlock);
		std::vector<std::function<void()>> tasks;
		for (int k = 0; k < used; k++) tasks.emplace_back([&fmt, k] { fmt(k); });
		runAll(tasks);
		for (int k = 0; k < used; k++) o.put(blk[k].data(), blk[k].size());
	}
}

void writeValues(Out &o, const double *v, size_t n, size_t stride) {
	writeLines(o, n, [&](size_t i, char *p) { return std::snprintf(p, 160, "%.12e\n", v[i * stride]); });
}

// SolutionPost::parseFile semantics (SolutionPost.cpp:40-59): sol[k][i], k = rho u v w p E.
std::vector<double> readSol(const std::string &path, long nElem) {
	std::ifstream in(path, std::ios::binary);
	if (!in) throw std::runtime_error("Error while opening solution file ! " + path);
	std::string buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
	// line starts after the header line; lines past the end of the file read as empty
	std::vector<size_t> start;
	start.reserve(nElem + 1);
	size_t pos = buf.find('\n');
	pos = pos == std::string::npos ? buf.size() : pos + 1;
	for (long i = 0; i < nElem && pos < buf.size(); i++) {
		start.push_back(pos);
		size_t nl = buf.find('\n', pos);
		pos = nl == std::string::npos ? buf.size() : nl + 1;
	}
	start.push_back(pos);
	const long nLines = (long) start.size() - 1;
	std::vector<double> sol(6 * (size_t) nElem, 0.0);
	std::vector<unsigned char> got(nElem, 0); // bit k: number k was found on the line
	parallelFor((size_t) nElem, [&](size_t a, size_t b) {
		std::string line;
		for (size_t i = a; i < b; i++) {
			if ((long) i >= nLines) continue;
			line.assign(buf, start[i], start[i + 1] - start[i]);
			const char *q = line.c_str();
			unsigned char mask = 0;
			for (int k = 0; k < 6; k++) {
				while (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n' || *q == '\v' || *q == '\f') ++q;
				if (!*q) break; // end of line: this and the later numbers keep their values
				// iostream extraction takes decimal digits, a point and an exponent only: "inf",
				// "nan" and hex forms do not read as those values
				const char *r = q + (*q == '+' || *q == '-');
				bool decimal = (*r >= '0' && *r <= '9') || *r == '.';
				char *end = const_cast<char *>(q);
				double v = 0.0;
				if (decimal && r[0] == '0' && (r[1] == 'x' || r[1] == 'X')) {
					end = const_cast<char *>(r + 1); // "0" reads, the next number starts at 'x'
				} else if (decimal) {
					v = std::strtod(q, &end);
				}
				if (end == q) { // not a number: reads as 0, the line ends there
					sol[6 * i + k] = 0.0;
					mask |= (unsigned char) (1u << k);
					break;
				}
				sol[6 * i + k] = v;
				mask |= (unsigned char) (1u << k);
				q = end;
			}
			got[i] = mask;
		}
	});
	// values missing from a line repeat the previous line's (zero before the first line)
	for (long i = 1; i < nElem; i++)
		if (got[i] != 0x3f)
			for (int k = 0; k < 6; k++)
				if (!(got[i] >> k & 1)) sol[6 * i + k] = sol[6 * (i - 1) + k];
	return sol;
}

int elementLine(const MeshFile &m, long e, char *p) {
	const int *n = &m.elemNodes[m.elemStart[e]];
	int cnt = m.elemStart[e + 1] - m.elemStart[e];
	static const int kTet[8] = {0, 1, 2, 2, 3, 3, 3, 3};
	static const int kPyr[8] = {0, 1, 2, 3, 4, 4, 4, 4};
	static const int kPri[8] = {0, 1, 1, 2, 3, 4, 4, 5};
	static const int kHex[8] = {0, 1, 2, 3, 4, 5, 6, 7};
	const int *map = nullptr;
	switch (m.elemVtk[e]) {
		case VTK_HEX: map = kHex; break;
		case VTK_TET: map = kTet; break;
		case VTK_PYRAMID: map = kPyr; break;
		case VTK_WEDGE: map = kPri; break;
		default: break;
	}
	int w = 0;
	if (map) {
		w = std::snprintf(p, 160, "%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\n", n[map[0]] + 1, n[map[1]] + 1, n[map[2]] + 1,
		                  n[map[3]] + 1, n[map[4]] + 1, n[map[5]] + 1, n[map[6]] + 1, n[map[7]] + 1);
	} else {
		for (int k = 0; k < cnt; k++) w += std::snprintf(p + w, 160 - w, "%d\t", n[k] + 1);
		p[w++] = '\n';
	}
	return w;
}

}// namespace

std::vector<std::string> postSolutionPaths(const Params &p) {
	std::vector<std::string> out;
	std::filesystem::path dir(p.outputDir);
	for (auto &par : p.partitionFiles) {
		auto name = std::filesystem::path(par).filename();
		name += ".sol";
		out.push_back((dir / name).string());
	}
	return out;
}

void writeTecplot(const std::string &outPath, const std::vector<std::string> &parPaths,
                  const std::vector<std::string> &solPaths) {
	if (parPaths.size() != solPaths.size()) throw std::invalid_argument("one .sol file per partition is required");
	std::FILE *f = std::fopen(outPath.c_str(), "w");
	if (!f) throw std::runtime_error("cannot open " + outPath + " for writing");
	std::vector<char> iobuf(1 << 22);
	std::setvbuf(f, iobuf.data(), _IOFBF, iobuf.size());
	Out o{f, outPath};
	try {
		static const char kHead[] = "VARIABLES=\"X\",\"Y\",\"Z\",\"Rho\",\"U\",\"V\",\"W\",\"P\",\"E\"\n";
		o.put(kHead, sizeof(kHead) - 1);
		for (size_t ip = 0; ip < parPaths.size(); ip++) {
			MeshFile m = readSu2(parPaths[ip], ParBlock::Ignore);
			const long nNodes = m.nNode(), nElem = m.nElemDeclared;
			char zone[256];
			int zl = std::snprintf(zone, sizeof(zone),
			                       "ZONE T=\"Element\"\nNodes=%ld, Elements=%ld, ZONETYPE=FEBRICK\n"
			                       "DATAPACKING=BLOCK, VARLOCATION=([4-9]=CELLCENTERED)\n",
			                       nNodes, nElem);
			o.put(zone, zl);
			for (int k = 0; k < 3; k++) writeValues(o, m.xyz.data() + k, nNodes, 3);
			std::vector<double> sol = readSol(solPaths[ip], nElem);
			for (int k = 0; k < 6; k++) writeValues(o, sol.data() + k, nElem, 6);
			writeLines(o, m.nElem(), [&](size_t e, char *p) { return elementLine(m, (long) e, p); });
		}
	} catch (...) {
		std::fclose(f);
		throw;
	}
	if (std::fclose(f) != 0) throw std::runtime_error("write failed: " + outPath);
}

}// namespace e3d
