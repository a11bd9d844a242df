// SU2 / .par mesh reader and writer.
//
// Reader semantics follow src/parser/SU2MeshParser.cpp:32-240 (NDIME must be 3; NELEM, NPOIN and
// NMARK blocks, each keyword searched forward from the previous block) and
// src/parser/MeshPartition.cpp:39-112 (the .par NINTERNAL_MARKER block, which is required for a
// partition file, and its interface markers when several partitions are read together).
//
// Writer layout follows src/partition/SU2Writer.cpp:29-123 (headers, setw(3) vtk ids,
// setw(10) node ids and element index, setw(20) scientific coordinates at the requested
// precision, MARKER_TAG / MARKER_ELEMS, NINTERNAL_MARKER).
//
// The parser works on the whole file in memory with strtol/strtod, so multi-million-cell meshes
// load in seconds instead of going through one stringstream per line.
#include "e3d_host.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace e3d {

namespace {

int nodesOfVolume(int vtk) {
	switch (vtk) {
		case VTK_TET: return 4;
		case VTK_HEX: return 8;
		case 11: return 8; // voxel: parsed, rejected by connectivity (Connectivity.cpp:141-146)
		case VTK_WEDGE: return 6;
		case VTK_PYRAMID: return 5;
		default: return -1;
	}
}

int nodesOfSurface(int vtk) {
	switch (vtk) {
		case VTK_TRI: return 3;
		case VTK_QUAD: return 4;
		case 8: return 4; // pixel
		default: return -1;
	}
}

struct Cursor {
	const char *p;
	const char *end;

	bool eof() const { return p >= end; }
	// next line as [b, e), advancing past the newline
	bool line(const char *&b, const char *&e) {
		if (p >= end) return false;
		b = p;
		const char *nl = (const char *) std::memchr(p, '\n', end - p);
		e = nl ? nl : end;
		p = nl ? nl + 1 : end;
		return true;
	}
	// find a line containing key; returns pointer just after the line start
	bool findLine(const char *key, const char *&b, const char *&e) {
		size_t kl = std::strlen(key);
		while (line(b, e)) {
			if ((size_t) (e - b) >= kl) {
				for (const char *q = b; q + kl <= e; ++q)
					if (std::memcmp(q, key, kl) == 0) return true;
			}
		}
		return false;
	}
};

long parseLongAfter(const char *b, const char *e, size_t off) {
	std::string s(b, e);
	if (s.size() < off) return 0;
	return std::strtol(s.c_str() + off, nullptr, 10);
}

std::string parseWordAfter(const char *b, const char *e, size_t off) {
	std::string s(b, e);
	if (s.size() < off) return "";
	std::istringstream ss(s.substr(off));
	std::string w;
	ss >> w;
	return w;
}

// read integers from a line into out (up to max)
int lineInts(const char *b, const char *e, long *out, int max) {
	int n = 0;
	const char *q = b;
	while (n < max) {
		while (q < e && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
		if (q >= e) break;
		char *endp;
		long v = std::strtol(q, &endp, 10);
		if (endp == q) break;
		out[n++] = v;
		q = endp;
	}
	return n;
}

}// namespace

MeshFile readSu2(const std::string &path, ParBlock parBlock) {
	std::ifstream in(path, std::ios::binary);
	if (!in) throw std::runtime_error("Error while opening mesh file ! " + path);
	std::string buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
	Cursor c{buf.data(), buf.data() + buf.size()};
	MeshFile m;
	const char *b, *e;

	if (!c.findLine("NDIME=", b, e)) throw std::runtime_error("Dimension keyword not found \"NDIME=\" !");
	if (parseLongAfter(b, e, 6) != 3) throw std::runtime_error("Geometry is not 3D (NDIME != 3 )");

	if (!c.findLine("NELEM=", b, e)) throw std::runtime_error("Dimension keyword not found \"NELEM=\" !");
	long nElem = parseLongAfter(b, e, 6);
	m.nElemDeclared = nElem;
	m.elemVtk.reserve(nElem);
	m.elemStart.reserve(nElem + 1);
	m.elemStart.push_back(0);
	m.elemNodes.reserve(nElem * 8);
	long vals[16];
	for (long i = 0; i < nElem; i++) {
		if (!c.line(b, e)) throw std::runtime_error("unexpected end of file in NELEM block");
		int n = lineInts(b, e, vals, 16);
		if (n < 1) throw std::runtime_error("malformed element line");
		int nn = nodesOfVolume((int) vals[0]);
		if (nn < 0) continue; // unknown ids are skipped, as SU2MeshParser.cpp:89-109 does
		if (n < 1 + nn) throw std::runtime_error("element line has too few nodes");
		m.elemVtk.push_back((int) vals[0]);
		for (int k = 0; k < nn; k++) m.elemNodes.push_back((int) vals[1 + k]);
		m.elemStart.push_back((int) m.elemNodes.size());
	}

	if (!c.findLine("NPOIN=", b, e)) throw std::runtime_error("Dimension keyword not found \"NPOIN=\" !");
	long nPoin = parseLongAfter(b, e, 6);
	m.xyz.resize(3 * nPoin);
	for (long i = 0; i < nPoin; i++) {
		if (!c.line(b, e)) throw std::runtime_error("unexpected end of file in NPOIN block");
		std::string s(b, e);
		const char *q = s.c_str();
		char *endp;
		for (int k = 0; k < 3; k++) {
			m.xyz[3 * i + k] = std::strtod(q, &endp);
			if (endp == q) throw std::runtime_error("malformed coordinate line");
			q = endp;
		}
	}

	if (!c.findLine("NMARK=", b, e)) throw std::runtime_error("Dimension keyword not found \"NMARK=\" !");
	long nMark = parseLongAfter(b, e, 6);
	for (long im = 0; im < nMark; im++) {
		Marker mk;
		if (!c.line(b, e)) throw std::runtime_error("missing MARKER_TAG");
		mk.tag = parseWordAfter(b, e, 11);
		if (!c.line(b, e)) throw std::runtime_error("missing MARKER_ELEMS");
		long ne = parseLongAfter(b, e, 13);
		mk.nodeStart.push_back(0);
		for (long i = 0; i < ne; i++) {
			if (!c.line(b, e)) throw std::runtime_error("unexpected end of file in marker block");
			int n = lineInts(b, e, vals, 16);
			int nn = n > 0 ? nodesOfSurface((int) vals[0]) : -1;
			if (nn < 0) continue;
			if (n < 1 + nn) throw std::runtime_error("marker element line has too few nodes");
			mk.vtk.push_back((int) vals[0]);
			for (int k = 0; k < nn; k++) mk.nodes.push_back((int) vals[1 + k]);
			mk.nodeStart.push_back((int) mk.nodes.size());
		}
		m.markers.push_back(std::move(mk));
	}

	if (parBlock == ParBlock::Ignore) return m; // SU2MeshParser reads no further
	Cursor c2 = c;
	if (c2.findLine("NINTERNAL_MARKER=", b, e)) {
		m.nInternalMarker = (int) parseLongAfter(b, e, 17);
		if (parBlock == ParBlock::Partition) {
			// MeshPartition::parseMpiBoundaryElement (MeshPartition.cpp:39-112): INTERNAL_MARKER=,
			// NELEM_MPI=, then "this other vtk nodes..." per element; a face VTK id it does not
			// know carries no nodes
			for (int im = 0; im < m.nInternalMarker; im++) {
				InterfaceMarker mk;
				if (!c2.line(b, e)) throw std::runtime_error("missing INTERNAL_MARKER");
				mk.partner = (int) parseLongAfter(b, e, 16);
				if (!c2.line(b, e)) throw std::runtime_error("missing NELEM_MPI");
				long ne = parseLongAfter(b, e, 10);
				for (long i = 0; i < ne; i++) {
					if (!c2.line(b, e)) throw std::runtime_error("unexpected end of file in NELEM_MPI block");
					int n = lineInts(b, e, vals, 16);
					if (n < 3) throw std::runtime_error("malformed interface element line");
					int nn = nodesOfSurface((int) vals[2]);
					if (nn < 0) nn = 0;
					if (n < 3 + nn) throw std::runtime_error("interface element line has too few nodes");
					mk.thisElem.push_back((int) vals[0]);
					mk.otherElem.push_back((int) vals[1]);
					mk.vtk.push_back((int) vals[2]);
					for (int k = 0; k < nn; k++) mk.nodes.push_back((int) vals[3 + k]);
					mk.nodeStart.push_back((int) mk.nodes.size());
				}
				m.interfaces.push_back(std::move(mk));
			}
		} else if (m.nInternalMarker != 0) {
			throw std::runtime_error("NINTERNAL_MARKER > 0: this partition has neighbours; read all partitions together");
		}
	} else if (parBlock == ParBlock::Required || parBlock == ParBlock::Partition) {
		throw std::runtime_error("MPI Boundary Elements keyword \"NINTERNAL_MARKER=\" not found !");
	}
	return m;
}

void writeSu2(const MeshFile &m, const std::string &path, int precision, bool parBlock) {
	FILE *f = std::fopen(path.c_str(), "w");
	if (!f) throw std::runtime_error("cannot open " + path + " for writing");
	std::vector<char> iobuf(1 << 20);
	std::setvbuf(f, iobuf.data(), _IOFBF, iobuf.size());
	std::fprintf(f, "%%\n%% Problem dimension\n%%\nNDIME= 3\n");
	std::fprintf(f, "%%\n%% Inner element connectivity\n%%\n");
	std::fprintf(f, "NELEM= %d\n", m.nElem());
	for (int i = 0; i < m.nElem(); i++) {
		std::fprintf(f, "%3d", m.elemVtk[i]);
		for (int k = m.elemStart[i]; k < m.elemStart[i + 1]; k++) std::fprintf(f, "%10d", m.elemNodes[k]);
		std::fprintf(f, "%10d\n", i);
	}
	std::fprintf(f, "%%\n%% Node coordinates\n%%\n");
	std::fprintf(f, "NPOIN= %d\n", m.nNode());
	// setw(20) as SU2Writer uses; wider when the precision would make neighbouring fields touch
	// ("-d.<precision>e+dd" needs precision + 7 characters plus one blank)
	const int width = precision + 8 > 20 ? precision + 8 : 20;
	for (int i = 0; i < m.nNode(); i++)
		std::fprintf(f, "%*.*e%*.*e%*.*e%10d\n", width, precision, m.xyz[3 * i], width, precision, m.xyz[3 * i + 1], width,
		             precision, m.xyz[3 * i + 2], i);
	std::fprintf(f, "%%\n%% Boundary elements\n%%\n");
	std::fprintf(f, "NMARK= %d\n", (int) m.markers.size());
	for (auto &mk : m.markers) {
		std::fprintf(f, "MARKER_TAG= %s\n", mk.tag.c_str());
		std::fprintf(f, "MARKER_ELEMS= %d\n", (int) mk.vtk.size());
		for (size_t i = 0; i < mk.vtk.size(); i++) {
			std::fprintf(f, "%3d", mk.vtk[i]);
			for (int k = mk.nodeStart[i]; k < mk.nodeStart[i + 1]; k++) std::fprintf(f, "%10d", mk.nodes[k]);
			std::fprintf(f, "\n");
		}
	}
	if (parBlock) {
		// SU2Writer::WriteInternalMarker (SU2Writer.cpp:98-121): one block per neighbour partition,
		// each element as setw(10) fields this, other, vtk, face nodes
		std::fprintf(f, "%%\n%% Internal Boundary elements\n%%\n");
		std::fprintf(f, "NINTERNAL_MARKER= %d\n", (int) m.interfaces.size());
		for (auto &mk : m.interfaces) {
			std::fprintf(f, "INTERNAL_MARKER= %d\n", mk.partner);
			std::fprintf(f, "NELEM_MPI= %d\n", (int) mk.thisElem.size());
			for (size_t i = 0; i < mk.thisElem.size(); i++) {
				std::fprintf(f, "%10d%10d%10d", mk.thisElem[i], mk.otherElem[i], mk.vtk[i]);
				for (int k = mk.nodeStart[i]; k < mk.nodeStart[i + 1]; k++) std::fprintf(f, "%10d", mk.nodes[k]);
				std::fprintf(f, "\n");
			}
		}
	}
	std::fclose(f);
}

}// namespace e3d
