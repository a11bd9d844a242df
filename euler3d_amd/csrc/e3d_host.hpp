// Host-side data model of the explicit Euler solver: mesh as read from SU2/.par, the face/cell
// structure-of-arrays derived from it, the .e3d solver parameters, and the mesh generators.
//
// Reference counterparts (file:line in the Euler3D tree):
//   mesh file model     src/parser/SU2MeshParser.cpp:32-240, src/parser/MeshPartition.cpp:39-112
//   connectivity        src/connectivity/Connectivity.cpp:70-262, include/connectivity/Connectivity.hpp:37-45
//   ghost classification include/mesh/Mesh.hpp:471-707
//   metrics             src/mesh/Metrics.cpp:9-309
//   solver parameters   src/parser/SimConfig.cpp:49-236
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace e3d {

// VTK cell / face type ids used by SU2 files.
enum : int { VTK_TRI = 5, VTK_QUAD = 9, VTK_TET = 10, VTK_HEX = 12, VTK_WEDGE = 13, VTK_PYRAMID = 14 };

// Boundary kind of a ghost cell. The numbering is this library's own.
enum BoundaryKind : int { BND_FARFIELD = 0, BND_WALL = 1, BND_SYM = 2, BND_OTHER = 3, BND_MPI = 4 };

struct Marker {
	std::string tag;
	std::vector<int> vtk;       // per boundary element
	std::vector<int> nodeStart; // CSR into nodes, size n+1
	std::vector<int> nodes;
};

// One INTERNAL_MARKER block of a .par file: the faces this partition shares with partition
// `partner` (MeshPartition.cpp:39-112). Element i is (thisElem[i], otherElem[i]), local ids in the
// two partitions, and the face's VTK id and nodes (local to this partition).
struct InterfaceMarker {
	int partner = 0;
	std::vector<int> thisElem, otherElem, vtk;
	std::vector<int> nodeStart{0}, nodes;
};

// The mesh as a file holds it (SU2 / .par, one partition).
struct MeshFile {
	std::vector<double> xyz;       // 3 * nNode
	std::vector<int> elemVtk;      // nElem
	std::vector<int> elemStart;    // nElem + 1
	std::vector<int> elemNodes;
	std::vector<Marker> markers;
	int nInternalMarker = 0;       // .par NINTERNAL_MARKER
	std::vector<InterfaceMarker> interfaces; // read with ParBlock::Partition
	long nElemDeclared = 0;        // NELEM of the file, unknown VTK ids included
	int nNode() const { return (int) xyz.size() / 3; }
	int nElem() const { return (int) elemVtk.size(); }
	int nBoundaryElems() const {
		int n = 0;
		for (auto &m : markers) n += (int) m.vtk.size();
		return n;
	}
	int nInterfaceElems() const {
		int n = 0;
		for (auto &m : interfaces) n += (int) m.thisElem.size();
		return n;
	}
};

// Derived face/cell structure in the reference's numbering (face ids, ghost ids, orientation).
struct SolverMesh {
	int nElem = 0;     // interior cells
	int nFace = 0;
	int nFaceInt = 0;  // interior faces are ids [0, nFaceInt), boundary faces [nFaceInt, nFace)
	int nGhost = 0;    // ghost k <-> boundary face nFaceInt + k, ghost cell id nElem + k

	std::vector<int> face2element;      // 2 * nFace: (elem0, elem1); elem1 >= nElem is a ghost
	std::vector<int> face2nodeStart;    // nFace + 1
	std::vector<int> face2node;
	std::vector<int> element2faceStart; // nElem + 1
	std::vector<int> element2face;      // local-face order of each cell
	std::vector<int> element2element;

	// ghost lists in the reference's marker order (Mesh.hpp:594-707)
	std::vector<int> wallGhost, wallCell, wallFace;
	std::vector<int> farGhost, farCell, farFace;
	std::vector<int> symGhost, symCell, symFace;
	std::vector<int> ghostKind;         // nGhost, BoundaryKind of ghost nElem + k

	// partition interfaces (Mesh::updateMPIGhostCells, Mesh.hpp:471-584), in the reference's order
	std::vector<int> mpiGhost, mpiCell;           // MPIGhostCellIDs, MPIadjacentToGhostCellIDs
	std::vector<std::vector<int>> interfaceGhost; // per interface marker and element: its ghost
	// cells [partStart[r], partStart[r + 1]) belong to partition r of a composite mesh (empty for
	// a mesh read on its own)
	std::vector<int> partStart;
	// the interface exchange of a composite mesh (MPIHandler::updateFlowField), one message per
	// (sending, receiving) partition pair: entries [haloStart[k], haloStart[k + 1]) of message k
	// carry the state of cell haloSrc[i] into ghost haloDst[i] (composite ids), in the order the
	// reference's sorted requester and sender lists pair them; haloFrom[k] / haloTo[k] are the
	// sending and receiving partitions. A ghost that no exchange reaches gets its own cell
	// (a message from its partition to itself, last).
	std::vector<int> haloStart, haloFrom, haloTo, haloSrc, haloDst;

	// metrics (FP64), non-dimensionalised by CREF
	std::vector<double> faceNormal;     // 3 * nFace, area-weighted, oriented elem0 -> elem1
	std::vector<double> faceUnitNormal; // 3 * nFace
	std::vector<double> faceArea;       // nFace
	std::vector<double> faceCenter;     // 3 * nFace
	std::vector<double> cellVolume;     // nElem
	std::vector<double> cellCentroid;   // 3 * nElem
};

// Solver parameters from the .e3d file (SimConfig), with explicit defaults for the keys the
// reference leaves uninitialised when absent (include/parser/SimConfig.hpp:149-179).
struct Params {
	std::string initialMesh;
	int nPartition = 1;
	std::vector<std::string> partitionFiles;
	int orientCL[2] = {1, 1};
	int orientCD[2] = {0, 1};
	int orientCM[2] = {2, 1};
	double refPoint[3] = {0, 0, 0};
	double sref = 1.0;
	double cref = 1.0;
	int sampling = 1;
	int speedOption = 0;
	double mach = 0.0;
	double velocity = 0.0;
	double aoa = 0.0;
	double pressure = 101325.0;
	double temperature = 288.15;
	double viscosity = 1.853e-5;
	double density = 1.225;
	double gamma = 1.4;
	double gasConstant = 287.058;
	double specificHeat = 1004.7;
	int scheme = 0;          // 0 Roe, 1 AUSM
	int timeIntegration = 2; // 0 explicit Euler, 1 implicit Euler, 2 RK5
	int residualSmoothing = 0;
	double cfl = 1.0;
	double minResidual = 0.0;
	double minAeroCoeffError = 0.0;
	int maxIter = 1;
	std::string solverLog;
	std::string outputFile;
	std::string outputDir;
	std::string preLog, postLog;
};

// ---- mesh_io.cpp
// parBlock: the reader's handling of the .par NINTERNAL_MARKER block. Optional and Required
// accept only NINTERNAL_MARKER= 0 (a mesh solved on its own); Partition requires the block and
// reads its interface markers (one partition of several).
enum class ParBlock { Ignore, Optional, Required, Partition };
MeshFile readSu2(const std::string &path, ParBlock parBlock);
void writeSu2(const MeshFile &m, const std::string &path, int precision, bool parBlock);

// ---- partitions.cpp
// Where partition r of a composite mesh sits in it.
struct PartRange {
	int elem0 = 0, nElem = 0;       // interior cells
	int node0 = 0, nNode = 0;       // nodes (composite node ids = node0 + local id)
	int faceInt0 = 0, nFaceInt = 0; // interior faces
	int ghost0 = 0, nGhost = 0;     // ghosts; boundary face nFaceInt + ghost0 + k of the composite
	int wall0 = 0, nWall = 0;       // entries of the composite wall list
};
// Several partitions of one mesh (each preprocessed on its own, in its reference numbering) as one
// resident problem. Cells, nodes and ghosts are concatenated partition by partition; faces are
// [every partition's interior faces][every partition's boundary faces], so each cell still sums
// its faces in its partition's ascending face order. An interface face keeps its ghost, and its
// elem1 becomes the cell of the other partition whose state the reference's ghost exchange
// delivers to that ghost (MPIHandler.cpp:16-267).
SolverMesh mergePartitions(const std::vector<MeshFile> &files, const std::vector<SolverMesh> &parts,
                           std::vector<PartRange> &ranges, MeshFile &merged);
// Element-to-partition assignment by recursive coordinate bisection of the cell centroids.
std::vector<int> partitionRcb(const SolverMesh &sm, int nPart);
// The partition files E3D_PRE writes for an assignment (Partition.cpp:123-460).
std::vector<MeshFile> splitMesh(const MeshFile &global, const SolverMesh &sm, const std::vector<int> &elem2part,
                                int nPart);

// ---- config.cpp
Params readConfig(const std::string &path);

// ---- post.cpp
// Tecplot ASCII file of the given partitions and their .sol files (src/post/post.cpp:49-133).
void writeTecplot(const std::string &outPath, const std::vector<std::string> &parPaths,
                  const std::vector<std::string> &solPaths);
// the .sol files Post::Post reads: <dir of OUTPUT_FILE>/<partition file name>.sol (post.cpp:18-30)
std::vector<std::string> postSolutionPaths(const Params &p);

// ---- preprocess.cpp
// Builds SolverMesh from a MeshFile. O(N log N) in the number of cells; numbering, orientation
// and every metric follow the reference exactly.
SolverMesh preprocess(const MeshFile &m, double cref);

// local faces of a cell type (Connectivity.hpp:37-45): count, and the cell-local node indices of
// face f written to out (returns the node count)
int localFaceCount(int vtk);
int localFaceNodes(int vtk, int f, int out[4]);

// ---- generators.cpp
struct GenInfo {
	double sref = 1.0, cref = 1.0;
};
MeshFile genBox(int nx, int ny, int nz, double lx, double ly, double lz, const char *tagsXYZ6[6]);
MeshFile genNacaOgrid(int ni, int nj, int nk, double thicknessFactor, double camber, double camberPos,
                      double wallSpacing, double farfieldRadius, double dz, uint64_t seed_unused);
MeshFile genBumpTets(int nx, int ny, int nz, double jitter, uint64_t seed, bool permute);
MeshFile genWing(int ni, int nj, int nk, double thickness, double sweepDeg, double taper, double semispan,
                 double farfield);

// splitmix64 (Steele, Lea & Flood 2014) — deterministic across compilers, unlike std distributions.
struct SplitMix64 {
	uint64_t s;
	explicit SplitMix64(uint64_t seed) : s(seed) {}
	uint64_t next() {
		uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
		return z ^ (z >> 31);
	}
	double uniform() { return (double) (next() >> 11) * 0x1.0p-53; }
};

}// namespace e3d
