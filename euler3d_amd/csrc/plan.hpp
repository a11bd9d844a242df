// Device data plan of one case: a locality order of the cells, the block structure of the fused
// stage kernel, and every per-face / per-cell array in the device numbering.
#pragma once

#include "e3d_host.hpp"

#include <cstdint>
#include <vector>

namespace e3d {

enum CellOrder : int { ORDER_REFERENCE = 0, ORDER_MORTON = 1, ORDER_BISECT = 2 };

struct HostPlan {
	int nElem = 0, nFace = 0, nFaceInt = 0, nGhost = 0;
	std::vector<int> new2old, old2new; // interior cells: device id <-> reference id

	// per-face arrays in reference face order, cell ids in device numbering
	std::vector<int> fe0, fe1;
	std::vector<double> fnx, fny, fnz, fS;
	std::vector<int8_t> gkind; // per ghost: 0 farfield, 1 wall, 2 symmetry, 3 partition interface

	// per-cell arrays in device order
	std::vector<double> vol, sx, sy, sz;
	std::vector<int> fstart, flist; // reference face ids ascending, bit 31 = subtract

	// fused-stage block plan
	int cellsPerBlock = 0, nBlocks = 0, maxSlots = 0, csrStride = 0;
	std::vector<int> blockSlot;
	std::vector<int> se0, se1;
	std::vector<int8_t> skind;
	// block-local view of the slots: bits 0-10 local index of e0, 11-21 local index of e1, 22-24
	// kind (0 interior, 1 wall, 2 symmetry, 3 farfield, 4 partition interface read from its ghost). Local index: own cell i -> i - c0; halo cell -> cellsPerBlock + its rank in the
	// block's halo list; e1 of a boundary slot is a ghost and has no local index (0x7FF).
	std::vector<uint32_t> spack;
	std::vector<int> blockHalo;  // nBlocks + 1
	std::vector<int> haloCell;   // per block: the interior cells outside it that its faces read, ascending
	int maxHalo = 0;
	std::vector<double> snx, sny, snz, sS;
	std::vector<uint16_t> csr;

	// interface exchange of a composite mesh when its faces read ghost slots (buildPlan exchange =
	// true): entry i carries interior cell xSrc[i] (device numbering) into ghost xDst[i]; its values
	// sit in the message buffer at xSlot[2 i] + k * xSlot[2 i + 1] for k = rho, u, v, w, E
	std::vector<int> xSrc, xDst, xSlot;
	int xBufLen = 0;

	// wall list (reference order) for the force integration
	std::vector<int> wallCell;
	std::vector<double> wallN[3], wallC[3];
};

// exchange: the faces of partition interfaces read the ghost that the interface exchange fills
// (slot kind 4) instead of the other partition's cell, and the plan carries that exchange
HostPlan buildPlan(const SolverMesh &sm, int order, int cellsPerBlock, bool exchange = false);
// the device order of the interior cells (new2old) that buildPlan uses; the cells of each
// partition of a composite mesh keep their range [partStart[r], partStart[r + 1])
std::vector<int> cellOrder(const SolverMesh &sm, int order, int cellsPerBlock);

}// namespace e3d
