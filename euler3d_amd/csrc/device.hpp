// Device-resident solver data for one case and the launchers of the HIP kernels that advance it.
//
// Cells are stored in a locality order (a space-filling-curve sort of the centroids) so that a
// block of consecutive cells is a compact cluster; ghost cells keep the reference's numbering
// after the interior ones. The reference numbering survives where it decides bits: every face
// keeps its reference orientation (elem0 -> elem1) and every cell sums its faces in ascending
// reference face id.
//
// Layout in HBM (FP64 unless noted; structure-of-arrays, so 64 lanes load 64 consecutive values):
//   faces  [nFace]   e0, e1 (int32, device cell ids; e1 >= nElem is a ghost), unit normal nx, ny,
//                    nz, area S (reference face order; used by the per-face parity entry points)
//   ghosts [nGhost]  kind (int8: 0 farfield, 1 wall, 2 symmetry); ghost k sits on face nFaceInt + k
//   cells  [nElem]   volume, sum|Nx|, sum|Ny|, sum|Nz| (static parts of the local time step)
//   state  2 x [nElem + nGhost][8] (rho, u, v, w, p, E, H, M) per cell, 64 B: ping-pong buffers, a
//                    stage reads one and writes the other (the ghost part carries the farfield
//                    ghosts' previous state)
//   W0     [5][nElem] conservative state at the start of the iteration
//   plan   per block of kCellsPerBlock cells: its faces (every face touching the block, faces cut
//                    by the block boundary appear in both blocks) as slot records e0, e1, kind,
//                    n, S; per cell the block-local slots of its faces in ascending reference
//                    face id, sign in bit 15
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace e3d {

constexpr int kCellsPerBlock = 256;

struct DevFaces {
	const int *e0, *e1;
	const double *nx, *ny, *nz, *S;
};

struct DevCells {
	const double *vol, *sx, *sy, *sz;
	const int *fstart; // nElem + 1
	const int *flist;  // face id, bit 31 set when the cell is the face's elem1 (subtract)
};

// cell state, array of structures: kStateWidth doubles per cell in the order of StateField
constexpr int kStateWidth = 8;
enum StateField : int { S_RHO = 0, S_U, S_V, S_W, S_P, S_E, S_H, S_M };
struct DevState {
	double *a; // (nElem + nGhost) * kStateWidth, 64 B aligned per cell
};

struct DevFree {
	double gamma, cfl;
	double rho_inf, p_inf, u_inf, v_inf, w_inf, M_inf, E_inf, H_inf;
};

struct DevMesh {
	int nElem, nFace, nFaceInt, nGhost;
	DevFaces f;
	DevCells c;
	const int8_t *gkind; // nGhost: 0 farfield, 1 wall, 2 symmetry, 3 partition interface
};

// block plan of the fused stage kernel
struct DevPlan {
	int nBlocks;
	int maxSlots;            // largest slot count of a block (LDS size)
	int csrStride;           // slots per cell in the CSR (4 for tets, 6 with hexes), 0xFFFF = none
	const int *blockSlot;    // nBlocks + 1: slot range of each block
	const int *e0, *e1;      // per slot: device cell ids (e1 >= nElem: ghost)
	const int8_t *kind;      // per slot: 0 interior, 1 wall, 2 symmetry, 3 farfield
	const double *nx, *ny, *nz, *S;
	const uint16_t *csr;     // nElem * csrStride
	const uint32_t *pack;    // per slot: local e0 | local e1 << 11 | kind << 22
	const int *blockHalo;    // nBlocks + 1
	const int *haloCell;     // outside cells read by each block, device ids
	int maxHalo;
	int cellsPerBlock;       // 256, or 128 where a 256-cell block needs too much LDS
	// per block: first slot, slot count | leading interior (kind 0) slots << 16, first halo entry,
	// halo count
	const int4 *blockMeta;
};

void launchBc(const DevMesh &m, DevState s, const DevFree &q, int machDerived, hipStream_t st);
// per-face flux output: F*S into fs[5][nFace]; when raw != nullptr also the unscaled F
void launchFlux(const DevMesh &m, DevState s, const DevFree &q, int scheme, double *fs, double *raw, hipStream_t st);
// gather-only: R[5][nElem] from fs (parity hook)
void launchGather(const DevMesh &m, const double *fs, double *R, hipStream_t st);
void launchTimestep(const DevMesh &m, DevState s, const DevFree &q, double *dt, hipStream_t st);
// Unfused stage (needs launchBc + launchFlux before it): residual gathered from fs (or read from
// Rin, SoA 5 x nElem), local dt and the update in place. integ 0: RK, W_k = W0 - alpha*dt*R/V
// (firstStage also stores W0); integ 1: explicit Euler, W0 is the persistent conservative state;
// integ 2: no update. When rmsPartial != nullptr each block writes its partial sum of R_rho^2.
void launchStage(const DevMesh &m, DevState s, const DevFree &q, const double *fs, const double *Rin, double *W0,
                 double alpha, int firstStage, int integ, double *rmsPartial, double *Rout, hipStream_t st);
// residual smoothing R0 -> R2 (R1 scratch, blockMax one double per 256 cells, flag[2] device ints:
// flag[1] is sticky NaN and must start at 0)
// flag: [nPart] decisions then the NaN flag; with nPart > 1, partStart (nPart + 1, device order)
// gives each partition's cells and cellMax (nElem) holds the per-cell changes
void launchSmooth(const double *R0, double *R1, double *R2, const int *nbr, const uint8_t *nface, int stride, int nElem,
                  double *blockMax, int *flag, const int *partStart, int nPart, double *cellMax, hipStream_t st);
int stageBlocks(int nElem);
// dynamic LDS of one fused-stage workgroup for a plan's largest block, per state layout
size_t stageFusedLdsBytes(int cellsPerBlock, int maxSlots, int maxHalo, int layout);
// waves the fused RK stage keeps resident per CU for such a plan (0: a block does not fit)
int stageFusedWavesPerCU(int cellsPerBlock, int maxSlots, int maxHalo, int scheme);
// workgroups of the persistent, software-pipelined stage for this plan (every one resident at
// once), or 0 when the plan runs the one-block-per-workgroup kernel
int stagePipeGrid(const DevPlan &p, int scheme, int device, int xcds);
// launch choices of the fused stage (they change speed only, never results)
struct StageTuning {
	bool nontemporal = true; // non-temporal loads for arrays read once per stage
	int xcds = 8;            // XCD count for the L2-aware block order (1: plain dispatch order)
	// quotients sharing a denominator use one IEEE reciprocal plus Markstein's FMA correction,
	// which is correctly rounded and so bit-identical to the divisions it replaces
	bool sharedRecip = true;
	// every other stage walks each XCD's block range backwards, so that it starts on the cells
	// the previous stage wrote last, which are still in the memory-side cache
	bool alternate = true;
	// > 0: the persistent, software-pipelined stage with this many workgroups (stagePipeGrid)
	int pipeGrid = 0;
	// threads of a 256-cell block whose LDS fits one workgroup per CU, when the pipelined stage
	// does not run it (768: 12 waves; 149 against 161 us per stage on the 1.18M-hex wing)
	int wideThreads = 768;
	// 256-cell blocks at the fixed LDS strides: 0 the update's inputs loaded with the block's state,
	// 256 threads; 1 loaded after the flux loop, 320 threads at 128 registers (15 waves per CU);
	// 2 loaded after the flux loop, 256 threads; -1 by scheme: AUSM 1 (6.3M-tet bump 345 -> 324 us
	// per stage), Roe 2 (380 -> 377 us; 320 threads took 454 us)
	int fixedShape = -1;
};

// Fused RK stage: boundary ghosts, face fluxes (once per block, in LDS), residual gather, local
// dt and update, reading `in` and writing `out` (interior and ghosts). Before the first update
// after a state upload (machDerived == 0) the interior Mach number comes from the state's M slot.
// integ 0: RK stage (W0 = the iteration's start state, written in the first stage); integ 1:
// explicit Euler step (W0 = the persistent conservative state rho, rho u, rho v, rho w, rho E).
// parity: the stage's ping-pong parity (the walk direction when t.alternate is set).
// layout (RK stages only; the explicit Euler step is always 0): where the interior state lives.
//   0: 64 B primitive records in `in` -> `out`;
//   1: primitive records in `in` -> conservative states (5 doubles per cell) in Wout;
//   2: conservative states Win -> Wout.
// Ghost records are read from `in` and written to `out` in every layout. The conservative buffers
// hold 5 * nElem doubles plus two of padding.
void launchStageFused(const DevMesh &m, const DevPlan &p, DevState in, DevState out, const double *Win, double *Wout,
                      int layout, const DevFree &q, int scheme, int integ, double *W0, double alpha, int firstStage,
                      int machDerived, double *rmsPartial, int parity, const StageTuning &t, hipStream_t st);
// bitwise comparison of the short square root, reciprocal and quotient with the library's on n
// pseudo-random inputs; bad[0..2] (device, zeroed by the caller) count the mismatches
void launchMathCheck(unsigned long long seed, long long n, unsigned long long *bad, hipStream_t st);
// the interface exchange of a partitioned mesh on one device: pack (rho, u, v, w, E) of the n send
// cells src[] into the message buffer (from W when W is not null, else from the primitive
// records of s), then unpack into the ghost records dst[] of s with p, H and M re-formed; slot
// holds 2 ints per entry (buffer offset, message length)
void launchHaloExchange(int n, const int *src, const int *dst, const int *slot, DevState s, const double *W,
                        double gamma, double *buf, hipStream_t st);
// dst[i] = src[i] for `chunks` 16 B chunks (grid-stride; both 16 B aligned)
void launchCopy16(const void *src, void *dst, long long chunks, hipStream_t st);
// interior primitive records of s from the conservative states W (the stage's expressions), with
// M = c / |u|
void launchWToPrim(const DevMesh &m, const double *W, DevState s, const DevFree &q, hipStream_t st);
// Static per-device scratch of the allocation-free iteration entry point: kIterScratchMax doubles
// (512 KiB, one RMS sum per iteration). iterScratchInit resolves its address on `device`, which
// must be current, once per device; iterScratch returns that address (null before the init).
constexpr int kIterScratchMax = 65536;
hipError_t iterScratchInit(int device);
double *iterScratch(int device);
// deterministic sum of n doubles into out[0] (fixed order, over several workgroups); tmp holds
// sumScratch(n) doubles
int sumScratch(int n);
void launchSum(const double *in, int n, double *tmp, double *out, hipStream_t st);
// interior Mach number from the primitives (FlowField.cpp:157), for state read-back
void launchMach(const DevMesh &m, DevState s, const DevFree &q, hipStream_t st);
// wall-face force/moment: out[6] = (Fx,Fy,Fz, CPx,CPy,CPz) summed in a fixed order over several
// workgroups (wall cells' p from W when W is not null, else from the primitive records of s);
// cpOut (may be null) receives each wall face's cp; tmp holds aeroScratch(nWall) doubles
int aeroScratch(int nWall);
void launchAero(int nWall, const int *wallCell, const double *Nx, const double *Ny, const double *Nz,
                const double *Cx, const double *Cy, const double *Cz, DevState s, const double *W, double gamma,
                double cpFactor, double pInf, double cref, double sref, double *tmp, double *out6, double *cpOut,
                hipStream_t st);

}// namespace e3d
