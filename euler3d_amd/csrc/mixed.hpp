// Mixed-precision multi-case solver: device-side case table, block table and the launchers of the
// kernels that advance every resident case together.
//
// Precision split. Face fluxes, boundary ghosts and the local time step run in FP32 on FP32
// copies of the state and of the face geometry; the conservative state W, the RK update, the
// per-cell residual accumulation (Kahan-compensated), the residual norms and the force
// integration stay FP64.
//
// Perturbation form. The FP32 copy of a cell holds its primitives minus the freestream
// (d_rho, d_u, d_v, d_w, d_p, d_H) and each face contributes F - F_inf(n) instead of F, with
// F_inf(n) the freestream flux through the face. The subtracted terms sum to F_inf . sum(n S) = 0
// over a closed cell, so the residual is unchanged in exact arithmetic, while the FP32 rounding
// error scales with the local departure from freestream instead of with |F|. The residual still
// flattens where the FP32 rounding of those perturbations dominates, typically 5.5 to 6.5 orders
// below the first residual; the stall test below retires a case there so that the FP64 path can
// finish it.
//
// State. The FP64 conservative state W is read and written once per iteration, by the last RK
// stage. The stages before it only need their state for the FP32 records that the fluxes read,
// so they form it from an FP32 copy of the iteration's start state, dW0 = W - W_inf (the
// perturbation form again), instead of reading W.
//
// Cases. Every resident case owns its arrays (plan, metrics, W, dW0, two FP32 record buffers,
// per-block partial sums, residual history). A device table of MixCase descriptors and a table
// of (case, block) pairs let one launch walk the blocks of all cases; a case's
// arithmetic does not depend on which other cases are resident or where its blocks sit in the
// table, so its result is the same alone and in any mix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace e3d {

constexpr int kMixBlock = 256;  // cells per block = threads per workgroup
constexpr int kRecWidth = 8;    // floats per FP32 cell record

// FP32 cell record (32 B): perturbations from freestream and the stored Mach number
enum RecField : int { R_DRHO = 0, R_DU, R_DV, R_DW, R_DP, R_DH, R_M, R_PAD };

struct MixCase {
	// sizes
	int nElem, nGhost, nBlocks, csrStride;
	int scheme; // 0 Roe, 1 AUSM
	int histCap;
	int maxIter;
	int stallWindow;    // iterations per window of the stall test, 0: off
	double drop;        // converged when rms < drop * rms of the first iteration
	double minResidual; // or when rms < minResidual (absolute)
	// stalled when a window's lowest rms is not below stallRatio x the previous window's lowest and
	// below stallLevel x the first rms: the residual has reached the floor of the FP32 fluxes
	double stallRatio, stallLevel;
	// block plan: each block's slots list its interior faces first (blockInt of them), then its
	// boundary faces
	const int *blockSlot, *blockInt, *blockHalo, *haloCell, *ghostOf;
	const uint32_t *pack;
	const float *nx, *ny, *nz, *S;
	const uint16_t *csr;
	// cell metrics (FP32): the sums of |N_x|, |N_y|, |N_z| over the faces (the volume cancels
	// from the update, W -= alpha CFL R / lambda)
	const float *sx, *sy, *sz;
	// closure defect sum(+-n S) of each cell (3 floats), null when every cell is closed: the
	// perturbation form adds F_inf . defect back so that unclosed cells keep the reference's residual
	const float *defect;
	// state
	double *W;      // 5 x nElem conservative state (rho, rho u, rho v, rho w, rho E), FP64
	float *dW0;     // 5 x nElem FP32 W - W_inf of the iteration's start state (stages 1 to 4)
	float *rec[2];  // (nElem + nGhost) x kRecWidth FP32 records, ping-pong
	double *partial; // nBlocks partial sums of R_rho^2 (last stage)
	double *hist;    // residual RMS per iteration, histCap entries
	// freestream (FP32 for the flux, FP64 for the update)
	float gamma, cfl, rInf, uInf, vInf, wInf, pInf, hInf, mInf;
	float pad1;
	double gammaD, rInfD, uInfD, vInfD, wInfD, pInfD, hInfD;
	double consInf[5]; // W_inf (rho, rho u, rho v, rho w, rho E) as W is initialised
};

// device-written progress of a case
struct MixStatus {
	int parity; // FP32 record buffer holding the state at the start of an iteration
	int iter;   // iterations done
	int done;   // nonzero: finished, its blocks are skipped. 1 converged, 2 iteration cap, 3 non-finite
	            // residual, 4 stalled (E3D_CASE_CONVERGED, _CAPPED, _FAILED, _STALLED)
	int active; // 0: the slot holds no case
	double rms0, rmsLast;
	int winStart, pad; // first iteration of the current stall-test window
	double winMin, prevMin; // lowest rms of the current and of the previous window (0: none yet)
};

// launch choices of the mixed stage (they change speed only, never results)
struct MixTuning {
	int wgPerCU = 0;        // 0: one workgroup per (case, block) pair; else persistent workgroups per CU
	int xcds = 8;           // XCD count for the L2-aware block order
	bool nontemporal = true;
};

struct MixLaunch {
	const MixCase *cases;   // case table (device)
	MixStatus *status;      // per case slot (device)
	const int2 *blocks;     // (case slot, block of the case)
	const int *nBlocks;     // number of entries of `blocks` in use (device scalar)
	int nSlots;             // case slots
	int grid;               // workgroups (a multiple of the XCD count)
	int cap;                // LDS cells per block: kMixBlock + the largest halo of any resident case
	int ms;                 // LDS face slots per block: the largest slot count, even
	int schemes;            // bit s set: some resident case uses flux scheme s (one launch each)
};

// LDS bytes of a block plan with these extremes
size_t mixLdsBytes(int maxSlots, int maxHalo);
// workgroups of the stage kernel that fit on one CU at once with this LDS layout (registers and
// LDS); a persistent launch larger than that would leave a tail of workgroups running alone
int mixResidentPerCU(int ms, int cap, int schemes, bool nontemporal);
// one RK stage (0..4) of every active, unfinished case
void launchMixStage(const MixLaunch &L, int stage, const MixTuning &t, hipStream_t st);
// after the 5 stages: per-case RMS, history, convergence and stall tests, buffer parity
void launchMixIterEnd(const MixLaunch &L, hipStream_t st);
// FP64 primitives (rho, u, v, w, p, E, H, M) of the interior cells from W: out[8][nElem]
// wall and symmetry ghost records of the last completed iteration into rec[parity]
void launchMixMirrorGhosts(const MixCase &hc, int parity, hipStream_t st);
void launchMixPrimitives(const MixCase &hc, double *out, hipStream_t st);
// wall-face force/moment of one case from its FP64 state (same reduction as the FP64 path)
void launchMixAero(int nWall, const int *wallCell, const double *Nx, const double *Ny, const double *Nz,
                   const double *Cx, const double *Cy, const double *Cz, const MixCase *dc, double cpFactor,
                   double pInf, double cref, double sref, double *out6, hipStream_t st);

}// namespace e3d
