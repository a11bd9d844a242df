// Device-resident solver data for one case and the launchers of the HIP kernels that advance it.
//
// Layout in HBM (all FP64 unless noted, structure-of-arrays so that a wavefront's 64 lanes load
// 64 consecutive values):
//   faces  [nFace]   e0, e1 (int32; e1 >= nElem is ghost e1 - nElem), unit normal nx, ny, nz, area
//   ghosts [nGhost]  kind (int8: 0 farfield, 1 wall, 2 symmetry); ghost k sits on face nFaceInt + k
//   cells  [nElem]   volume, sum|Nx|, sum|Ny|, sum|Nz| (static parts of the local time step),
//                    face CSR (int32 face id | sign bit, ascending face id: the reference's
//                    accumulation order)
//   state  [nElem + nGhost]  rho, u, v, w, p, E, H, M  (ghost part persists between stages, as the
//                    farfield classifier reads the ghost's previous state)
//   W0     [5][nElem] conservative state at the start of the iteration
//   flux   [5][nFace] F * S of every face for the current stage
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace e3d {

struct DevFaces {
	const int *e0, *e1;
	const double *nx, *ny, *nz, *S;
};

struct DevCells {
	const double *vol, *sx, *sy, *sz;
	const int *fstart; // nElem + 1
	const int *flist;  // face id, bit 31 set when the cell is the face's elem1 (subtract)
};

struct DevState {
	double *rho, *u, *v, *w, *p, *E, *H, *M;
};

struct DevFree {
	double gamma, cfl;
	double rho_inf, p_inf, u_inf, v_inf, w_inf, M_inf, E_inf, H_inf;
};

struct DevMesh {
	int nElem, nFace, nFaceInt, nGhost;
	DevFaces f;
	DevCells c;
	const int8_t *gkind; // nGhost
};

// per-face flux output: F*S into fs[5][nFace]; when raw != nullptr also the unscaled F
void launchBc(const DevMesh &m, DevState s, const DevFree &q, int machDerived, hipStream_t st);
void launchFlux(const DevMesh &m, DevState s, const DevFree &q, int scheme, double *fs, double *raw, hipStream_t st);
// gather-only: R[5][nElem] from fs (parity hook)
void launchGather(const DevMesh &m, const double *fs, double *R, hipStream_t st);
void launchTimestep(const DevMesh &m, DevState s, const DevFree &q, double *dt, hipStream_t st);
// one RK stage: gather residual, local dt, W_k = W0 - alpha*dt*R/V, primitive update.
// stage 0 also stores W0 from the current primitives; when rmsPartial != nullptr each block writes
// its partial sum of R_rho^2; when Rout != nullptr the residual is stored.
void launchStage(const DevMesh &m, DevState s, const DevFree &q, const double *fs, double *W0, double alpha,
                 int firstStage, double *rmsPartial, double *Rout, hipStream_t st);
int stageBlocks(int nElem);
// deterministic sum of n doubles into out[0] (single workgroup, fixed order)
void launchSum(const double *in, int n, double *out, hipStream_t st);
// interior Mach number from the primitives (FlowField.cpp:157), for state read-back
void launchMach(const DevMesh &m, DevState s, const DevFree &q, hipStream_t st);
// wall-face force/moment: out[6] = (Fx,Fy,Fz, CPx,CPy,CPz) summed in a fixed order
void launchAero(int nWall, const int *wallCell, const double *Nx, const double *Ny, const double *Nz,
                const double *Cx, const double *Cy, const double *Cz, const double *p, double cpFactor, double pInf,
                double cref, double sref, double *out6, hipStream_t st);

}// namespace e3d
