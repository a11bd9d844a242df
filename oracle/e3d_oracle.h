/*
 * CPU ORACLE of the Euler3D explicit-solver hot path. TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the reference's FP64 arithmetic, in the reference's
 * evaluation order, used to check the HIP path (tests/, the smoke check and the cpu_baseline leg
 * of the benchmark). The product never links, loads or calls it.
 *
 * Parity status of each function (what pins it):
 *   or_freestream      FlowField.cpp:68-101        golden fixtures (reference harness "mesh"/"run")
 *   or_update_bc       EulerSolver.cpp:201-251,
 *                      PhysicalBC.cpp:11-231       golden fixtures ("flux": bc_* states)
 *   or_roe             ConvectiveFluxesSchemes.cpp:10-165   golden fixtures (per-face flux)
 *   or_fc              ConvectiveFluxesSchemes.cpp:167-187  golden fixtures (per-face flux)
 *   or_face_flux       EulerSolver.cpp:306-358     golden fixtures (per-face flux)
 *   or_residual        EulerSolver.cpp:303-359     golden fixtures (residual)
 *   or_timestep        TimeIntegration.cpp:8-40    golden fixtures (dt)
 *   or_rk5             EulerSolver.cpp:425-472, FlowField.cpp:140-171  golden fixtures (rk_*)
 *   or_rms_sum         EulerSolver.cpp:382-389     golden fixtures (rk_rms_sum, run_rms_sum)
 *   or_aero            AeroCoefficients.cpp:5-52   golden fixtures (force/moment coeff, CL/CD/CM)
 *   or_run             EulerSolver.cpp:60-167      golden fixtures (residuals.dat, final state)
 *   or_smooth          EulerSolver.cpp:483-518     golden fixtures (runs with RESIDUAL_SMOOTHING= 1)
 *   or_euler           EulerSolver.cpp:419-423, TimeIntegration.cpp:78-91, FlowField.cpp:103-126
 *                                                  golden fixtures (runs with TIME_INTEGRATION= 0)
 *   or_run_any         EulerSolver.cpp:60-167      golden fixtures (TIME_INTEGRATION 0/1/2, smoothing)
 *   partitioned meshes (nPart > 1): interface faces, per-partition coefficients and smoothing
 *                      stops (MPIHandler.cpp:16-313, FlowField.cpp:129-169)
 *                                                  golden fixtures of mpiexec -np N runs (r<r>_*)
 *   or_ausm            no reference implementation exists (ConvectiveFluxesSchemes.h:27-30 only
 *                      declares it), so no fixture can pin it. It restates Liou & Steffen,
 *                      J. Comput. Phys. 107 (1993) 23-39, and is pinned by what the mathematics
 *                      fixes (tests/test_ausm_invariants.py): the splittings are the unique
 *                      polynomials with M+ + M- = M, P+ + P- = 1 and C1 joins at |M| = 1;
 *                      consistency F(U,U,n) = physical flux; exact upwinding for M_L, M_R > 1 and
 *                      < -1; mirror antisymmetry F(L,R,n) = -F(R,L,-n) (bitwise); rotation
 *                      invariance; conservation on a closed mesh. Each planted mistake
 *                      (OR_AUSM_MUTANT 1-4) fails at least one of them.
 */
#ifndef E3D_ORACLE_H
#define E3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	int nElem, nFace, nFaceInt, nGhost;
	const int *face2element;      /* 2*nFace, (elem0, elem1) */
	const int *element2faceStart; /* nElem+1 */
	const int *element2face;      /* local-face order */
	const double *faceNormal;     /* 3*nFace area-weighted */
	const double *faceUnitNormal; /* 3*nFace */
	const double *faceArea;       /* nFace */
	const double *faceCenter;     /* 3*nFace */
	const double *cellVolume;     /* nElem */
	int nWall, nFar, nSym;
	const int *wallGhost, *wallCell, *wallFace;
	const int *farGhost, *farCell, *farFace;
	const int *symGhost, *symCell, *symFace;
	/* several partitions as one mesh (zero / NULL otherwise): interface faces are boundary faces
	 * whose elem1 is the other partition's cell; each interface ghost copies its own cell after an
	 * update; cells [partElemStart[r], partElemStart[r+1]) and wall entries [partWallStart[r], ...)
	 * belong to partition r */
	int nMpi;
	const int *mpiGhost, *mpiCell;
	int nPart;
	const int *partElemStart, *partWallStart;
} or_mesh;

typedef struct {
	double gamma, cfl, mach, aoa_deg;
	double sref, cref, ref_point[3];
	int orient[3][2]; /* CL, CD, CM: (axis, sign) */
	int scheme;       /* 0 Roe, 1 AUSM */
	int max_iter, sampling;
	double min_residual, min_aerocoeff_error;
	/* filled by or_freestream */
	double rho_inf, p_inf, u_inf, v_inf, w_inf, E_inf, H_inf, M_inf;
	int time_integration;   /* TIME_INTEGRATION: 0 explicit Euler, 1 implicit Euler (no update), 2 RK5 */
	int residual_smoothing; /* RESIDUAL_SMOOTHING */
	const int *element2element; /* neighbour per local face (CSR by element2faceStart), for smoothing */
} or_params;

typedef struct {
	double *rho, *u, *v, *w, *p, *E, *H, *M; /* nElem + nGhost */
} or_state;

/* Threads for the parallel loops (default 1: the plain sequential restatement). Results keep
 * their bits for any count: cells still sum their faces in ascending face id. */
void or_set_threads(int n);
void or_freestream(or_params *prm);
void or_init_state(const or_mesh *m, const or_params *prm, or_state *s);
void or_update_bc(const or_mesh *m, const or_params *prm, or_state *s);
void or_roe(const or_mesh *m, const or_params *prm, const or_state *s, int face, double out[5]);
void or_ausm(const or_mesh *m, const or_params *prm, const or_state *s, int face, double out[5]);
/* {M+, M-, P+, P-} of or_ausm's splittings at Mach number M */
void or_ausm_split(double M, double out[4]);
void or_fc(const or_mesh *m, const or_state *s, int ghost, int face, double V, double out[5]);
/* per-face flux as the residual dispatches it; F unscaled and FS = F*S (5 per face), kind per face
 * (0 interior, 1 wall, 2 symmetry, 3 farfield) */
void or_face_flux(const or_mesh *m, const or_params *prm, const or_state *s, double *F, double *FS, int *kind);
void or_residual(const or_mesh *m, const or_params *prm, const or_state *s, double *R /* 5*nElem */);
void or_timestep(const or_mesh *m, const or_params *prm, const or_state *s, double *dt);
/* One RK5 step; R holds the residual of the current state on entry (as Run computes it before
 * calling the integrator) and the last-stage residual on exit. */
void or_rk5(const or_mesh *m, const or_params *prm, or_state *s, double *R);
double or_rms_sum(const double *R, int nElem);
/* wall-face force and moment coefficients (AeroCoefficients::Update) */
void or_aero(const or_mesh *m, const or_params *prm, const or_state *s, double cf[3], double cm[3]);
/* Run loop. Writes per-row history (up to max_rows rows of 8 doubles: iter, rms, Cx, Cy, Cz,
 * CMx, CMy, CMz as residuals.dat holds them) and returns the number of rows; *iters receives the
 * final iteration count and coeffs[3] CL, CD, CM. R (5*nElem) is scratch and holds the last
 * residual on exit. */
/* Implicit residual smoothing as the reference runs it: one Jacobi sweep when the first changes
 * nothing, otherwise exactly two (its convergence test scans an emptied vector from the second
 * pass on). Returns the number of sweeps, or -1 when the first sweep meets a NaN residual (the
 * reference exits the process). */
int or_smooth(const or_mesh *m, const or_params *prm, double *R);
/* One explicit Euler step: W (5*nElem conservative, kept between steps) += R * (-dt / V), then the
 * primitives from W. */
void or_euler(const or_mesh *m, const or_params *prm, or_state *s, double *W, const double *R);
/* Run loop for every TIME_INTEGRATION / RESIDUAL_SMOOTHING choice. W (5*nElem) is the explicit
 * Euler conservative state (FlowField's rho u, rho v, rho w, rho E arrays; set to the freestream
 * by or_init_conservative). Returns the number of rows as or_run does, or -1 on a NaN residual in
 * the smoothing. */
void or_init_conservative(const or_mesh *m, const or_params *prm, const or_state *s, double *W);
int or_run_any(const or_mesh *m, const or_params *prm, or_state *s, double *W, double *R, double *rows,
               int max_rows, int *iters, double coeffs[3], double cf[3], double cm[3]);
int or_run(const or_mesh *m, const or_params *prm, or_state *s, double *R, double *rows, int max_rows,
           int *iters, double coeffs[3], double cf[3], double cm[3]);

#ifdef __cplusplus
}
#endif
#endif
