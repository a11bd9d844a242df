/*
 * CPU ORACLE of the Euler3D explicit-solver hot path. TEST INFRASTRUCTURE ONLY (see
 * e3d_oracle.h for what pins each function). Every expression keeps the reference's evaluation
 * order; x*x stands for std::pow(x, 2), which g++ folds to the same single multiply. Build with
 * -O2 -ffp-contract=off so that no multiply-add is fused.
 */
#include "e3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* threads of the parallel loops; 1 runs the sequential restatement (or_set_threads) */
static int g_threads = 1;

#define E3D_PI 3.141592653589793238

/* FlowField::Initialize, FlowField.cpp:68-101 (rho_inf = p_inf = 1, FlowField.hpp:138-139) */
void or_freestream(or_params *prm) {
	double g = prm->gamma;
	prm->rho_inf = 1.0;
	prm->p_inf = 1.0;
	prm->M_inf = prm->mach;
	prm->u_inf = prm->M_inf * sqrt(g) * cos(prm->aoa_deg * (E3D_PI / 180));
	prm->v_inf = prm->M_inf * sqrt(g) * sin(prm->aoa_deg * (E3D_PI / 180));
	prm->w_inf = 0.0;
	prm->E_inf = prm->p_inf / ((g - 1) * prm->rho_inf) +
	             ((prm->u_inf * prm->u_inf + prm->v_inf * prm->v_inf + prm->w_inf * prm->w_inf) / 2.0);
	prm->H_inf = prm->E_inf + (prm->p_inf / prm->rho_inf);
}

void or_init_state(const or_mesh *m, const or_params *prm, or_state *s) {
	int n = m->nElem + m->nGhost;
	for (int i = 0; i < n; i++) {
		s->rho[i] = prm->rho_inf;
		s->p[i] = prm->p_inf;
		s->u[i] = prm->u_inf;
		s->v[i] = prm->v_inf;
		s->w[i] = prm->w_inf;
		s->M[i] = prm->M_inf;
		s->E[i] = prm->E_inf;
		s->H[i] = prm->H_inf;
	}
}

static void set_ghost(or_state *s, int g, double rho, double u, double v, double w, double M, double p, double E, double H) {
	s->rho[g] = rho;
	s->u[g] = u;
	s->v[g] = v;
	s->w[g] = w;
	s->M[g] = M;
	s->p[g] = p;
	s->E[g] = E;
	s->H[g] = H;
}

/* PhysicalBC.cpp:43-79 */
static void bc_sub_inflow(const or_params *q, or_state *s, int g, int i, const double *n) {
	double gm = q->gamma;
	double c = sqrt(gm * (s->p[i] / s->rho[i]));
	double p = 0.5 * (q->p_inf + s->p[i] -
	                  s->rho[i] * c * ((n[0] * (q->u_inf - s->u[i])) + (n[1] * (q->v_inf - s->v[i])) + (n[2] * (q->w_inf - s->w[i]))));
	double rho = q->rho_inf + (p - q->p_inf) / (c * c);
	double u = q->u_inf - n[0] * (q->p_inf - p) / (s->rho[i] * c);
	double v = q->v_inf - n[1] * (q->p_inf - p) / (s->rho[i] * c);
	double w = q->w_inf - n[2] * (q->p_inf - p) / (s->rho[i] * c);
	double M = sqrt(u * u + v * v + w * w) / c;
	double E = p / ((gm - 1) * rho) + ((u * u + v * v + w * w) / 2);
	double H = E + (p / rho);
	set_ghost(s, g, rho, u, v, w, M, p, E, H);
}

/* PhysicalBC.cpp:82-121 */
static void bc_sub_outflow(const or_params *q, or_state *s, int g, int i, const double *n) {
	double gm = q->gamma;
	double c = sqrt(gm * (s->p[i] / s->rho[i]));
	double p = q->p_inf;
	double rho = s->rho[i] + (p - s->p[i]) / (c * c);
	double u = s->u[i] + n[0] * (s->p[i] - p) / (s->rho[i] * c);
	double v = s->v[i] + n[1] * (s->p[i] - p) / (s->rho[i] * c);
	double w = s->w[i] + n[2] * (s->p[i] - p) / (s->rho[i] * c);
	double M = sqrt(u * u + v * v + w * w) / c;
	double E = p / ((gm - 1) * rho) + ((u * u + v * v + w * w) / 2);
	double H = E + (p / rho);
	set_ghost(s, g, rho, u, v, w, M, p, E, H);
}

/* PhysicalBC.cpp:124-162 (wall) and 165-231 (symmetry): identical arithmetic */
static void bc_mirror(const or_params *q, or_state *s, int g, int i, const double *n) {
	double gm = q->gamma;
	double c = sqrt(gm * (s->p[i] / s->rho[i]));
	double V = s->u[i] * n[0] + s->v[i] * n[1] + s->w[i] * n[2];
	double u = s->u[i] - 2 * V * n[0];
	double v = s->v[i] - 2 * V * n[1];
	double w = s->w[i] - 2 * V * n[2];
	double p = s->p[i];
	double rho = s->rho[i];
	double M = sqrt(u * u + v * v + w * w) / c;
	double E = p / ((gm - 1) * rho) + ((u * u + v * v + w * w) / 2.0);
	double H = E + (p / rho);
	set_ghost(s, g, rho, u, v, w, M, p, E, H);
}

/* EulerSolver::updateBC, EulerSolver.cpp:201-247 */
void or_update_bc(const or_mesh *m, const or_params *q, or_state *s) {
	for (int k = 0; k < m->nFar; k++) {
		int g = m->farGhost[k], f = m->farFace[k], i = m->farCell[k];
		const double *n = &m->faceUnitNormal[3 * (size_t) f];
		double V = s->u[g] * n[0] + s->v[g] * n[1] + s->w[g] * n[2];
		double M = s->M[g];
		if (M >= 1.0 && V > 0.0) {
			set_ghost(s, g, s->rho[i], s->u[i], s->v[i], s->w[i], s->M[i], s->p[i], s->E[i], s->H[i]);
		} else if (M < 1.0 && V > 0.0) {
			bc_sub_outflow(q, s, g, i, n);
		} else if (M >= 1.0 && V <= 0.0) {
			set_ghost(s, g, q->rho_inf, q->u_inf, q->v_inf, q->w_inf, q->M_inf, q->p_inf, q->E_inf, q->H_inf);
		} else if (M < 1.0 && V <= 0.0) {
			bc_sub_inflow(q, s, g, i, n);
		}
	}
	for (int k = 0; k < m->nWall; k++)
		bc_mirror(q, s, m->wallGhost[k], m->wallCell[k], &m->faceUnitNormal[3 * (size_t) m->wallFace[k]]);
	for (int k = 0; k < m->nSym; k++)
		bc_mirror(q, s, m->symGhost[k], m->symCell[k], &m->faceUnitNormal[3 * (size_t) m->symFace[k]]);
}

/* Solver::Roe, ConvectiveFluxesSchemes.cpp:10-165 */
void or_roe(const or_mesh *m, const or_params *q, const or_state *s, int face, double out[5]) {
	int L = m->face2element[2 * (size_t) face], R = m->face2element[2 * (size_t) face + 1];
	double gamma = q->gamma;
	double rL = s->rho[L], rR = s->rho[R];
	double sL = sqrt(rL), sR = sqrt(rR);
	double rhoT = sqrt(rL * rR);
	double uL = s->u[L], uR = s->u[R], vL = s->v[L], vR = s->v[R], wL = s->w[L], wR = s->w[R];
	double hL = s->H[L], hR = s->H[R], pL = s->p[L], pR = s->p[R];
	double uT = (uL * sL + uR * sR) / (sL + sR);
	double vT = (vL * sL + vR * sR) / (sL + sR);
	double wT = (wL * sL + wR * sR) / (sL + sR);
	double hT = (hL * sL + hR * sR) / (sL + sR);
	double q2 = uT * uT + vT * vT + wT * wT;
	double cT = sqrt((gamma - 1.0) * (hT - q2 / 2));
	const double *n = &m->faceUnitNormal[3 * (size_t) face];
	double VT = uT * n[0] + vT * n[1] + wT * n[2];
	double VL = uL * n[0] + vL * n[1] + wL * n[2];
	double VR = uR * n[0] + vR * n[1] + wR * n[2];
	double dRho = rR - rL, dU = uR - uL, dV = vR - vL, dW = wR - wL, dP = pR - pL, dVc = VR - VL;
	double Fr[5] = {rR * VR, rR * VR * uR + pR * n[0], rR * VR * vR + pR * n[1], rR * VR * wR + pR * n[2], rR * VR * hR};
	double Fl[5] = {rL * VL, rL * VL * uL + pL * n[0], rL * VL * vL + pL * n[1], rL * VL * wL + pL * n[2], rL * VL * hL};
	double avg[5];
	for (int i = 0; i < 5; i++) avg[i] = (Fr[i] + Fl[i]) * 0.5;
	double localC = sqrt(gamma * pL / rL);
	double delta = (1.0 / 15.0) * localC;
	double l1, l5;
	if (fabs(VT - cT) > delta) l1 = fabs(VT - cT);
	else l1 = ((VT - cT) * (VT - cT) + delta * delta) / (2.0 * delta);
	if (fabs(VT + cT) > delta) l5 = fabs(VT + cT);
	else l5 = ((VT + cT) * (VT + cT) + delta * delta) / (2.0 * delta);
	double f1[5] = {1.0, uT - cT * n[0], vT - cT * n[1], wT - cT * n[2], hT - cT * VT};
	for (int i = 0; i < 5; i++) f1[i] = f1[i] * l1 * (dP - rhoT * cT * dVc) / (2.0 * (cT * cT));
	double f234[5] = {1.0, uT, vT, wT, q2 * 0.5};
	for (int i = 0; i < 5; i++) f234[i] = f234[i] * (dRho - dP / (cT * cT));
	double t2[5] = {0.0, rhoT * (dU - dVc * n[0]), rhoT * (dV - dVc * n[1]), rhoT * (dW - dVc * n[2]),
	                rhoT * (uT * dU + vT * dV + wT * dW - VT * dVc)};
	for (int i = 0; i < 5; i++) f234[i] = (f234[i] + t2[i]) * fabs(VT);
	double f5[5] = {1.0, uT + cT * n[0], vT + cT * n[1], wT + cT * n[2], hT + cT * VT};
	for (int i = 0; i < 5; i++) f5[i] = f5[i] * l5 * (dP + rhoT * cT * dVc) / (2.0 * (cT * cT));
	for (int i = 0; i < 5; i++) out[i] = avg[i] - 0.5 * (f1[i] + f234[i] + f5[i]);
}

/* AUSM, Liou & Steffen, J. Comput. Phys. 107 (1993) 23-39: Mach-number splitting of degree 2
 * and pressure splitting of degree 3, convective quantities Phi = rho a (1, u, v, w, H) upwinded by
 * the sign of the interface Mach number m = M+(M_L) + M-(M_R), M = V / a with each side's own
 * sound speed. The reference only declares it (ConvectiveFluxesSchemes.h:27-30); what pins it is
 * listed in e3d_oracle.h.
 *
 * OR_AUSM_MUTANT (compile-time, tests only) plants one plausible mistake so that the tests can show
 * that their checks catch it: 1 swapped pressure splittings, 2 the left sound speed for both sides,
 * 3 the |m| dissipation term dropped, 4 the pressure polynomial (2 - M) written (2 + M). */
#ifndef OR_AUSM_MUTANT
#define OR_AUSM_MUTANT 0
#endif
static double mach_plus(double M) { return fabs(M) <= 1.0 ? 0.25 * (M + 1.0) * (M + 1.0) : 0.5 * (M + fabs(M)); }
static double mach_minus(double M) { return fabs(M) <= 1.0 ? -0.25 * (M - 1.0) * (M - 1.0) : 0.5 * (M - fabs(M)); }
static double pres_plus(double M) {
#if OR_AUSM_MUTANT == 4
	return fabs(M) <= 1.0 ? 0.25 * (M + 1.0) * (M + 1.0) * (2.0 + M) : (M > 0.0 ? 1.0 : 0.0);
#else
	return fabs(M) <= 1.0 ? 0.25 * (M + 1.0) * (M + 1.0) * (2.0 - M) : (M > 0.0 ? 1.0 : 0.0);
#endif
}
static double pres_minus(double M) {
#if OR_AUSM_MUTANT == 4
	return fabs(M) <= 1.0 ? 0.25 * (M - 1.0) * (M - 1.0) * (2.0 - M) : (M < 0.0 ? 1.0 : 0.0);
#else
	return fabs(M) <= 1.0 ? 0.25 * (M - 1.0) * (M - 1.0) * (2.0 + M) : (M < 0.0 ? 1.0 : 0.0);
#endif
}

/* The four splittings at M: out = {M+, M-, P+, P-} (for the identity checks of the tests). */
void or_ausm_split(double M, double out[4]) {
	out[0] = mach_plus(M);
	out[1] = mach_minus(M);
	out[2] = pres_plus(M);
	out[3] = pres_minus(M);
}

void or_ausm(const or_mesh *m, const or_params *q, const or_state *s, int face, double out[5]) {
	int L = m->face2element[2 * (size_t) face], R = m->face2element[2 * (size_t) face + 1];
	const double *n = &m->faceUnitNormal[3 * (size_t) face];
	double g = q->gamma;
	double rL = s->rho[L], rR = s->rho[R], pL = s->p[L], pR = s->p[R];
	double uL = s->u[L], vL = s->v[L], wL = s->w[L], hL = s->H[L];
	double uR = s->u[R], vR = s->v[R], wR = s->w[R], hR = s->H[R];
#if OR_AUSM_MUTANT == 2
	double aL = sqrt(g * pL / rL), aR = aL;
#else
	double aL = sqrt(g * pL / rL), aR = sqrt(g * pR / rR);
#endif
	double VL = uL * n[0] + vL * n[1] + wL * n[2];
	double VR = uR * n[0] + vR * n[1] + wR * n[2];
	double ML = VL / aL, MR = VR / aR;
	double mh = mach_plus(ML) + mach_minus(MR);
#if OR_AUSM_MUTANT == 1
	double ph = pres_minus(ML) * pL + pres_plus(MR) * pR;
#else
	double ph = pres_plus(ML) * pL + pres_minus(MR) * pR;
#endif
	double phiL[5] = {rL * aL, rL * aL * uL, rL * aL * vL, rL * aL * wL, rL * aL * hL};
	double phiR[5] = {rR * aR, rR * aR * uR, rR * aR * vR, rR * aR * wR, rR * aR * hR};
	double pn[5] = {0.0, ph * n[0], ph * n[1], ph * n[2], 0.0};
#if OR_AUSM_MUTANT == 3
	for (int i = 0; i < 5; i++) out[i] = mh * 0.5 * (phiL[i] + phiR[i]) + pn[i];
#else
	for (int i = 0; i < 5; i++) out[i] = mh * 0.5 * (phiL[i] + phiR[i]) - 0.5 * fabs(mh) * (phiR[i] - phiL[i]) + pn[i];
#endif
}

/* Solver::Fc, ConvectiveFluxesSchemes.cpp:167-187 */
void or_fc(const or_mesh *m, const or_state *s, int g, int face, double V, double out[5]) {
	const double *n = &m->faceUnitNormal[3 * (size_t) face];
	double rho = s->rho[g], u = s->u[g], v = s->v[g], w = s->w[g], H = s->H[g], p = s->p[g];
	out[0] = rho * V;
	out[1] = rho * u * V + p * n[0];
	out[2] = rho * v * V + p * n[1];
	out[3] = rho * w * V + p * n[2];
	out[4] = rho * H * V;
}

static int ghost_kind_of(const or_mesh *m, int ghost, const unsigned char *kind) {
	return kind[ghost - m->nElem];
}

/* ghost kind table: 1 wall, 2 symmetry, 3 farfield, with the precedence of
 * EulerSolver.cpp:331-357 (wall first, then symmetry/MPI, everything else farfield) */
static unsigned char *ghost_kinds(const or_mesh *m) {
	unsigned char *k = (unsigned char *) malloc((size_t) m->nGhost + 1);
	memset(k, 3, (size_t) m->nGhost + 1);
	for (int i = 0; i < m->nSym; i++) k[m->symGhost[i] - m->nElem] = 2;
	for (int i = 0; i < m->nWall; i++) k[m->wallGhost[i] - m->nElem] = 1;
	return k;
}

static void face_flux_one(const or_mesh *m, const or_params *q, const or_state *s, const unsigned char *gk, int f,
                          double F[5], int *kindOut) {
	int e2 = m->face2element[2 * (size_t) f + 1];
	int kind;
	if (f < m->nFaceInt) {
		kind = 0;
	} else if (e2 < m->nElem) {
		kind = 4; /* partition interface: Roe with the other partition's cell (EulerSolver.cpp:343-346) */
	} else {
		kind = ghost_kind_of(m, e2, gk);
	}
	if (kind == 0 || kind == 2 || kind == 4) {
		if (q->scheme == 1) or_ausm(m, q, s, f, F);
		else or_roe(m, q, s, f, F);
	} else if (kind == 1) {
		const double *n = &m->faceUnitNormal[3 * (size_t) f];
		F[0] = 0;
		F[1] = n[0] * s->p[e2];
		F[2] = n[1] * s->p[e2];
		F[3] = n[2] * s->p[e2];
		F[4] = 0;
	} else {
		const double *n = &m->faceUnitNormal[3 * (size_t) f];
		double V = s->u[e2] * n[0] + s->v[e2] * n[1] + s->w[e2] * n[2];
		or_fc(m, s, e2, f, V, F);
	}
	if (kindOut) *kindOut = kind;
}

void or_face_flux(const or_mesh *m, const or_params *q, const or_state *s, double *F, double *FS, int *kind) {
	unsigned char *gk = ghost_kinds(m);
	for (int f = 0; f < m->nFace; f++) {
		double r[5];
		face_flux_one(m, q, s, gk, f, r, kind ? &kind[f] : NULL);
		double S = m->faceArea[f];
		for (int k = 0; k < 5; k++) {
			if (F) F[5 * (size_t) f + k] = r[k];
			if (FS) FS[5 * (size_t) f + k] = r[k] * S;
		}
	}
	free(gk);
}

/* Threads of the per-face, per-cell and per-ghost loops (1: the sequential restatement). With
 * more threads the faces' F*S are formed in parallel and each cell then sums its faces in
 * ascending face id, the order of the sequential face loop, so every result keeps its bits. */
void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }

static void residual_parallel(const or_mesh *m, const or_params *q, const or_state *s, double *R,
                              const unsigned char *gk) {
	const int nF = m->nFace, nE = m->nElem;
	double *FS = (double *) malloc(sizeof(double) * 5 * (size_t) nF);
#pragma omp parallel for num_threads(g_threads) schedule(static)
	for (int f = 0; f < nF; f++) {
		double r[5];
		face_flux_one(m, q, s, gk, f, r, NULL);
		double S = m->faceArea[f];
		for (int k = 0; k < 5; k++) FS[5 * (size_t) f + k] = r[k] * S;
	}
#pragma omp parallel for num_threads(g_threads) schedule(static)
	for (int e = 0; e < nE; e++) {
		int b = m->element2faceStart[e], en = m->element2faceStart[e + 1];
		int fl[16], n = 0;
		for (int j = b; j < en && n < 16; j++) fl[n++] = m->element2face[j];
		for (int i = 1; i < n; i++)
			for (int j = i; j > 0 && fl[j - 1] > fl[j]; j--) {
				int t = fl[j];
				fl[j] = fl[j - 1];
				fl[j - 1] = t;
			}
		double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
		for (int i = 0; i < n; i++) {
			int f = fl[i];
			int first = m->face2element[2 * (size_t) f] == e;
			for (int k = 0; k < 5; k++) {
				if (first) acc[k] += FS[5 * (size_t) f + k];
				else acc[k] -= FS[5 * (size_t) f + k];
			}
		}
		for (int k = 0; k < 5; k++) R[5 * (size_t) e + k] = acc[k];
	}
	free(FS);
}

/* EulerSolver::computeResidual, EulerSolver.cpp:303-359 (after resetResiduals, :391-395) */
void or_residual(const or_mesh *m, const or_params *q, const or_state *s, double *R) {
	unsigned char *gk = ghost_kinds(m);
	if (g_threads > 1) {
		residual_parallel(m, q, s, R, gk);
		free(gk);
		return;
	}
	memset(R, 0, sizeof(double) * 5 * (size_t) m->nElem);
	for (int f = 0; f < m->nFace; f++) {
		double r[5];
		face_flux_one(m, q, s, gk, f, r, NULL);
		double S = m->faceArea[f];
		int e1 = m->face2element[2 * (size_t) f], e2 = m->face2element[2 * (size_t) f + 1];
		for (int k = 0; k < 5; k++) {
			double rs = r[k] * S;
			R[5 * (size_t) e1 + k] += rs;
			if (f < m->nFaceInt) R[5 * (size_t) e2 + k] -= rs;
		}
	}
	free(gk);
}

/* TimeStepMethod2, TimeIntegration.cpp:8-40, for every interior cell */
void or_timestep(const or_mesh *m, const or_params *q, const or_state *s, double *dt) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
	for (int e = 0; e < m->nElem; e++) {
		double sx = 0, sy = 0, sz = 0;
		for (int j = m->element2faceStart[e]; j < m->element2faceStart[e + 1]; j++) {
			const double *N = &m->faceNormal[3 * (size_t) m->element2face[j]];
			sx += fabs(N[0]);
			sy += fabs(N[1]);
			sz += fabs(N[2]);
		}
		double c = sqrt(q->gamma * s->p[e] / s->rho[e]);
		double lx = 0.5 * (fabs(s->u[e]) + c) * sx;
		double ly = 0.5 * (fabs(s->v[e]) + c) * sy;
		double lz = 0.5 * (fabs(s->w[e]) + c) * sz;
		dt[e] = q->cfl * m->cellVolume[e] / (lx + ly + lz);
	}
}

/* FlowField::updateWRungeKutta, FlowField.cpp:140-158, with RHS = alpha*dt*R/V formed as in
 * EulerSolver.cpp:441-449 */
/* FlowField.cpp:129-138, 159-169: every interface ghost takes a copy of its own cell */
static void copy_mpi_ghosts(const or_mesh *m, or_state *s) {
	for (int i = 0; i < m->nMpi; i++) {
		int g = m->mpiGhost[i], c = m->mpiCell[i];
		s->rho[g] = s->rho[c], s->u[g] = s->u[c], s->v[g] = s->v[c], s->w[g] = s->w[c];
		s->E[g] = s->E[c], s->p[g] = s->p[c], s->H[g] = s->H[c], s->M[g] = s->M[c];
	}
}

static void rk_update(const or_mesh *m, const or_params *q, or_state *s, const double *W0, const double *R,
                      const double *dt, double alpha) {
	double g = q->gamma;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
	for (int e = 0; e < m->nElem; e++) {
		double V = m->cellVolume[e];
		double d = dt[e];
		double rhs[5];
		for (int k = 0; k < 5; k++) rhs[k] = alpha * d * R[5 * (size_t) e + k] / V;
		double rho = W0[5 * (size_t) e] - rhs[0];
		double ru = W0[5 * (size_t) e + 1] - rhs[1];
		double rv = W0[5 * (size_t) e + 2] - rhs[2];
		double rw = W0[5 * (size_t) e + 3] - rhs[3];
		double rE = W0[5 * (size_t) e + 4] - rhs[4];
		double u = ru / rho, v = rv / rho, w = rw / rho, E = rE / rho;
		double p = (g - 1) * rho * (E - (u * u + v * v + w * w) / 2.0);
		double H = E + (p / rho);
		double M = sqrt(g * (p / rho)) / sqrt((u * u + v * v + w * w));
		s->rho[e] = rho;
		s->u[e] = u;
		s->v[e] = v;
		s->w[e] = w;
		s->E[e] = E;
		s->p[e] = p;
		s->H[e] = H;
		s->M[e] = M;
	}
	copy_mpi_ghosts(m, s);
}

/* EulerSolver::RungeKutta, EulerSolver.cpp:425-472 */
void or_rk5(const or_mesh *m, const or_params *q, or_state *s, double *R) {
	static const double alphas[4] = {0.1263, 0.2375, 0.4414, 1.0};
	size_t n = (size_t) m->nElem;
	double *W0 = (double *) malloc(sizeof(double) * 5 * n);
	double *dt = (double *) malloc(sizeof(double) * n);
	for (size_t i = 0; i < n; i++) {
		double rho = s->rho[i];
		W0[5 * i] = rho;
		W0[5 * i + 1] = s->u[i] * rho;
		W0[5 * i + 2] = s->v[i] * rho;
		W0[5 * i + 3] = s->w[i] * rho;
		W0[5 * i + 4] = s->E[i] * rho;
	}
	or_timestep(m, q, s, dt);
	rk_update(m, q, s, W0, R, dt, 0.0533);
	for (int a = 0; a < 4; a++) {
		or_update_bc(m, q, s);
		or_residual(m, q, s, R);
		or_timestep(m, q, s, dt);
		rk_update(m, q, s, W0, R, dt, alphas[a]);
	}
	free(W0);
	free(dt);
}

/* EulerSolver::computeRMS, EulerSolver.cpp:382-389 */
double or_rms_sum(const double *R, int nElem) {
	double err = 0.0;
	for (int e = 0; e < nElem; e++) err += R[5 * (size_t) e] * R[5 * (size_t) e];
	return err;
}

/* AeroCoefficients ctor + Update, AeroCoefficients.cpp:5-52. The centre of pressure is never
 * divided (Vector3::operator/=(Vector3) returns a new vector, Vector3.h:107-109). */
static void aero_range(const or_mesh *m, const or_params *q, const or_state *s, int w0, int w1, double cf[3],
                       double cm[3]);

/* With several partitions each computes its own force and moment vectors, and rank 0 adds them
 * in rank order (MPIHandler::UpdateAerodynamicCoefficients, MPIHandler.cpp:268-313). */
void or_aero(const or_mesh *m, const or_params *q, const or_state *s, double cf[3], double cm[3]) {
	if (m->nPart <= 1) {
		aero_range(m, q, s, 0, m->nWall, cf, cm);
		return;
	}
	for (int r = 0; r < m->nPart; r++) {
		double f[3], c[3];
		aero_range(m, q, s, m->partWallStart[r], m->partWallStart[r + 1], f, c);
		for (int k = 0; k < 3; k++) {
			cf[k] = r ? cf[k] + f[k] : f[k];
			cm[k] = r ? cm[k] + c[k] : c[k];
		}
	}
}

static void aero_range(const or_mesh *m, const or_params *q, const or_state *s, int w0, int w1, double cf[3],
                       double cm[3]) {
	double uinf2 = q->u_inf * q->u_inf + q->v_inf * q->v_inf + q->w_inf * q->w_inf;
	double cpFactor = 1 / (0.5 * q->rho_inf * uinf2);
	double F[3] = {0, 0, 0}, CP[3] = {0, 0, 0};
	for (int i = w0; i < w1; i++) {
		int cell = m->wallCell[i];
		int f = m->wallFace[i];
		double cp = cpFactor * (s->p[cell] - q->p_inf);
		const double *N = &m->faceNormal[3 * (size_t) f];
		const double *C = &m->faceCenter[3 * (size_t) f];
		double co[3];
		for (int k = 0; k < 3; k++) co[k] = N[k] * cp * q->cref / q->sref;
		for (int k = 0; k < 3; k++) F[k] += co[k];
		for (int k = 0; k < 3; k++) CP[k] += co[k] * C[k];
	}
	double lever[3] = {q->ref_point[0] - CP[0], q->ref_point[1] - CP[1], q->ref_point[2] - CP[2]};
	double M[3] = {F[1] * lever[2] - F[2] * lever[1], F[2] * lever[0] - F[0] * lever[2], F[0] * lever[1] - F[1] * lever[0]};
	for (int k = 0; k < 3; k++) {
		cf[k] = F[k];
		cm[k] = M[k] / q->cref;
	}
}

/* EulerSolver::smoothResiduals, EulerSolver.cpp:483-518: Jacobi sweeps
 * R_e = (R0_e + sum over interior neighbours of 0.25 R_s) / (1 + 0.25 * (number of faces)), the
 * neighbours taken in local-face order. The loop is written as "repeat until max |change| <=
 * 1e-16", but the vector of changes is cleared at the end of each pass (:515) and the next pass
 * writes its changes through begin() of the empty vector and scans zero elements, so the second
 * pass always reports no change: one sweep when the first changes nothing, else exactly two.
 * Only the first pass checks for NaN (ResidualVar::findMax exits the process on one). */
static int smooth_range(const or_mesh *m, const or_params *q, double *R, size_t e0, size_t e1);

/* Several partitions: each rank runs its own loop over its own cells (interface neighbours are
 * ghosts there), so each partition stops on its own first sweep. The return value is the largest
 * number of sweeps, or -1 on a NaN. */
int or_smooth(const or_mesh *m, const or_params *q, double *R) {
	if (m->nPart <= 1) return smooth_range(m, q, R, 0, (size_t) m->nElem);
	int most = 0;
	for (int r = 0; r < m->nPart; r++) {
		int k = smooth_range(m, q, R, (size_t) m->partElemStart[r], (size_t) m->partElemStart[r + 1]);
		if (k < 0) return -1;
		if (k > most) most = k;
	}
	return most;
}

/* one partition's sweeps over cells [e0, e1) (whose interior neighbours are all in that range) */
static int smooth_range(const or_mesh *m, const or_params *q, double *R, size_t e0, size_t e1) {
	const double epsilon = 0.25;
	const size_t n = (size_t) m->nElem;
	double *orig = (double *) malloc(sizeof(double) * 5 * n);
	double *last = (double *) malloc(sizeof(double) * 5 * n);
	memcpy(orig, R, sizeof(double) * 5 * n);
	int sweeps = 0;
	double error;
	do {
		error = 0.0;
		memcpy(last, R, sizeof(double) * 5 * n);
		for (size_t e = e0; e < e1; e++) {
			double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
			int b = m->element2faceStart[e], en = m->element2faceStart[e + 1];
			int ns = en - b;
			for (int i = b; i < en; i++) {
				int s = q->element2element[i];
				if (s < m->nElem)
					for (int k = 0; k < 5; k++) sum[k] += last[5 * (size_t) s + k] * epsilon;
			}
			double den = 1 + ns * epsilon;
			for (int k = 0; k < 5; k++) R[5 * e + k] = (orig[5 * e + k] + sum[k]) / den;
		}
		if (sweeps == 0) {
			for (size_t i = 5 * e0; i < 5 * e1; i++) {
				double d = fabs(last[i] - R[i]);
				if (isnan(d)) {
					free(orig);
					free(last);
					return -1;
				}
				if (d > error) error = d;
			}
		}
		sweeps++;
	} while (error > 1e-16);
	free(orig);
	free(last);
	return sweeps;
}

/* FlowField::Initialize, FlowField.cpp:83-95: the conservative arrays start at rho_inf * (u, v, w, E)_inf */
void or_init_conservative(const or_mesh *m, const or_params *q, const or_state *s, double *W) {
	(void) s;
	for (int e = 0; e < m->nElem; e++) {
		W[5 * (size_t) e] = q->rho_inf;
		W[5 * (size_t) e + 1] = q->rho_inf * q->u_inf;
		W[5 * (size_t) e + 2] = q->rho_inf * q->v_inf;
		W[5 * (size_t) e + 3] = q->rho_inf * q->w_inf;
		W[5 * (size_t) e + 4] = q->rho_inf * q->E_inf;
	}
}

/* EulerSolver::EulerExplicit (EulerSolver.cpp:419-423): dt, dW = R * (-dt / V) (ExplicitEuler,
 * TimeIntegration.cpp:78-91), FlowField::Update (FlowField.cpp:103-126): rho += dW.rho (the
 * density array is both primitive and conservative), rho u += dW.rhoU, ..., then the primitives */
void or_euler(const or_mesh *m, const or_params *q, or_state *s, double *W, const double *R) {
	double g = q->gamma;
	size_t n = (size_t) m->nElem;
	double *dt = (double *) malloc(sizeof(double) * n);
	or_timestep(m, q, s, dt);
	for (size_t e = 0; e < n; e++) {
		double coeff = -dt[e] / m->cellVolume[e];
		double d[5];
		for (int k = 0; k < 5; k++) d[k] = R[5 * e + k] * coeff;
		double rho = s->rho[e] + d[0];
		W[5 * e] = rho;
		W[5 * e + 1] += d[1];
		W[5 * e + 2] += d[2];
		W[5 * e + 3] += d[3];
		W[5 * e + 4] += d[4];
		double u = W[5 * e + 1] / rho, v = W[5 * e + 2] / rho, w = W[5 * e + 3] / rho, E = W[5 * e + 4] / rho;
		double p = (g - 1) * rho * (E - (u * u + v * v + w * w) / 2.0);
		double H = E + (p / rho);
		double M = sqrt(g * (p / rho)) / sqrt((u * u + v * v + w * w));
		s->rho[e] = rho;
		s->u[e] = u;
		s->v[e] = v;
		s->w[e] = w;
		s->E[e] = E;
		s->p[e] = p;
		s->H[e] = H;
		s->M[e] = M;
	}
	copy_mpi_ghosts(m, s);
	free(dt);
}

/* EulerSolver::RungeKutta with RESIDUAL_SMOOTHING (EulerSolver.cpp:450-456: smoothing after each
 * stage's residual) */
static int rk5_smooth(const or_mesh *m, const or_params *q, or_state *s, double *R) {
	static const double alphas[4] = {0.1263, 0.2375, 0.4414, 1.0};
	size_t n = (size_t) m->nElem;
	double *W0 = (double *) malloc(sizeof(double) * 5 * n);
	double *dt = (double *) malloc(sizeof(double) * n);
	for (size_t i = 0; i < n; i++) {
		double rho = s->rho[i];
		W0[5 * i] = rho;
		W0[5 * i + 1] = s->u[i] * rho;
		W0[5 * i + 2] = s->v[i] * rho;
		W0[5 * i + 3] = s->w[i] * rho;
		W0[5 * i + 4] = s->E[i] * rho;
	}
	or_timestep(m, q, s, dt);
	rk_update(m, q, s, W0, R, dt, 0.0533);
	int rc = 0;
	for (int a = 0; a < 4 && rc >= 0; a++) {
		or_update_bc(m, q, s);
		or_residual(m, q, s, R);
		if (q->residual_smoothing) rc = or_smooth(m, q, R);
		or_timestep(m, q, s, dt);
		rk_update(m, q, s, W0, R, dt, alphas[a]);
	}
	free(W0);
	free(dt);
	return rc < 0 ? -1 : 0;
}

/* EulerSolver::Run, EulerSolver.cpp:60-167, any TIME_INTEGRATION and RESIDUAL_SMOOTHING */
int or_run_any(const or_mesh *m, const or_params *q, or_state *s, double *W, double *R, double *rows, int max_rows,
               int *iters, double coeffs[3], double cfOut[3], double cmOut[3]) {
	int nb = 0, counter = 0, nrows = 0;
	double CL = 0, CD = 0, CM = 0;
	double cf[3] = {0, 0, 0}, cm[3] = {0, 0, 0};
#define E3D_ROW(it, rms)                                                          \
	do {                                                                          \
		CL = cf[q->orient[0][0]] * q->orient[0][1];                               \
		CD = cf[q->orient[1][0]] * q->orient[1][1];                               \
		CM = cm[q->orient[2][0]] * q->orient[2][1];                               \
		if (nrows < max_rows) {                                                   \
			double *r_ = &rows[8 * (size_t) nrows];                               \
			r_[0] = (it);                                                         \
			r_[1] = (rms);                                                        \
			r_[2] = cf[0], r_[3] = cf[1], r_[4] = cf[2];                          \
			r_[5] = cm[0], r_[6] = cm[1], r_[7] = cm[2];                          \
		}                                                                         \
		nrows++;                                                                  \
	} while (0)
	while (nb < q->max_iter) {
		or_update_bc(m, q, s);
		or_residual(m, q, s, R);
		if (q->residual_smoothing && or_smooth(m, q, R) < 0) return -1;
		if (q->time_integration == 2) {
			if (q->residual_smoothing) {
				if (rk5_smooth(m, q, s, R) < 0) return -1;
			} else {
				or_rk5(m, q, s, R);
			}
		} else if (q->time_integration == 0) {
			or_euler(m, q, s, W, R);
		}
		/* TIME_INTEGRATION 1 (EulerImplicit, EulerSolver.cpp:474-480) only recomputes dt */
		double err = or_rms_sum(R, m->nElem);
		double rms = sqrt(err / m->nElem);
		if (rms < q->min_residual) {
			E3D_ROW(nb, rms);
			break;
		}
		nb += 1;
		if (nb % q->sampling == 0) {
			double oCL = CL, oCD = CD, oCM = CM;
			or_aero(m, q, s, cf, cm);
			E3D_ROW(nb, rms);
			double tol = q->min_aerocoeff_error;
			int crit = fabs(oCL - CL) < tol && fabs(oCD - CD) < tol && fabs(oCM - CM) < tol;
			counter = crit ? counter + 1 : 0;
			if (counter >= 2) break;
		}
	}
#undef E3D_ROW
	*iters = nb;
	coeffs[0] = CL;
	coeffs[1] = CD;
	coeffs[2] = CM;
	for (int k = 0; k < 3; k++) {
		cfOut[k] = cf[k];
		cmOut[k] = cm[k];
	}
	return nrows;
}

/* EulerSolver::Run, EulerSolver.cpp:60-167, TIME_INTEGRATION = 2 (RK5), no residual smoothing */
int or_run(const or_mesh *m, const or_params *q, or_state *s, double *R, double *rows, int max_rows, int *iters,
           double coeffs[3], double cfOut[3], double cmOut[3]) {
	int nb = 0, counter = 0, nrows = 0;
	double CL = 0, CD = 0, CM = 0;
	double cf[3] = {0, 0, 0}, cm[3] = {0, 0, 0};
#define E3D_ROW(it, rms)                                                          \
	do {                                                                          \
		CL = cf[q->orient[0][0]] * q->orient[0][1];                               \
		CD = cf[q->orient[1][0]] * q->orient[1][1];                               \
		CM = cm[q->orient[2][0]] * q->orient[2][1];                               \
		if (nrows < max_rows) {                                                   \
			double *r_ = &rows[8 * (size_t) nrows];                               \
			r_[0] = (it);                                                         \
			r_[1] = (rms);                                                        \
			r_[2] = cf[0], r_[3] = cf[1], r_[4] = cf[2];                          \
			r_[5] = cm[0], r_[6] = cm[1], r_[7] = cm[2];                          \
		}                                                                         \
		nrows++;                                                                  \
	} while (0)
	while (nb < q->max_iter) {
		or_update_bc(m, q, s);
		or_residual(m, q, s, R);
		or_rk5(m, q, s, R);
		double err = or_rms_sum(R, m->nElem);
		double rms = sqrt(err / m->nElem);
		if (rms < q->min_residual) {
			E3D_ROW(nb, rms);
			break;
		}
		nb += 1;
		if (nb % q->sampling == 0) {
			double oCL = CL, oCD = CD, oCM = CM;
			or_aero(m, q, s, cf, cm);
			E3D_ROW(nb, rms);
			double tol = q->min_aerocoeff_error;
			int crit = fabs(oCL - CL) < tol && fabs(oCD - CD) < tol && fabs(oCM - CM) < tol;
			counter = crit ? counter + 1 : 0;
			if (counter >= 2) break;
		}
	}
#undef E3D_ROW
	*iters = nb;
	coeffs[0] = CL;
	coeffs[1] = CD;
	coeffs[2] = CM;
	for (int k = 0; k < 3; k++) {
		cfOut[k] = cf[k];
		cmOut[k] = cm[k];
	}
	return nrows;
}
