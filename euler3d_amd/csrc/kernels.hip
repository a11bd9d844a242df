// HIP kernels of the explicit FP64 solver path for gfx950 (wave64).
//
// Arithmetic follows the reference expression by expression (file:line given per function) and
// the file is compiled with -ffp-contract=off, so every add and multiply rounds as the
// reference's does; FP64 division and sqrt are the IEEE correctly rounded sequences, and x / 2
// is written x * 0.5 (the same correctly rounded value). Assembly is a per-cell gather in
// ascending reference face id, the reference's accumulation order, so the result is
// deterministic and bit-identical to the sequential face loop without atomics.
//
// Cell state is array-of-structures, 8 doubles = 64 B per cell (rho, u, v, w, p, E, H, M): a
// face reading a neighbour cell touches one 64 B chunk with four 16 B loads instead of six
// scattered cache lines. Between RK stages the fused kernel keeps the interior cells as the
// conservative state instead (rho, rho u, rho v, rho w, rho E: 40 B per cell) and forms the
// primitives on load with the update's own expressions (primOf), which gives the same bits.
//
// Two forms of an RK stage exist:
//   kStageFused  the production path: one workgroup per block of kCellsPerBlock consecutive cells
//                computes every face flux touching the block once into LDS (boundary ghosts
//                included), then each thread gathers its cell's fluxes from LDS, forms the local
//                time step and updates the cell, ping-ponging between two state buffers.
//   kBc, kFlux, kStage (+ kGather, kTimestep)  one pass per step through HBM, used by the
//                per-face parity entry points (update_bc, face_flux, residual, timestep).
#include "device.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <mutex>

namespace e3d {

namespace {

constexpr int kBlock = 256;
static_assert(kBlock == kCellsPerBlock, "one thread per cell in the fused stage");

struct Prim {
	double rho, u, v, w, p, H;
};

struct Ghost {
	double rho, u, v, w, p, E, H, M;
};

// streamed (read-once-per-stage) arrays: optionally loaded non-temporally so that they do not
// evict the cell state that neighbouring faces and blocks re-read from L2
template<bool NT, typename T>
__device__ __forceinline__ T ldStream(const T *p) {
	if constexpr (NT) return __builtin_nontemporal_load(p);
	else return *p;
}

__device__ __forceinline__ double *cellPtr(const DevState &s, int i) { return s.a + (size_t) i * kStateWidth; }

__device__ __forceinline__ Prim loadPrim(const DevState &s, int i) {
	const double2 *c = reinterpret_cast<const double2 *>(cellPtr(s, i));
	const double2 a = c[0], b = c[1], d = c[2], h = c[3]; // (rho,u) (v,w) (p,E) (H,M)
	return {a.x, a.y, b.x, b.y, d.x, h.x};
}

__device__ __forceinline__ void storeGhost(const DevState &s, int g, const Ghost &G) {
	double2 *c = reinterpret_cast<double2 *>(cellPtr(s, g));
	c[0] = make_double2(G.rho, G.u);
	c[1] = make_double2(G.v, G.w);
	c[2] = make_double2(G.p, G.E);
	c[3] = make_double2(G.H, G.M);
}

// Solver::Roe, ConvectiveFluxesSchemes.cpp:10-165
__device__ __forceinline__ void roeFlux(const Prim &L, const Prim &R, double nx, double ny, double nz, double gamma,
                                        double out[5]) {
	double sL = sqrt(L.rho), sR = sqrt(R.rho);
	double rhoT = sqrt(L.rho * R.rho);
	double uT = (L.u * sL + R.u * sR) / (sL + sR);
	double vT = (L.v * sL + R.v * sR) / (sL + sR);
	double wT = (L.w * sL + R.w * sR) / (sL + sR);
	double hT = (L.H * sL + R.H * sR) / (sL + sR);
	double q2 = uT * uT + vT * vT + wT * wT;
	double cT = sqrt((gamma - 1.0) * (hT - q2 * 0.5));
	double VT = uT * nx + vT * ny + wT * nz;
	double VL = L.u * nx + L.v * ny + L.w * nz;
	double VR = R.u * nx + R.v * ny + R.w * nz;
	double dRho = R.rho - L.rho, dU = R.u - L.u, dV = R.v - L.v, dW = R.w - L.w, dP = R.p - L.p, dVc = VR - VL;
	double avg[5];
	avg[0] = (R.rho * VR + L.rho * VL) * 0.5;
	avg[1] = ((R.rho * VR * R.u + R.p * nx) + (L.rho * VL * L.u + L.p * nx)) * 0.5;
	avg[2] = ((R.rho * VR * R.v + R.p * ny) + (L.rho * VL * L.v + L.p * ny)) * 0.5;
	avg[3] = ((R.rho * VR * R.w + R.p * nz) + (L.rho * VL * L.w + L.p * nz)) * 0.5;
	avg[4] = (R.rho * VR * R.H + L.rho * VL * L.H) * 0.5;
	double localC = sqrt(gamma * L.p / L.rho);
	double delta = (1.0 / 15.0) * localC;
	double a1 = VT - cT, a5 = VT + cT;
	double l1 = fabs(a1) > delta ? fabs(a1) : (a1 * a1 + delta * delta) / (2.0 * delta);
	double l5 = fabs(a5) > delta ? fabs(a5) : (a5 * a5 + delta * delta) / (2.0 * delta);
	double c2x2 = 2.0 * (cT * cT);
	double k1 = dP - rhoT * cT * dVc;
	double k5 = dP + rhoT * cT * dVc;
	double kr = dRho - dP / (cT * cT);
	double aV = fabs(VT);
	double e1[5] = {1.0, uT - cT * nx, vT - cT * ny, wT - cT * nz, hT - cT * VT};
	double e2[5] = {1.0, uT, vT, wT, q2 * 0.5};
	double t2[5] = {0.0, rhoT * (dU - dVc * nx), rhoT * (dV - dVc * ny), rhoT * (dW - dVc * nz),
	                rhoT * (uT * dU + vT * dV + wT * dW - VT * dVc)};
	double e5[5] = {1.0, uT + cT * nx, vT + cT * ny, wT + cT * nz, hT + cT * VT};
#pragma unroll
	for (int i = 0; i < 5; i++) {
		double f1 = e1[i] * l1 * k1 / c2x2;
		double f234 = (e2[i] * kr + t2[i]) * aV;
		double f5 = e5[i] * l5 * k5 / c2x2;
		out[i] = avg[i] - 0.5 * (f1 + f234 + f5);
	}
}

// Quotient x / d given y = RN(1 / d), correctly rounded (Markstein, IBM J. Res. Dev. 34 (1990)
// 111-119): q = RN(x y) is within one ulp of x / d, the remainder x - q d is exact under an FMA,
// and RN(q + (x - q d) y) is RN(x / d). Several quotients by one denominator then cost one IEEE
// division plus a multiply and two FMAs each. The result takes q's sign (one bitfield insert), so
// that a zero quotient keeps the sign IEEE division gives it. (Finite, normal operands; the
// solver's values are.)
__device__ __forceinline__ double divRecip(double x, double d, double y) {
	const double q = x * y;
	const double e = fma(-q, d, x);
	const double r = fma(e, y, q);
	// r has the sign of q whenever q is not zero; for a zero q (x = +-0) the sign must be q's
	return copysign(r, q);
}

// x / d either as the IEEE division or through the shared reciprocal y of d
template<bool MK>
__device__ __forceinline__ double quot(double x, double d, double y) {
	if constexpr (MK) return divRecip(x, d, y);
	else return x / d;
}

// Correctly rounded square root of a positive normal x >= 2^-767, and of +0, -0 and +inf: the
// reciprocal square-root estimate refined by the coupled Newton iteration for g ~ sqrt(x) and
// h ~ 1 / (2 sqrt(x)), then two FMA residual corrections g += (x - g^2) h (Markstein, "IA-64 and
// Elementary Functions", 2000). It is the library square root without the input scaling that
// only inputs below 2^-767 need; a GPU check of 1.07e9 inputs over 2^-700 .. 2^700 found the
// same bits everywhere.
__device__ __forceinline__ double sqrtN(double x) {
	const double y = __builtin_amdgcn_rsq(x);
	double g = x * y;
	double h = y * 0.5;
	const double r = fma(-h, g, 0.5);
	g = fma(g, r, g);
	h = fma(h, r, h);
	double d = fma(-g, g, x);
	g = fma(d, h, g);
	d = fma(-g, g, x);
	g = fma(d, h, g);
	return __builtin_amdgcn_class(x, 0x260) ? x : g; // +-0 and +inf are their own roots
}

// Correctly rounded 1 / d and x / d for normal operands of moderate size (|x|, |d| and the
// quotient within 2^-400 .. 2^400): the reciprocal estimate, two Newton steps y += y (1 - d y) and
// one FMA residual correction (Markstein 1990). These are the library divisions without the
// operand scaling and special-case fix-up that only extreme or non-finite operands need; the
// quotient takes q's sign so that 0 / d keeps IEEE's signed zero. A GPU check of 1.07e9 operand
// pairs (exponents -400 .. 400, either sign, an eighth of the numerators +-0) found the same bits.
__device__ __forceinline__ double recipN(double d) {
	double y = __builtin_amdgcn_rcp(d);
	double e = fma(-d, y, 1.0);
	y = fma(y, e, y);
	e = fma(-d, y, 1.0);
	y = fma(y, e, y);
	const double r = fma(-d, y, 1.0);
	return fma(r, y, y);
}
__device__ __forceinline__ double divN(double x, double d) {
	double y = __builtin_amdgcn_rcp(d);
	double e = fma(-d, y, 1.0);
	y = fma(y, e, y);
	e = fma(-d, y, 1.0);
	y = fma(y, e, y);
	const double q = x * y;
	const double r = fma(-d, q, x);
	return copysign(fma(r, y, q), q);
}
// x / d: the short sequence with MK, else the library's IEEE division
template<bool MK>
__device__ __forceinline__ double divK(double x, double d) {
	if constexpr (MK) return divN(x, d);
	else return x / d;
}

// primitives of a conservative state as FlowField::updateWRungeKutta forms them
// (FlowField.cpp:140-158): u = rho u / rho, ..., E = rho E / rho, p = (gamma - 1) rho
// (E - |u|^2 / 2), H = E + p / rho
template<bool MK>
__device__ __forceinline__ void primOf(double r, double ru, double rv, double rw, double rE, double g, double &u,
                                       double &v, double &w, double &pr, double &E, double &H) {
	const double yr = MK ? recipN(r) : 0.0;
	u = quot<MK>(ru, r, yr);
	v = quot<MK>(rv, r, yr);
	w = quot<MK>(rw, r, yr);
	E = quot<MK>(rE, r, yr);
	pr = (g - 1) * r * (E - (u * u + v * v + w * w) * 0.5);
	H = E + quot<MK>(pr, r, yr);
}

// The same Roe flux, component by component: each of the five values is formed with exactly the
// operations of roeFlux and handed to emit(k, F_k) at once, so no 5-vector of temporaries is live.
// With MK the quotients by sL + sR and by 2 c~^2 go through divRecip. The per-cell values
// sL = sqrt(rho_L), sR = sqrt(rho_R) and cL = sqrt(gamma p_L / rho_L) come from the caller.
template<bool MK = false, typename Emit>
__device__ __forceinline__ void roeFluxEmit(const Prim &L, const Prim &R, double sL, double sR, double cL, double nx,
                                            double ny, double nz, double gamma, Emit emit) {
	double rhoT = sqrtN(L.rho * R.rho);
	const double den = sL + sR;
	const double yden = MK ? recipN(den) : 0.0;
	double uT = quot<MK>(L.u * sL + R.u * sR, den, yden);
	double vT = quot<MK>(L.v * sL + R.v * sR, den, yden);
	double wT = quot<MK>(L.w * sL + R.w * sR, den, yden);
	double hT = quot<MK>(L.H * sL + R.H * sR, den, yden);
	double q2 = uT * uT + vT * vT + wT * wT;
	double cT = sqrtN((gamma - 1.0) * (hT - q2 * 0.5));
	double VT = uT * nx + vT * ny + wT * nz;
	double VL = L.u * nx + L.v * ny + L.w * nz;
	double VR = R.u * nx + R.v * ny + R.w * nz;
	double dP = R.p - L.p, dVc = VR - VL;
	double delta = (1.0 / 15.0) * cL;
	double a1 = VT - cT, a5 = VT + cT;
	// Harten's fix on the acoustic waves; the quotients are formed only where a lane needs them
	double l1 = fabs(a1), l5 = fabs(a5);
	if (!(l1 > delta) || !(l5 > delta)) {
		if (!(l1 > delta)) l1 = (a1 * a1 + delta * delta) / (2.0 * delta);
		if (!(l5 > delta)) l5 = (a5 * a5 + delta * delta) / (2.0 * delta);
	}
	const double cc = cT * cT;
	double c2x2 = 2.0 * cc;
	double k1 = dP - rhoT * cT * dVc;
	double k5 = dP + rhoT * cT * dVc;
	// RN(1 / (2 cc)) scaled by 2 is RN(1 / cc) exactly, so dP / cc shares the reciprocal
	const double yc = MK ? recipN(c2x2) : 0.0;
	double kr = (R.rho - L.rho) - quot<MK>(dP, cc, 2.0 * yc);
	double aV = fabs(VT);
	double rVR = R.rho * VR, lVL = L.rho * VL;
	auto comp = [&](int k, double avg, double e1, double e2, double t2, double e5) {
		double f1 = quot<MK>(e1 * l1 * k1, c2x2, yc);
		double f234 = (e2 * kr + t2) * aV;
		double f5 = quot<MK>(e5 * l5 * k5, c2x2, yc);
		emit(k, avg - 0.5 * (f1 + f234 + f5));
	};
	comp(0, (rVR + lVL) * 0.5, 1.0, 1.0, 0.0, 1.0);
	comp(1, ((rVR * R.u + R.p * nx) + (lVL * L.u + L.p * nx)) * 0.5, uT - cT * nx, uT, rhoT * ((R.u - L.u) - dVc * nx),
	     uT + cT * nx);
	comp(2, ((rVR * R.v + R.p * ny) + (lVL * L.v + L.p * ny)) * 0.5, vT - cT * ny, vT, rhoT * ((R.v - L.v) - dVc * ny),
	     vT + cT * ny);
	comp(3, ((rVR * R.w + R.p * nz) + (lVL * L.w + L.p * nz)) * 0.5, wT - cT * nz, wT, rhoT * ((R.w - L.w) - dVc * nz),
	     wT + cT * nz);
	comp(4, (rVR * R.H + lVL * L.H) * 0.5, hT - cT * VT, q2 * 0.5,
	     rhoT * (uT * (R.u - L.u) + vT * (R.v - L.v) + wT * (R.w - L.w) - VT * dVc), hT + cT * VT);
}

// AUSM, Liou & Steffen, J. Comput. Phys. 107 (1993) 23-39 (the reference only declares it,
// ConvectiveFluxesSchemes.h:27-30)
__device__ __forceinline__ double machPlus(double M) { return fabs(M) <= 1.0 ? 0.25 * (M + 1.0) * (M + 1.0) : 0.5 * (M + fabs(M)); }
__device__ __forceinline__ double machMinus(double M) { return fabs(M) <= 1.0 ? -0.25 * (M - 1.0) * (M - 1.0) : 0.5 * (M - fabs(M)); }
__device__ __forceinline__ double presPlus(double M) {
	return fabs(M) <= 1.0 ? 0.25 * (M + 1.0) * (M + 1.0) * (2.0 - M) : (M > 0.0 ? 1.0 : 0.0);
}
__device__ __forceinline__ double presMinus(double M) {
	return fabs(M) <= 1.0 ? 0.25 * (M - 1.0) * (M - 1.0) * (2.0 + M) : (M < 0.0 ? 1.0 : 0.0);
}

// with the sound speeds aL = sqrt(g p_L / rho_L), aR = sqrt(g p_R / rho_R) from the caller
__device__ __forceinline__ void ausmFluxA(const Prim &L, const Prim &R, double aL, double aR, double nx, double ny,
                                          double nz, double g, double out[5]) {
	double VL = L.u * nx + L.v * ny + L.w * nz;
	double VR = R.u * nx + R.v * ny + R.w * nz;
	double ML = VL / aL, MR = VR / aR;
	double mh = machPlus(ML) + machMinus(MR);
	double ph = presPlus(ML) * L.p + presMinus(MR) * R.p;
	double phiL[5] = {L.rho * aL, L.rho * aL * L.u, L.rho * aL * L.v, L.rho * aL * L.w, L.rho * aL * L.H};
	double phiR[5] = {R.rho * aR, R.rho * aR * R.u, R.rho * aR * R.v, R.rho * aR * R.w, R.rho * aR * R.H};
	double pn[5] = {0.0, ph * nx, ph * ny, ph * nz, 0.0};
#pragma unroll
	for (int i = 0; i < 5; i++) out[i] = mh * 0.5 * (phiL[i] + phiR[i]) - 0.5 * fabs(mh) * (phiR[i] - phiL[i]) + pn[i];
}

__device__ __forceinline__ void ausmFlux(const Prim &L, const Prim &R, double nx, double ny, double nz, double g,
                                         double out[5]) {
	ausmFluxA(L, R, sqrt(g * L.p / L.rho), sqrt(g * R.p / R.rho), nx, ny, nz, g, out);
}

template<int SCHEME>
__device__ __forceinline__ void schemeFlux(const Prim &L, const Prim &R, double nx, double ny, double nz, double g,
                                           double out[5]) {
	if (SCHEME == 1) ausmFlux(L, R, nx, ny, nz, g, out);
	else roeFlux(L, R, nx, ny, nz, g, out);
}

// interior Mach number as FlowField::updateWRungeKutta stores it (FlowField.cpp:157, c/|u|)
__device__ __forceinline__ double machOf(double rho, double u, double v, double w, double p, double g) {
	return sqrt(g * (p / rho)) / sqrt((u * u + v * v + w * w));
}

// Wall (PhysicalBC.cpp:124-162) and symmetry (:165-231): mirror the interior velocity
__device__ __forceinline__ Ghost mirrorGhost(double ri, double ui, double vi, double wi, double pi, double nx, double ny,
                                             double nz, double gm) {
	double c = sqrt(gm * (pi / ri));
	double V = ui * nx + vi * ny + wi * nz;
	double u = ui - 2 * V * nx;
	double v = vi - 2 * V * ny;
	double w = wi - 2 * V * nz;
	double M = sqrt(u * u + v * v + w * w) / c;
	double E = pi / ((gm - 1) * ri) + ((u * u + v * v + w * w) * 0.5);
	double H = E + (pi / ri);
	return {ri, u, v, w, pi, E, H, M};
}

// Farfield: branch on the ghost's previous velocity and Mach (EulerSolver.cpp:205-226), then
// PhysicalBC.cpp:11-121. `in` holds the ghost's previous state; I, Ei, Mi are the interior
// cell's primitives, energy and stored Mach number.
__device__ __forceinline__ Ghost farfieldGhost(const DevState &in, int g, double Mi, double nx, double ny, double nz,
                                               const DevFree &q, const Prim &I, double Ei) {
	double gm = q.gamma;
	const double *gc = cellPtr(in, g);
	double Vg = gc[S_U] * nx + gc[S_V] * ny + gc[S_W] * nz;
	double Mg = gc[S_M];
	double ri = I.rho, pi = I.p, ui = I.u, vi = I.v, wi = I.w;
	if (Mg >= 1.0 && Vg > 0.0) return {ri, ui, vi, wi, pi, Ei, I.H, Mi};
	if (Mg >= 1.0 && Vg <= 0.0) return {q.rho_inf, q.u_inf, q.v_inf, q.w_inf, q.p_inf, q.E_inf, q.H_inf, q.M_inf};
	double c = sqrt(gm * (pi / ri));
	double p, rho, u, v, w;
	if (Vg > 0.0) {
		// subsonic outflow, PhysicalBC.cpp:82-121
		p = q.p_inf;
		rho = ri + (p - pi) / (c * c);
		u = ui + nx * (pi - p) / (ri * c);
		v = vi + ny * (pi - p) / (ri * c);
		w = wi + nz * (pi - p) / (ri * c);
	} else {
		// subsonic inflow, PhysicalBC.cpp:43-79
		p = 0.5 * (q.p_inf + pi - ri * c * ((nx * (q.u_inf - ui)) + (ny * (q.v_inf - vi)) + (nz * (q.w_inf - wi))));
		rho = q.rho_inf + (p - q.p_inf) / (c * c);
		u = q.u_inf - nx * (q.p_inf - p) / (ri * c);
		v = q.v_inf - ny * (q.p_inf - p) / (ri * c);
		w = q.w_inf - nz * (q.p_inf - p) / (ri * c);
	}
	double M = sqrt(u * u + v * v + w * w) / c;
	double E = p / ((gm - 1) * rho) + ((u * u + v * v + w * w) * 0.5);
	double H = E + (p / rho);
	return {rho, u, v, w, p, E, H, M};
}

// EulerSolver::updateBC (EulerSolver.cpp:201-247), one thread per ghost, in place
__global__ void __launch_bounds__(kBlock) kBc(DevMesh m, DevState s, DevFree q, int machDerived) {
	int k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= m.nGhost || m.gkind[k] == 3) return; // an interface face reads the other partition's cell
	int f = m.nFaceInt + k;
	int g = m.nElem + k;
	int i = m.f.e0[f];
	double nx = m.f.nx[f], ny = m.f.ny[f], nz = m.f.nz[f];
	const double *ic = cellPtr(s, i);
	Prim I = loadPrim(s, i);
	Ghost G;
	if (m.gkind[k] == 0) {
		double Mi = machDerived ? machOf(I.rho, I.u, I.v, I.w, I.p, q.gamma) : ic[S_M];
		G = farfieldGhost(s, g, Mi, nx, ny, nz, q, I, ic[S_E]);
	} else {
		G = mirrorGhost(I.rho, I.u, I.v, I.w, I.p, nx, ny, nz, q.gamma);
	}
	storeGhost(s, g, G);
}

// per-face flux as EulerSolver::computeResidual dispatches it (EulerSolver.cpp:306-358)
__global__ void __launch_bounds__(kBlock) kFlux(DevMesh m, DevState s, DevFree q, int scheme, double *fs, double *raw) {
	int f = blockIdx.x * kBlock + threadIdx.x;
	if (f >= m.nFace) return;
	int a = m.f.e0[f], b = m.f.e1[f];
	double nx = m.f.nx[f], ny = m.f.ny[f], nz = m.f.nz[f];
	double F[5];
	int kind = f < m.nFaceInt ? 3 : (int) m.gkind[f - m.nFaceInt];
	Prim R = loadPrim(s, b);
	if (kind == 3 || kind == 2) {
		Prim L = loadPrim(s, a);
		if (scheme == 1) ausmFlux(L, R, nx, ny, nz, q.gamma, F);
		else roeFlux(L, R, nx, ny, nz, q.gamma, F);
	} else if (kind == 1) {
		F[0] = 0;
		F[1] = nx * R.p;
		F[2] = ny * R.p;
		F[3] = nz * R.p;
		F[4] = 0;
	} else {
		// Solver::Fc with V = ghost velocity . n (ConvectiveFluxesSchemes.cpp:167-187)
		double V = R.u * nx + R.v * ny + R.w * nz;
		F[0] = R.rho * V;
		F[1] = R.rho * R.u * V + R.p * nx;
		F[2] = R.rho * R.v * V + R.p * ny;
		F[3] = R.rho * R.w * V + R.p * nz;
		F[4] = R.rho * R.H * V;
	}
	double S = m.f.S[f];
	size_t nF = (size_t) m.nFace;
#pragma unroll
	for (int k = 0; k < 5; k++) fs[k * nF + f] = F[k] * S;
	if (raw) {
#pragma unroll
		for (int k = 0; k < 5; k++) raw[k * nF + f] = F[k];
	}
}

__device__ __forceinline__ void gatherCell(const DevMesh &m, const double *fs, int e, double r[5]) {
	size_t nF = (size_t) m.nFace;
#pragma unroll
	for (int k = 0; k < 5; k++) r[k] = 0.0;
	int b = m.c.fstart[e], en = m.c.fstart[e + 1];
	for (int j = b; j < en; j++) {
		int code = m.c.flist[j];
		int f = code & 0x7fffffff;
		if (code < 0) {
#pragma unroll
			for (int k = 0; k < 5; k++) r[k] -= fs[k * nF + f];
		} else {
#pragma unroll
			for (int k = 0; k < 5; k++) r[k] += fs[k * nF + f];
		}
	}
}

__global__ void __launch_bounds__(kBlock) kGather(DevMesh m, const double *fs, double *R) {
	int e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= m.nElem) return;
	double r[5];
	gatherCell(m, fs, e, r);
	size_t n = (size_t) m.nElem;
#pragma unroll
	for (int k = 0; k < 5; k++) R[k * n + e] = r[k];
}

// TimeStepMethod2, TimeIntegration.cpp:8-40 (sum|N| over the cell's faces is static: precomputed
// on the host in the same local-face order)
template<bool NT = false>
__device__ __forceinline__ double localDt(const DevMesh &m, const DevFree &q, double rho, double u, double v, double w,
                                          double p, int e) {
	double c = sqrt(q.gamma * p / rho);
	double lx = 0.5 * (fabs(u) + c) * ldStream<NT>(m.c.sx + e);
	double ly = 0.5 * (fabs(v) + c) * ldStream<NT>(m.c.sy + e);
	double lz = 0.5 * (fabs(w) + c) * ldStream<NT>(m.c.sz + e);
	return q.cfl * ldStream<NT>(m.c.vol + e) / (lx + ly + lz);
}

__global__ void __launch_bounds__(kBlock) kTimestep(DevMesh m, DevState s, DevFree q, double *dt) {
	int e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= m.nElem) return;
	Prim P = loadPrim(s, e);
	dt[e] = localDt(m, q, P.rho, P.u, P.v, P.w, P.p, e);
}

// deterministic reductions of one double per thread: wavefront butterflies, then the wave totals
// in order (reduce.hpp)
template<int N>
__device__ __forceinline__ double blockSumN(double x, double *lds) { return groupSum<N>(x, lds); }
__device__ __forceinline__ double blockSum(double x, double *lds) { return groupSum<kBlock>(x, lds); }

// RK stage update of one cell from its residual r: EulerSolver.cpp:431-449 / 459-470 and
// FlowField.cpp:140-158. (rho, u, v, w, p, E) is the cell's state before the stage; writes `out`
// (the Mach slot is not written: it is recomputed from the primitives when needed).
template<bool NT = false>
__device__ __forceinline__ void updateCell(const DevMesh &m, const DevState &out, const DevFree &q, double *W0,
                                           double alpha, int firstStage, int e, const double r[5], double rho, double u,
                                           double v, double w, double p, double E, int integ = 0) {
	size_t n = (size_t) m.nElem;
	double w0[5];
	if (integ == 1) {
		// explicit Euler (TimeIntegration.cpp:78-91, FlowField.cpp:103-126), W0 = persistent W
		double dt = localDt<NT>(m, q, rho, u, v, w, p, e);
		double coeff = -dt / ldStream<NT>(m.c.vol + e);
		double nr = rho + r[0] * coeff;
		double nru = W0[n + e] + r[1] * coeff;
		double nrv = W0[2 * n + e] + r[2] * coeff;
		double nrw = W0[3 * n + e] + r[3] * coeff;
		double nrE = W0[4 * n + e] + r[4] * coeff;
		W0[e] = nr, W0[n + e] = nru, W0[2 * n + e] = nrv, W0[3 * n + e] = nrw, W0[4 * n + e] = nrE;
		double nu = nru / nr, nv = nrv / nr, nw = nrw / nr, nE = nrE / nr;
		double g = q.gamma;
		double np = (g - 1) * nr * (nE - (nu * nu + nv * nv + nw * nw) * 0.5);
		double nH = nE + (np / nr);
		double2 *c = reinterpret_cast<double2 *>(cellPtr(out, e));
		c[0] = make_double2(nr, nu);
		c[1] = make_double2(nv, nw);
		c[2] = make_double2(np, nE);
		c[3].x = nH;
		return;
	}
	if (firstStage) {
		w0[0] = rho;
		w0[1] = u * rho;
		w0[2] = v * rho;
		w0[3] = w * rho;
		w0[4] = E * rho;
#pragma unroll
		for (int k = 0; k < 5; k++) W0[k * n + e] = w0[k];
	} else {
#pragma unroll
		for (int k = 0; k < 5; k++) w0[k] = ldStream<NT>(W0 + k * n + e);
	}
	double dt = localDt<NT>(m, q, rho, u, v, w, p, e);
	double V = ldStream<NT>(m.c.vol + e);
	double nr = w0[0] - alpha * dt * r[0] / V;
	double nru = w0[1] - alpha * dt * r[1] / V;
	double nrv = w0[2] - alpha * dt * r[2] / V;
	double nrw = w0[3] - alpha * dt * r[3] / V;
	double nrE = w0[4] - alpha * dt * r[4] / V;
	double nu = nru / nr, nv = nrv / nr, nw = nrw / nr, nE = nrE / nr;
	double g = q.gamma;
	double np = (g - 1) * nr * (nE - (nu * nu + nv * nv + nw * nw) * 0.5);
	double nH = nE + (np / nr);
	double2 *c = reinterpret_cast<double2 *>(cellPtr(out, e));
	c[0] = make_double2(nr, nu);
	c[1] = make_double2(nv, nw);
	c[2] = make_double2(np, nE);
	c[3].x = nH;
}

// integ 0: RK stage, 1: explicit Euler step, 2: no update (TIME_INTEGRATION= 1 only recomputes
// dt, EulerSolver.cpp:474-480). The residual is gathered from fs, or read from Rin when given.
__global__ void __launch_bounds__(kBlock) kStage(DevMesh m, DevState s, DevFree q, const double *fs, const double *Rin,
                                                 double *W0, double alpha, int firstStage, int integ,
                                                 double *rmsPartial, double *Rout) {
	__shared__ double lds[kBlock];
	int e = blockIdx.x * kBlock + threadIdx.x;
	double r2 = 0.0;
	if (e < m.nElem) {
		size_t n = (size_t) m.nElem;
		double r[5];
		if (Rin) {
#pragma unroll
			for (int k = 0; k < 5; k++) r[k] = Rin[k * n + e];
		} else {
			gatherCell(m, fs, e, r);
		}
		if (Rout) {
#pragma unroll
			for (int k = 0; k < 5; k++) Rout[k * n + e] = r[k];
		}
		if (integ != 2) {
			const double *c = cellPtr(s, e);
			updateCell(m, s, q, W0, alpha, firstStage, e, r, c[S_RHO], c[S_U], c[S_V], c[S_W], c[S_P], c[S_E], integ);
		}
		r2 = r[0] * r[0];
	}
	if (rmsPartial) {
		double t = blockSum(r2, lds);
		if (threadIdx.x == 0) rmsPartial[blockIdx.x] = t;
	}
}

// EulerSolver::smoothResiduals (EulerSolver.cpp:483-518) as it runs: one Jacobi sweep
// R1_e = (R0_e + sum over interior neighbours s of R0_s * 0.25) / (1 + 0.25 * faces), neighbours in
// local-face order; a second sweep from R1 follows when the first changed any value by more than
// 1e-16 (the loop's own test scans an emptied vector from its second pass on, so it ends there).
// nbr: per cell (device ids) the neighbour of each local face, -2 for a ghost (counted, not summed),
// -1 padding; nface: faces per cell.
__device__ __forceinline__ void smoothCell(const double *R0, const double *last, const int *nbr, const uint8_t *nface,
                                           int stride, size_t n, int e, double out[5]) {
	double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
	const int *nb = nbr + (size_t) e * stride;
	for (int i = 0; i < stride; i++) {
		const int s = nb[i];
		if (s == -1) break;
		if (s >= 0) {
#pragma unroll
			for (int k = 0; k < 5; k++) sum[k] += last[k * n + s] * 0.25;
		}
	}
	const double den = 1 + nface[e] * 0.25;
#pragma unroll
	for (int k = 0; k < 5; k++) out[k] = (R0[k * n + e] + sum[k]) / den;
}

// first sweep: R1 and the largest change per block (NaN when a change is NaN), or per cell into
// cellMax when the cells belong to several partitions (each decides on its own)
__global__ void __launch_bounds__(kBlock) kSmooth1(const double *R0, double *R1, const int *nbr, const uint8_t *nface,
                                                   int stride, int nElem, double *blockMax, double *cellMax) {
	__shared__ double lds[kBlock];
	const int e = blockIdx.x * kBlock + threadIdx.x;
	const size_t n = (size_t) nElem;
	double d = 0.0;
	if (e < nElem) {
		double o[5];
		smoothCell(R0, R0, nbr, nface, stride, n, e, o);
#pragma unroll
		for (int k = 0; k < 5; k++) {
			R1[k * n + e] = o[k];
			const double c = fabs(R0[k * n + e] - o[k]);
			d = (c > d || c != c) ? c : d;
		}
		if (cellMax) cellMax[e] = d;
	}
	if (cellMax) return;
	lds[threadIdx.x] = d;
	__syncthreads();
	for (int off = kBlock / 2; off > 0; off >>= 1) {
		if ((int) threadIdx.x < off) {
			const double a = lds[threadIdx.x], b = lds[threadIdx.x + off];
			lds[threadIdx.x] = (b > a || b != b) ? b : a;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) blockMax[blockIdx.x] = lds[0];
}

// per-partition decisions (one workgroup per partition, over its contiguous cells): flag[r] = 1
// when partition r's first sweep changed something by more than 1e-16; flag[nPart] = 1 on a NaN
__global__ void __launch_bounds__(kBlock) kSmoothDecideParts(const double *cellMax, const int *partStart, int nPart,
                                                             int *flag) {
	__shared__ double lds[kBlock];
	const int r = blockIdx.x;
	double d = 0.0;
	for (int i = partStart[r] + (int) threadIdx.x; i < partStart[r + 1]; i += kBlock) {
		const double c = cellMax[i];
		d = (c > d || c != c) ? c : d;
	}
	lds[threadIdx.x] = d;
	__syncthreads();
	for (int off = kBlock / 2; off > 0; off >>= 1) {
		if ((int) threadIdx.x < off) {
			const double a = lds[threadIdx.x], b = lds[threadIdx.x + off];
			lds[threadIdx.x] = (b > a || b != b) ? b : a;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		const double err = lds[0];
		flag[r] = err > 1e-16 ? 1 : 0;
		if (err != err) atomicOr(flag + nPart, 1);
	}
}

// flag[0] = 1 when the first sweep changed something by more than 1e-16, flag[1] = 1 on a NaN change
__global__ void __launch_bounds__(kBlock) kSmoothDecide(const double *blockMax, int nb, int *flag) {
	__shared__ double lds[kBlock];
	double d = 0.0;
	for (int i = threadIdx.x; i < nb; i += kBlock) {
		const double c = blockMax[i];
		d = (c > d || c != c) ? c : d;
	}
	lds[threadIdx.x] = d;
	__syncthreads();
	for (int off = kBlock / 2; off > 0; off >>= 1) {
		if ((int) threadIdx.x < off) {
			const double a = lds[threadIdx.x], b = lds[threadIdx.x + off];
			lds[threadIdx.x] = (b > a || b != b) ? b : a;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		const double err = lds[0];
		flag[0] = err > 1e-16 ? 1 : 0;
		flag[1] = flag[1] | (err != err ? 1 : 0);
	}
}

// second sweep from R1 into R2 when the cell's partition decided so (flag[partition]), else R2 = R1
__global__ void __launch_bounds__(kBlock) kSmooth2(const double *R0, const double *R1, double *R2, const int *nbr,
                                                   const uint8_t *nface, int stride, int nElem, const int *flag,
                                                   const int *partStart, int nPart) {
	const int e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= nElem) return;
	const size_t n = (size_t) nElem;
	int r = 0;
	for (int lo = 0, hi = nPart - 1; lo < hi;) { // last r with partStart[r] <= e
		const int mid = (lo + hi + 1) >> 1;
		if (partStart[mid] <= e) lo = mid;
		else hi = mid - 1;
		r = lo;
	}
	if (flag[r]) {
		double o[5];
		smoothCell(R0, R1, nbr, nface, stride, n, e, o);
#pragma unroll
		for (int k = 0; k < 5; k++) R2[k * n + e] = o[k];
	} else {
#pragma unroll
		for (int k = 0; k < 5; k++) R2[k * n + e] = R1[k * n + e];
	}
}

// Workgroups are dispatched round-robin over the XCDs (each with its own L2). With xcds > 1 the
// dispatch index is remapped so that every XCD runs a contiguous range of blocks: neighbouring
// blocks, which share cut faces and read each other's cells, then share an L2. With reverse set,
// each XCD walks its range from the end.
__device__ __forceinline__ int logicalBlock(int w, int nBlocks, int xcds, bool reverse) {
	if (xcds <= 1) return reverse ? nBlocks - 1 - w : w;
	const int x = w % xcds, k = w / xcds;
	const int q = nBlocks / xcds, r = nBlocks % xcds;
	const int len = q + (x < r ? 1 : 0); // dispatch indices w with w % xcds == x
	return x * q + min(x, r) + (reverse ? len - 1 - k : k);
}

// Fused RK stage over one block of cells (see the file comment). Every global load of the block
// is issued before any flux is computed, so a block waits for memory once: the block's own cells
// and its halo (the outside cells its faces read) are staged in LDS as rho, u, v, w, p, H; the
// update's inputs (W0, V, sum|S|, face slots) are loaded into registers. The flux loop then reads
// LDS only, except for boundary faces, whose ghost cells live in global memory. LAYOUT says where
// the interior state is read from and written to (0: primitive records in and out, 1: primitive
// records in and conservative states out, 2: conservative states in and out); ghost records are
// read from `in` and written to `out` in every layout.
// LDS strides of the hot configuration (256-cell blocks, conservative states in and out, halo
// and slot counts that fit): per-cell rows of kFixCap doubles and flux rows of kFixSlots. As
// compile-time constants every LDS address of a row is an immediate offset from one lane address.
// Both are odd, so the rows of one cell (and the five fluxes of one slot) sit in different banks.
constexpr int kFixCap = 441, kFixSlots = 637;

// NTH threads serve the CB cells of a block (NTH = CB, or 2 or 3 CB where a block's LDS leaves room
// for one workgroup per CU only: more waves then share the same LDS); threads t >= CB take part in
// the loads and the flux loop, and the gather and update run on the first CB. With LATE the
// update's own inputs (W0, metrics, CSR codes) are loaded after the flux loop instead of with the
// block's state, so that they hold no registers during it: with NTH = 5 CB / 4 the kernel then
// fits 128 registers, and three workgroups per CU run 15 waves instead of 12.
template<int SCHEME, bool NT, bool MK, int INTEG, int LAYOUT, int CB, bool FIXED = false, int NTH = CB, bool LATE = false>
__global__ void __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(LATE ? 4 : 1))) kStageFused(DevMesh m, DevPlan p, DevState in, DevState out,
                                                      const double *Win, double *Wout, DevFree q, double *W0,
                                                      double alpha, int firstStage, int machDerived,
                                                      double *rmsPartial, int xcds, int reverse) {
	extern __shared__ double lds[];
	const int b = logicalBlock(blockIdx.x, p.nBlocks, xcds, reverse != 0);
	const int t = threadIdx.x;
	const int c0 = b * CB;
	const int nc = min(CB, m.nElem - c0);
	const int s0 = p.blockSlot[b], ns = p.blockSlot[b + 1] - s0;
	const int h0 = p.blockHalo[b], nh = p.blockHalo[b + 1] - h0;
	const int cap = FIXED ? kFixCap : CB + p.maxHalo;
	const int ms = FIXED ? kFixSlots : p.maxSlots;
	// LDS: [6][cap] rho, u, v, w, p, H of own cells then halo cells; with LAYOUT 2 two more rows
	// [cap], sqrt(rho) and the sound speed sqrt(gamma p / rho) as the fluxes and the time step form
	// them, else [2][CB] E, M of own cells; then [5][ms] F*S per slot (first the own W of LAYOUT 2)
	double *ls = lds;
	double *lsq = ls + 6 * cap, *lcs = lsq + cap;
	double *lem = ls + 6 * cap;
	double *lf = LAYOUT == 2 ? lcs + cap : lem + 2 * CB;
	double Eown = 0.0; // LAYOUT 2: E of this thread's cell
	// LAYOUT 2: E of own cell l from its conservative state, as primOf forms it
	auto ownE = [&](int l) -> double {
		const double *w5 = Win + (size_t) 5 * (c0 + l);
		const double r = w5[0];
		return quot<MK>(w5[4], r, MK ? recipN(r) : 0.0);
	};
	const size_t n = (size_t) m.nElem;

	// ---- issue every load of the block
	const int e = c0 + t;
	const bool own = t < nc;
	// own records: the block's consecutive 64 B records are read as consecutive 16 B chunks,
	// thread t taking chunks t, t + NTH, ..., so that every wave instruction covers whole lines;
	// chunk j holds field pair j & 3 (rho u | v w | p E | H M) of cell j >> 2
	constexpr int kRec = (4 * CB + NTH - 1) / NTH, kCons = (5 * CB / 2 + NTH - 1) / NTH;
	double2 oc4[4];
	if constexpr (LAYOUT < 2) {
		const double2 *blk = reinterpret_cast<const double2 *>(cellPtr(in, c0));
#pragma unroll
		for (int i = 0; i < kRec; i++) {
			const int j = i * NTH + t;
			oc4[i] = j < 4 * nc ? blk[j] : make_double2(0, 0);
		}
	} else {
		// own conservative states: 5 doubles per cell, the block's range (16 B aligned) read as
		// consecutive 16 B chunks; chunk j holds doubles 2j and 2j + 1 of the range (the last
		// chunk of the last block may reach one double into the buffer's padding)
		const double2 *blk = reinterpret_cast<const double2 *>(Win + (size_t) 5 * c0);
#pragma unroll
		for (int i = 0; i < kCons; i++) {
			const int j = i * NTH + t;
			oc4[i] = 2 * j < 5 * nc ? blk[j] : make_double2(0, 0);
		}
	}
	double w0[5], vol = 0, sx = 0, sy = 0, sz = 0;
	uint16_t code[6];
	auto loadUpdateInputs = [&]() {
		if (!own) return;
		if (INTEG == 1 || !firstStage) {
#pragma unroll
			for (int k = 0; k < 5; k++) w0[k] = ldStream<NT>(W0 + k * n + e);
		}
		vol = ldStream<NT>(m.c.vol + e);
		sx = ldStream<NT>(m.c.sx + e);
		sy = ldStream<NT>(m.c.sy + e);
		sz = ldStream<NT>(m.c.sz + e);
		const uint16_t *cl = p.csr + (size_t) e * p.csrStride;
		if (p.csrStride == 4) {
			// tetrahedra: the cell's four codes as one 8 B load (rows are 8 B aligned)
			const uint64_t c4 = ldStream<NT>(reinterpret_cast<const uint64_t *>(cl));
#pragma unroll
			for (int i = 0; i < 4; i++) code[i] = (uint16_t) (c4 >> (16 * i));
			code[4] = code[5] = 0xFFFF;
		} else {
#pragma unroll
			for (int i = 0; i < 6; i++) code[i] = i < p.csrStride ? ldStream<NT>(cl + i) : (uint16_t) 0xFFFF;
		}
	};
	if constexpr (!LATE) loadUpdateInputs();
	for (int h = t; h < nh; h += NTH) {
		const int l = CB + h;
		if constexpr (LAYOUT < 2) {
			const double2 *hc = reinterpret_cast<const double2 *>(cellPtr(in, p.haloCell[h0 + h]));
			const double2 a0 = hc[0], a1 = hc[1], a2 = hc[2], a3 = hc[3];
			ls[0 * cap + l] = a0.x;
			ls[1 * cap + l] = a0.y;
			ls[2 * cap + l] = a1.x;
			ls[3 * cap + l] = a1.y;
			ls[4 * cap + l] = a2.x;
			ls[5 * cap + l] = a3.x;
		} else {
			const double *hw = Win + (size_t) 5 * p.haloCell[h0 + h];
			const double r = hw[0];
			double u, v, w, pr, E, H;
			primOf<MK>(r, hw[1], hw[2], hw[3], hw[4], q.gamma, u, v, w, pr, E, H);
			ls[0 * cap + l] = r;
			ls[1 * cap + l] = u;
			ls[2 * cap + l] = v;
			ls[3 * cap + l] = w;
			ls[4 * cap + l] = pr;
			ls[5 * cap + l] = H;
			lsq[l] = sqrtN(r);
			lcs[l] = sqrtN(divK<MK>(q.gamma * pr, r));
		}
	}
	if constexpr (LAYOUT == 2) {
		// the own W range as loaded, into LDS; turned into primitives after the barrier
#pragma unroll
		for (int i = 0; i < kCons; i++) {
			const int j = i * NTH + t;
			if (2 * j < 5 * nc) {
				lf[2 * j] = oc4[i].x;
				lf[2 * j + 1] = oc4[i].y;
			}
		}
	} else {
		// pair k = t & 3 of every chunk this thread loaded: rho|v|p|H into rows 0|2|4|5, the
		// second member into rows 1|3 (u, w) or the own-cell E|M rows
		const int k = t & 3;
		double *dA = ls + (k < 2 ? 2 * k : 2 + k) * cap;
		double *dB = k < 2 ? ls + (2 * k + 1) * cap : lem + (k - 2) * CB;
#pragma unroll
		for (int i = 0; i < kRec; i++) {
			const int j = i * NTH + t;
			if (j < 4 * nc) {
				dA[j >> 2] = oc4[i].x;
				dB[j >> 2] = oc4[i].y;
			}
		}
	}
	// face slot record of this thread's first round (later rounds are prefetched one ahead)
	uint32_t pkN = 0;
	double nxN = 0, nyN = 0, nzN = 0, SN = 0;
	if (t < ns) {
		pkN = ldStream<NT>(p.pack + s0 + t);
		nxN = ldStream<NT>(p.nx + s0 + t), nyN = ldStream<NT>(p.ny + s0 + t), nzN = ldStream<NT>(p.nz + s0 + t);
		SN = ldStream<NT>(p.S + s0 + t);
	}
	__syncthreads();
	if constexpr (LAYOUT == 2) {
		if (own) {
			const double *wt = lf + 5 * t;
			const double r = wt[0];
			double u, v, w, pr, E, H;
			primOf<MK>(r, wt[1], wt[2], wt[3], wt[4], q.gamma, u, v, w, pr, E, H);
			ls[t] = r;
			ls[cap + t] = u;
			ls[2 * cap + t] = v;
			ls[3 * cap + t] = w;
			ls[4 * cap + t] = pr;
			ls[5 * cap + t] = H;
			lsq[t] = sqrtN(r);
			lcs[t] = sqrtN(divK<MK>(q.gamma * pr, r));
			Eown = E;
		}
		__syncthreads();
	}

	// ---- face fluxes of the block, once each, into LDS
	auto lprim = [&](int l) -> Prim {
		return {ls[l], ls[cap + l], ls[2 * cap + l], ls[3 * cap + l], ls[4 * cap + l], ls[5 * cap + l]};
	};
	for (int j = t; j < ns; j += NTH) {
		const int s = s0 + j;
		const uint32_t pk = pkN;
		const double nx = nxN, ny = nyN, nz = nzN, S = SN;
		if (j + NTH < ns) {
			pkN = ldStream<NT>(p.pack + s + NTH);
			nxN = ldStream<NT>(p.nx + s + NTH), nyN = ldStream<NT>(p.ny + s + NTH);
			nzN = ldStream<NT>(p.nz + s + NTH), SN = ldStream<NT>(p.S + s + NTH);
		}
		const int l0 = (int) (pk & 0x7FFu), l1 = (int) ((pk >> 11) & 0x7FFu), kind = (int) (pk >> 22);
		const Prim L = lprim(l0);
		auto emit = [&](int k, double f) { lf[k * ms + j] = f * S; };
		if (SCHEME == 0 && kind == 0) {
			const Prim R = lprim(l1);
			double sL, sR, cL;
			if constexpr (LAYOUT == 2) {
				sL = lsq[l0], sR = lsq[l1], cL = lcs[l0];
			} else {
				sL = sqrt(L.rho), sR = sqrt(R.rho), cL = sqrt(q.gamma * L.p / L.rho);
			}
			roeFluxEmit<MK>(L, R, sL, sR, cL, nx, ny, nz, q.gamma, emit);
			continue;
		}
		double F[5];
		if (kind == 0) {
			if constexpr (SCHEME == 1 && LAYOUT == 2) ausmFluxA(L, lprim(l1), lcs[l0], lcs[l1], nx, ny, nz, q.gamma, F);
			else schemeFlux<SCHEME>(L, lprim(l1), nx, ny, nz, q.gamma, F);
		} else if (kind == 4) {
			// partition interface read from its ghost, which the interface exchange filled
			// (EulerSolver.cpp:343-346: the scheme's flux with the ghost, to the own cell only)
			schemeFlux<SCHEME>(L, loadPrim(in, p.e1[s]), nx, ny, nz, q.gamma, F);
		} else {
			const int a = c0 + l0;    // the interior cell of a boundary face is an own cell
			const int c = p.e1[s];    // its ghost
			Ghost G;
			if (kind == 3) {
				double Mi = (LAYOUT == 2 || machDerived) ? machOf(L.rho, L.u, L.v, L.w, L.p, q.gamma) : lem[CB + l0];
				G = farfieldGhost(in, c, Mi, nx, ny, nz, q, L, LAYOUT == 2 ? ownE(l0) : lem[l0]);
			} else {
				G = mirrorGhost(L.rho, L.u, L.v, L.w, L.p, nx, ny, nz, q.gamma);
			}
			// a farfield ghost's state is read by the next stage's classifier; wall and symmetry
			// ghosts are rebuilt from their cell every stage, so only the iteration's last stage
			// (the one that forms the RMS partials) stores them, as updateBC leaves them
			if (kind == 3 || rmsPartial) storeGhost(out, c, G);
			if (kind == 1) {
				F[0] = 0;
				F[1] = nx * G.p;
				F[2] = ny * G.p;
				F[3] = nz * G.p;
				F[4] = 0;
			} else if (kind == 2) {
				schemeFlux<SCHEME>(L, Prim{G.rho, G.u, G.v, G.w, G.p, G.H}, nx, ny, nz, q.gamma, F);
			} else {
				double V = G.u * nx + G.v * ny + G.w * nz;
				F[0] = G.rho * V;
				F[1] = G.rho * G.u * V + G.p * nx;
				F[2] = G.rho * G.v * V + G.p * ny;
				F[3] = G.rho * G.w * V + G.p * nz;
				F[4] = G.rho * G.H * V;
			}
		}
#pragma unroll
		for (int k = 0; k < 5; k++) emit(k, F[k]);
	}
	if constexpr (LATE) loadUpdateInputs();
	__syncthreads();

	// ---- one thread per cell: gather in reference order, local dt, RK update
	double r2 = 0.0;
	if (own) {
		double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
		for (int i = 0; i < 6; i++) {
			const unsigned cd = code[i];
			if (cd == 0xFFFFu) break;
			const int j = (int) (cd & 0x7FFFu);
			if (cd & 0x8000u) {
#pragma unroll
				for (int k = 0; k < 5; k++) r[k] -= lf[k * ms + j];
			} else {
#pragma unroll
				for (int k = 0; k < 5; k++) r[k] += lf[k * ms + j];
			}
		}
		const double rho = ls[t], u = ls[cap + t], v = ls[2 * cap + t], w = ls[3 * cap + t], pp = ls[4 * cap + t];
		if (INTEG == 0 && firstStage) {
			// W0 = (rho, u*rho, v*rho, w*rho, E*rho), EulerSolver.cpp:431-438
			w0[0] = rho;
			w0[1] = u * rho;
			w0[2] = v * rho;
			w0[3] = w * rho;
			w0[4] = (LAYOUT == 2 ? Eown : lem[t]) * rho;
#pragma unroll
			for (int k = 0; k < 5; k++) W0[k * n + e] = w0[k];
		}
		// TimeStepMethod2 (TimeIntegration.cpp:8-40) and FlowField::updateWRungeKutta
		// (FlowField.cpp:140-158), as updateCell
		double c = LAYOUT == 2 ? lcs[t] : sqrt(q.gamma * pp / rho);
		double lx = 0.5 * (fabs(u) + c) * sx;
		double ly = 0.5 * (fabs(v) + c) * sy;
		double lz = 0.5 * (fabs(w) + c) * sz;
		double dt = divK<MK>(q.cfl * vol, lx + ly + lz);
		double nr, nru, nrv, nrw, nrE;
		if (INTEG == 1) {
			// explicit Euler: dW = R * (-dt / V) (TimeIntegration.cpp:78-91), FlowField::Update
			// (FlowField.cpp:103-126): rho += dW.rho, (rho u, rho v, rho w, rho E) += dW
			const double coeff = -dt / vol;
			nr = rho + r[0] * coeff;
			nru = w0[1] + r[1] * coeff;
			nrv = w0[2] + r[2] * coeff;
			nrw = w0[3] + r[3] * coeff;
			nrE = w0[4] + r[4] * coeff;
			W0[e] = nr;
			W0[n + e] = nru;
			W0[2 * n + e] = nrv;
			W0[3 * n + e] = nrw;
			W0[4 * n + e] = nrE;
		} else {
			const double yv = MK ? recipN(vol) : 0.0;
			nr = w0[0] - quot<MK>(alpha * dt * r[0], vol, yv);
			nru = w0[1] - quot<MK>(alpha * dt * r[1], vol, yv);
			nrv = w0[2] - quot<MK>(alpha * dt * r[2], vol, yv);
			nrw = w0[3] - quot<MK>(alpha * dt * r[3], vol, yv);
			nrE = w0[4] - quot<MK>(alpha * dt * r[4], vol, yv);
		}
		if constexpr (LAYOUT == 0) {
			double nu, nv, nw, np, nE, nH;
			primOf<MK>(nr, nru, nrv, nrw, nrE, q.gamma, nu, nv, nw, np, nE, nH);
			double2 *oc = reinterpret_cast<double2 *>(cellPtr(out, e));
			oc[0] = make_double2(nr, nu);
			oc[1] = make_double2(nv, nw);
			oc[2] = make_double2(np, nE);
			// the whole 64 B record is written (the stored Mach number carried over), so no line
			// is left partly written
			oc[3] = make_double2(nH, lem[CB + t]);
		} else {
			// the conservative state only: the next stage (or the read-back) forms the primitives
			// from it with the same expressions
			double *ow = Wout + (size_t) 5 * e;
			ow[0] = nr;
			ow[1] = nru;
			ow[2] = nrv;
			ow[3] = nrw;
			ow[4] = nrE;
		}
		r2 = r[0] * r[0];
	}
	if (rmsPartial) {
		__syncthreads();
		double tt = blockSumN<NTH>(r2, lds);
		if (t == 0) rmsPartial[b] = tt;
	}
}

// Persistent, software-pipelined form of the fused RK stage (conservative states in and out, RK
// stages): a grid of as many workgroups as fit on the device at once, each walking a list of
// blocks, with the loads of the blocks ahead in flight in registers while a block computes (see
// the loop below). It serves plans whose blocks leave LDS for one workgroup per CU, where no other
// workgroup hides a block's load phase. The arithmetic of a block is kStageFused's, statement for
// statement, so the results are the same bits. KH: halo cells per thread held in flight (blocks' halos <= KH * NTH); KS: rounds of
// slot records held (blocks' slots <= KS * NTH). A workgroup's blocks are its XCD's range of the
// block order (as logicalBlock lays it out), strided by the XCD's workgroup count, backwards when
// `reverse` is set; gridDim.x is a multiple of xcds.
template<int SCHEME, bool FIXED, int NTH, int KH, int KS>
__global__ void __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(2)))
kStagePipe(DevMesh m, DevPlan p, DevState in, DevState out, const double *Win, double *Wout, DevFree q, double *W0,
           double alpha, int firstStage, double *rmsPartial, int xcds, int reverse) {
	constexpr int CB = 256;
	constexpr bool NT = true, MK = true;
	constexpr int kCons = (5 * CB / 2 + NTH - 1) / NTH;
	extern __shared__ double lds[];
	const int t = threadIdx.x;
	const int cap = FIXED ? kFixCap : CB + p.maxHalo;
	const int ms = FIXED ? kFixSlots : p.maxSlots;
	// LDS as kStageFused with conservative states: [6][cap] rho, u, v, w, p, H; [cap] sqrt(rho);
	// [cap] sound speed; [5][ms] F*S per slot, which first stages the block's own W
	double *ls = lds, *lsq = ls + 6 * cap, *lcs = lsq + cap, *lf = lcs + cap;
	const size_t n = (size_t) m.nElem;

	// this workgroup's blocks
	const int X = xcds > 1 ? xcds : 1;
	const int x = (int) blockIdx.x % X, k0 = (int) blockIdx.x / X, perX = (int) gridDim.x / X;
	const int qq = p.nBlocks / X, rr = p.nBlocks % X;
	const int len = qq + (x < rr ? 1 : 0), lo = x * qq + min(x, rr);
	if (k0 >= perX) return; // a grid that is not a multiple of xcds: the surplus workgroups idle
	auto blockAt = [&](int i) -> int {
		const int j = k0 + i * perX;
		if (j >= len) return -1;
		return reverse ? lo + len - 1 - j : lo + j;
	};

	// Loads in flight, three blocks deep. While block i computes: the own and halo states of block
	// i + 1 (issued before i's flux loop, from halo ids loaded during block i - 1), the halo ids of
	// block i + 2, the meta of block i + 3, and, after i's flux loop, block i + 1's slot records,
	// W0, metrics and CSR codes (which its flux loop and update need later). No load waits for
	// another load issued in the same block.
	double2 nW[kCons];
	double nw0[5] = {0, 0, 0, 0, 0}, nvol = 0, nsx = 0, nsy = 0, nsz = 0;
	// the CSR codes of the own cell as loaded, two per word (decoded where the gather uses them)
	uint32_t ncw[3];
	int aIdx[KH];
	double nHW[KH][5];
	uint32_t nPk[KS];
	double nNx[KS], nNy[KS], nNz[KS], nS[KS];
	auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
	// halo ids of block b (meta mb: first slot, slots | leading interior slots << 16, first halo
	// entry, halo cells)
	auto issueIds = [&](int b, int4 mb) {
		const int h0 = uni(mb.z), nh = b >= 0 ? uni(mb.w) : 0;
#pragma unroll
		for (int i = 0; i < KH; i++) {
			const int h = i * NTH + t;
			aIdx[i] = h < nh ? p.haloCell[h0 + h] : -1;
		}
	};
	// own and halo states of block b (its halo ids in aIdx)
	auto issueW = [&](int b) {
		if (b < 0) return;
		const int c0 = b * CB, nc = min(CB, m.nElem - c0);
		const double2 *blk = reinterpret_cast<const double2 *>(Win + (size_t) 5 * c0);
#pragma unroll
		for (int i = 0; i < kCons; i++) {
			const int j = i * NTH + t;
			nW[i] = 2 * j < 5 * nc ? blk[j] : make_double2(0, 0);
		}
#pragma unroll
		for (int i = 0; i < KH; i++) {
			if (aIdx[i] >= 0) {
				const double *hw = Win + (size_t) 5 * aIdx[i];
#pragma unroll
				for (int kk = 0; kk < 5; kk++) nHW[i][kk] = hw[kk];
			}
		}
	};
	// slot records of block b's leading interior slots (up to KS rounds), and W0, metrics and CSR
	// codes of its own cells
	auto issueRec = [&](int b, int4 mb) {
		if (b < 0) return;
		const int s0 = uni(mb.x), nr = min(uni(mb.y) >> 16, KS * NTH);
#pragma unroll
		for (int r = 0; r < KS; r++) {
			const int j = r * NTH + t;
			if (j < nr) {
				nPk[r] = ldStream<NT>(p.pack + s0 + j);
				nNx[r] = ldStream<NT>(p.nx + s0 + j), nNy[r] = ldStream<NT>(p.ny + s0 + j);
				nNz[r] = ldStream<NT>(p.nz + s0 + j), nS[r] = ldStream<NT>(p.S + s0 + j);
			}
		}
		const int c0 = b * CB, nc = min(CB, m.nElem - c0);
		const int e = c0 + t;
		if (t < nc) {
			if (!firstStage) {
#pragma unroll
				for (int kk = 0; kk < 5; kk++) nw0[kk] = ldStream<NT>(W0 + kk * n + e);
			}
			nvol = ldStream<NT>(m.c.vol + e);
			nsx = ldStream<NT>(m.c.sx + e);
			nsy = ldStream<NT>(m.c.sy + e);
			nsz = ldStream<NT>(m.c.sz + e);
			const uint16_t *cl = p.csr + (size_t) e * p.csrStride;
			if (p.csrStride == 4 || p.csrStride == 6) {
				// rows of 8 or 12 B, 4 B aligned
				const uint32_t *cw = reinterpret_cast<const uint32_t *>(cl);
				ncw[0] = ldStream<NT>(cw);
				ncw[1] = ldStream<NT>(cw + 1);
				ncw[2] = p.csrStride == 6 ? ldStream<NT>(cw + 2) : 0xFFFFFFFFu;
			} else {
				uint32_t h[6];
#pragma unroll
				for (int i = 0; i < 6; i++) h[i] = i < p.csrStride ? (uint32_t) ldStream<NT>(cl + i) : 0xFFFFu;
#pragma unroll
				for (int i = 0; i < 3; i++) ncw[i] = h[2 * i] | h[2 * i + 1] << 16;
			}
		} else {
#pragma unroll
			for (int i = 0; i < 3; i++) ncw[i] = 0xFFFFFFFFu;
		}
	};

	const int4 none = make_int4(0, 0, 0, 0);
	auto metaOf = [&](int bb) { return bb >= 0 ? p.blockMeta[bb] : none; };
	int b = blockAt(0);
	int4 mc = metaOf(b), mn = metaOf(blockAt(1)), mnn = metaOf(blockAt(2));
	issueIds(b, mc);
	issueW(b);
	issueRec(b, mc);
	issueIds(blockAt(1), mn);
	for (int it = 0; b >= 0; it++) {
		const int bn = blockAt(it + 1), bnn = blockAt(it + 2);
		const int4 mnnn = metaOf(blockAt(it + 3)); // used three blocks on
		const int c0 = b * CB, nc = min(CB, m.nElem - c0);
		const int s0 = uni(mc.x), ns = uni(mc.y) & 0xFFFF, nInt = uni(mc.y) >> 16, nh = uni(mc.w);
		// leading interior slots whose records are in registers; the rest read theirs from memory
		const int nReg = min(nInt, KS * NTH);
		const bool own = t < nc;
		const int e = c0 + t;
		// take over the prefetched block
		double w0[5], vol = nvol, sx = nsx, sy = nsy, sz = nsz;
#pragma unroll
		for (int kk = 0; kk < 5; kk++) w0[kk] = nw0[kk];
		uint32_t cw[3];
#pragma unroll
		for (int i = 0; i < 3; i++) cw[i] = ncw[i];
		uint32_t cPk[KS];
		double cNx[KS], cNy[KS], cNz[KS], cS[KS];
#pragma unroll
		for (int r = 0; r < KS; r++) cPk[r] = nPk[r], cNx[r] = nNx[r], cNy[r] = nNy[r], cNz[r] = nNz[r], cS[r] = nS[r];
		// halo states -> primitives, own W -> the flux rows (turned into primitives after the barrier)
#pragma unroll
		for (int i = 0; i < KH; i++) {
			const int h = i * NTH + t;
			if (h < nh) {
				const int l = CB + h;
				const double r = nHW[i][0];
				double u, v, w, pr, E, H;
				primOf<MK>(r, nHW[i][1], nHW[i][2], nHW[i][3], nHW[i][4], q.gamma, u, v, w, pr, E, H);
				ls[0 * cap + l] = r;
				ls[1 * cap + l] = u;
				ls[2 * cap + l] = v;
				ls[3 * cap + l] = w;
				ls[4 * cap + l] = pr;
				ls[5 * cap + l] = H;
				lsq[l] = sqrtN(r);
				lcs[l] = sqrtN(divK<MK>(q.gamma * pr, r));
			}
		}
#pragma unroll
		for (int i = 0; i < kCons; i++) {
			const int j = i * NTH + t;
			if (2 * j < 5 * nc) {
				lf[2 * j] = nW[i].x;
				lf[2 * j + 1] = nW[i].y;
			}
		}
		__syncthreads();
		double Eown = 0.0;
		if (own) {
			const double *wt = lf + 5 * t;
			const double r = wt[0];
			double u, v, w, pr, E, H;
			primOf<MK>(r, wt[1], wt[2], wt[3], wt[4], q.gamma, u, v, w, pr, E, H);
			ls[t] = r;
			ls[cap + t] = u;
			ls[2 * cap + t] = v;
			ls[3 * cap + t] = w;
			ls[4 * cap + t] = pr;
			ls[5 * cap + t] = H;
			lsq[t] = sqrtN(r);
			lcs[t] = sqrtN(divK<MK>(q.gamma * pr, r));
			Eown = E;
		}
		__syncthreads();
		issueW(bn);
		issueIds(bnn, mnn);

		// ---- face fluxes, once each, into LDS. The block's leading interior slots first, from LDS
		// and registers only (no load in this loop, so none waits for the next block's loads);
		// then the rest (boundary and partition interface faces), whose records and ghost states
		// are read from memory
		auto lprim = [&](int l) -> Prim {
			return {ls[l], ls[cap + l], ls[2 * cap + l], ls[3 * cap + l], ls[4 * cap + l], ls[5 * cap + l]};
		};
#pragma unroll 1
		for (int r = 0; r * NTH < nReg; r++) {
			const int j = r * NTH + t;
			const uint32_t pk = cPk[0];
			const double nx = cNx[0], ny = cNy[0], nz = cNz[0], S = cS[0];
#pragma unroll
			for (int z = 0; z + 1 < KS; z++)
				cPk[z] = cPk[z + 1], cNx[z] = cNx[z + 1], cNy[z] = cNy[z + 1], cNz[z] = cNz[z + 1], cS[z] = cS[z + 1];
			if (j >= nReg) continue;
			const int l0 = (int) (pk & 0x7FFu), l1 = (int) ((pk >> 11) & 0x7FFu);
			const Prim L = lprim(l0);
			auto emit = [&](int kk, double f) { lf[kk * ms + j] = f * S; };
			if constexpr (SCHEME == 0) {
				roeFluxEmit<MK>(L, lprim(l1), lsq[l0], lsq[l1], lcs[l0], nx, ny, nz, q.gamma, emit);
			} else {
				double F[5];
				ausmFluxA(L, lprim(l1), lcs[l0], lcs[l1], nx, ny, nz, q.gamma, F);
#pragma unroll
				for (int kk = 0; kk < 5; kk++) emit(kk, F[kk]);
			}
		}
		for (int j = nReg + t; j < ns; j += NTH) {
			const int s = s0 + j;
			const uint32_t pk = ldStream<NT>(p.pack + s);
			const double nx = ldStream<NT>(p.nx + s), ny = ldStream<NT>(p.ny + s), nz = ldStream<NT>(p.nz + s);
			const double S = ldStream<NT>(p.S + s);
			const int l0 = (int) (pk & 0x7FFu), l1 = (int) ((pk >> 11) & 0x7FFu), kind = (int) (pk >> 22);
			const Prim L = lprim(l0);
			auto emit = [&](int kk, double f) { lf[kk * ms + j] = f * S; };
			if (SCHEME == 0 && kind == 0) {
				roeFluxEmit<MK>(L, lprim(l1), lsq[l0], lsq[l1], lcs[l0], nx, ny, nz, q.gamma, emit);
				continue;
			}
			double F[5];
			if (kind == 0) {
				if constexpr (SCHEME == 1) ausmFluxA(L, lprim(l1), lcs[l0], lcs[l1], nx, ny, nz, q.gamma, F);
				else schemeFlux<SCHEME>(L, lprim(l1), nx, ny, nz, q.gamma, F);
			} else if (kind == 4) {
				schemeFlux<SCHEME>(L, loadPrim(in, p.e1[s]), nx, ny, nz, q.gamma, F);
			} else {
				const int c = p.e1[s];
				Ghost G;
				if (kind == 3) {
					const double *w5 = Win + (size_t) 5 * (c0 + l0);
					const double Ei = quot<MK>(w5[4], w5[0], recipN(w5[0]));
					G = farfieldGhost(in, c, machOf(L.rho, L.u, L.v, L.w, L.p, q.gamma), nx, ny, nz, q, L, Ei);
				} else {
					G = mirrorGhost(L.rho, L.u, L.v, L.w, L.p, nx, ny, nz, q.gamma);
				}
				if (kind == 3 || rmsPartial) storeGhost(out, c, G);
				if (kind == 1) {
					F[0] = 0;
					F[1] = nx * G.p;
					F[2] = ny * G.p;
					F[3] = nz * G.p;
					F[4] = 0;
				} else if (kind == 2) {
					schemeFlux<SCHEME>(L, Prim{G.rho, G.u, G.v, G.w, G.p, G.H}, nx, ny, nz, q.gamma, F);
				} else {
					double V = G.u * nx + G.v * ny + G.w * nz;
					F[0] = G.rho * V;
					F[1] = G.rho * G.u * V + G.p * nx;
					F[2] = G.rho * G.v * V + G.p * ny;
					F[3] = G.rho * G.w * V + G.p * nz;
					F[4] = G.rho * G.H * V;
				}
			}
#pragma unroll
			for (int kk = 0; kk < 5; kk++) emit(kk, F[kk]);
		}
		issueRec(bn, mn);
		__syncthreads();

		// ---- one thread per cell: gather in reference order, local dt, RK update
		double r2 = 0.0;
		if (own) {
			double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
			for (int i = 0; i < 6; i++) {
				const unsigned cd = (cw[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
				if (cd == 0xFFFFu) break;
				const int j = (int) (cd & 0x7FFFu);
				if (cd & 0x8000u) {
#pragma unroll
					for (int kk = 0; kk < 5; kk++) r[kk] -= lf[kk * ms + j];
				} else {
#pragma unroll
					for (int kk = 0; kk < 5; kk++) r[kk] += lf[kk * ms + j];
				}
			}
			const double rho = ls[t], u = ls[cap + t], v = ls[2 * cap + t], w = ls[3 * cap + t];
			if (firstStage) {
				w0[0] = rho;
				w0[1] = u * rho;
				w0[2] = v * rho;
				w0[3] = w * rho;
				w0[4] = Eown * rho;
#pragma unroll
				for (int kk = 0; kk < 5; kk++) W0[kk * n + e] = w0[kk];
			}
			const double c = lcs[t];
			double lx = 0.5 * (fabs(u) + c) * sx;
			double ly = 0.5 * (fabs(v) + c) * sy;
			double lz = 0.5 * (fabs(w) + c) * sz;
			double dt = divK<MK>(q.cfl * vol, lx + ly + lz);
			const double yv = recipN(vol);
			double *ow = Wout + (size_t) 5 * e;
			ow[0] = w0[0] - quot<MK>(alpha * dt * r[0], vol, yv);
			ow[1] = w0[1] - quot<MK>(alpha * dt * r[1], vol, yv);
			ow[2] = w0[2] - quot<MK>(alpha * dt * r[2], vol, yv);
			ow[3] = w0[3] - quot<MK>(alpha * dt * r[3], vol, yv);
			ow[4] = w0[4] - quot<MK>(alpha * dt * r[4], vol, yv);
			r2 = r[0] * r[0];
		}
		if (rmsPartial) {
			__syncthreads();
			const double tt = blockSumN<NTH>(r2, lds);
			if (t == 0) rmsPartial[b] = tt;
		}
		__syncthreads(); // the next block overwrites LDS
		b = bn;
		mc = mn;
		mn = mnn;
		mnn = mnnn;
	}
}

// self-check of sqrtN, recipN and divN against the library square root and division (bitwise)
__device__ __forceinline__ uint64_t mixBits(uint64_t z) {
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}
__device__ __forceinline__ double randDouble(uint64_t z, int span, bool sign) {
	const uint64_t ex = 1023 - span + (z >> 53) % (2 * span + 1);
	uint64_t b = (ex << 52) | (z & 0xFFFFFFFFFFFFFull);
	if (sign && ((z >> 40) & 1)) b |= 0x8000000000000000ull;
	return __longlong_as_double((long long) b);
}
__global__ void __launch_bounds__(kBlock) kMathCheck(uint64_t seed, uint64_t base, long long n,
                                                      unsigned long long *bad) {
	const uint64_t i = base + (uint64_t) blockIdx.x * kBlock + threadIdx.x;
	if ((long long) i >= n) return;
	const uint64_t z = mixBits(seed ^ (3 * i)), z1 = mixBits(seed ^ (3 * i + 1)), z2 = mixBits(seed ^ (3 * i + 2));
	const double xs = randDouble(z, 700, false);
	const double d = randDouble(z1, 400, true);
	double x = randDouble(z2, 400, true);
	if ((i & 15) == 0) x = 0.0;
	if ((i & 15) == 1) x = -0.0;
	unsigned int m = 0;
	if (__double_as_longlong(sqrtN(xs)) != __double_as_longlong(sqrt(xs))) m |= 1;
	if (__double_as_longlong(recipN(d)) != __double_as_longlong(1.0 / d)) m |= 2;
	if (__double_as_longlong(divN(x, d)) != __double_as_longlong(x / d)) m |= 4;
	if (m & 1) atomicAdd(bad, 1ull);
	if (m & 2) atomicAdd(bad + 1, 1ull);
	if (m & 4) atomicAdd(bad + 2, 1ull);
}

// interior primitives (64 B records) from the conservative states the fused RK stage keeps
__global__ void __launch_bounds__(kBlock) kWToPrim(DevMesh m, const double *W, DevState s, DevFree q) {
	const int e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= m.nElem) return;
	const double *w5 = W + (size_t) 5 * e;
	const double r = w5[0];
	double u, v, w, pr, E, H;
	primOf<false>(r, w5[1], w5[2], w5[3], w5[4], q.gamma, u, v, w, pr, E, H);
	double2 *c = reinterpret_cast<double2 *>(cellPtr(s, e));
	c[0] = make_double2(r, u);
	c[1] = make_double2(v, w);
	c[2] = make_double2(pr, E);
	c[3] = make_double2(H, machOf(r, u, v, w, pr, q.gamma));
}

// deterministic sum of n doubles in two launches over several workgroups: kSumChunks reduces
// chunk g = [1024 g, 1024 g + 1024) (each thread adds its four elements t, t + 256, t + 512,
// t + 768 in that order, then groupSum) into tmp[g]; kSumTotal adds the chunk sums the same way.
constexpr int kSumThreads = 256, kSumChunk = 4 * kSumThreads;
__global__ void __launch_bounds__(kSumThreads) kSumChunks(const double *in, int n, double *tmp) {
	__shared__ double lds[kSumThreads / 64];
	const int base = blockIdx.x * kSumChunk + threadIdx.x;
	double v[4];
#pragma unroll
	for (int i = 0; i < 4; i++) v[i] = base + i * kSumThreads < n ? in[base + i * kSumThreads] : 0.0;
	const double s = groupSum<kSumThreads>(((v[0] + v[1]) + v[2]) + v[3], lds);
	if (threadIdx.x == 0) tmp[blockIdx.x] = s;
}
__global__ void __launch_bounds__(kSumThreads) kSumTotal(const double *tmp, int chunks, double *out) {
	__shared__ double lds[kSumThreads / 64];
	double acc = 0.0;
	for (int i = threadIdx.x; i < chunks; i += kSumThreads) acc += tmp[i];
	const double s = groupSum<kSumThreads>(acc, lds);
	if (threadIdx.x == 0) out[0] = s;
}

__global__ void __launch_bounds__(kBlock) kMach(DevMesh m, DevState s, DevFree q) {
	int e = blockIdx.x * kBlock + threadIdx.x;
	if (e >= m.nElem) return;
	double *c = cellPtr(s, e);
	c[S_M] = machOf(c[S_RHO], c[S_U], c[S_V], c[S_W], c[S_P], q.gamma);
}

// AeroCoefficients::Update (AeroCoefficients.cpp:38-52): per wall face coeff = N*cp*Cref/Sref and
// coeff*x_face; summed per thread over a contiguous chunk, then a fixed tree
// AeroCoefficients::Update (AeroCoefficients.cpp:38-52) over the wall faces, in two launches over
// several workgroups: kAeroChunks sums chunk g of 4 kBlock faces (each thread its four faces t,
// t + kBlock, ... in that order, then groupSum) into tmp[6 g + k]; kAeroTotal adds the chunk sums
// the same way. The wall cells' p comes from W when W is not null (as the update forms it), else
// from the primitive records; cpOut (may be null) receives each face's cp.
constexpr int kAeroChunk = 4 * kBlock;
__global__ void __launch_bounds__(kBlock) kAeroChunks(int nWall, const int *wallCell, const double *Nx,
                                                      const double *Ny, const double *Nz, const double *Cx,
                                                      const double *Cy, const double *Cz, DevState s, const double *W,
                                                      double gamma, double cpFactor, double pInf, double cref,
                                                      double sref, double *tmp, double *cpOut) {
	__shared__ double lds[kBlock / 64];
	double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
	for (int j = 0; j < 4; j++) {
		const int i = blockIdx.x * kAeroChunk + j * kBlock + threadIdx.x;
		if (i >= nWall) break;
		double pw;
		if (W) {
			// p of the conservative state, as the update forms it
			const double *w5 = W + (size_t) 5 * wallCell[i];
			double u, v, w, E, H;
			primOf<false>(w5[0], w5[1], w5[2], w5[3], w5[4], gamma, u, v, w, pw, E, H);
		} else {
			pw = cellPtr(s, wallCell[i])[S_P];
		}
		double cp = cpFactor * (pw - pInf);
		if (cpOut) cpOut[i] = cp;
		double cx = Nx[i] * cp * cref / sref;
		double cy = Ny[i] * cp * cref / sref;
		double cz = Nz[i] * cp * cref / sref;
		acc[0] += cx;
		acc[1] += cy;
		acc[2] += cz;
		acc[3] += cx * Cx[i];
		acc[4] += cy * Cy[i];
		acc[5] += cz * Cz[i];
	}
	for (int k = 0; k < 6; k++) {
		const double t = groupSum<kBlock>(acc[k], lds);
		if (threadIdx.x == 0) tmp[6 * blockIdx.x + k] = t;
		__syncthreads();
	}
}
__global__ void __launch_bounds__(kBlock) kAeroTotal(const double *tmp, int chunks, double *out6) {
	__shared__ double lds[kBlock / 64];
	for (int k = 0; k < 6; k++) {
		double acc = 0.0;
		for (int g = threadIdx.x; g < chunks; g += kBlock) acc += tmp[6 * g + k];
		const double t = groupSum<kBlock>(acc, lds);
		if (threadIdx.x == 0) out6[k] = t;
		__syncthreads();
	}
}

// Interface exchange between partitions (MPIHandler::updateFlowField, MPIHandler.cpp:84-267).
// Pack: entry i puts (rho, u, v, w, E) of interior cell src[i] into its message at buf[slot.x +
// k * slot.y], k = rho, u, v, w, E (each message holds its rho's, then its u's, v's, w's and E's,
// as the reference sends them). The state is read from the conservative states W when W is given
// (primitives formed as the update forms them), else from the primitive records.
__global__ void __launch_bounds__(kBlock) kHaloPack(int n, const int *src, const int2 *slot, DevState s,
                                                    const double *W, double gamma, double *buf) {
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) return;
	const int c = src[i];
	double v[5];
	if (W) {
		const double *w5 = W + (size_t) 5 * c;
		double pr, H;
		v[0] = w5[0];
		primOf<true>(w5[0], w5[1], w5[2], w5[3], w5[4], gamma, v[1], v[2], v[3], pr, v[4], H);
	} else {
		const double *r = cellPtr(s, c);
		v[0] = r[S_RHO], v[1] = r[S_U], v[2] = r[S_V], v[3] = r[S_W], v[4] = r[S_E];
	}
	const int2 sl = slot[i];
#pragma unroll
	for (int k = 0; k < 5; k++) buf[sl.x + k * sl.y] = v[k];
}

// Unpack: ghost dst[i] takes (rho, u, v, w, E) from its message; p, H and M are formed as the
// reference's receiving rank forms them (MPIHandler.cpp:146-165): p = (gamma - 1) rho (E -
// (u^2 + v^2 + w^2) / 2), H = E + p / rho, M = |u| / sqrt(gamma (p / rho)). p and H are the update's
// own expressions, so a ghost carries the bits of the cell it copies.
__global__ void __launch_bounds__(kBlock) kHaloUnpack(int n, const int *dst, const int2 *slot, const double *buf,
                                                      double gamma, DevState s) {
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) return;
	const int2 sl = slot[i];
	const double r = buf[sl.x], u = buf[sl.x + sl.y], v = buf[sl.x + 2 * sl.y], w = buf[sl.x + 3 * sl.y];
	const double E = buf[sl.x + 4 * sl.y];
	const double p = (gamma - 1) * r * (E - (u * u + v * v + w * w) / 2.0);
	const double H = E + p / r;
	const double a = sqrt(gamma * (p / r));
	const double M = sqrt(u * u + v * v + w * w) / a;
	double2 *c = reinterpret_cast<double2 *>(cellPtr(s, dst[i]));
	c[0] = make_double2(r, u);
	c[1] = make_double2(v, w);
	c[2] = make_double2(p, E);
	c[3] = make_double2(H, M);
}

// plain copy of n 16 B chunks, each read once and written once (a kernel of known traffic for
// scaling the memory counters)
__global__ void __launch_bounds__(kBlock) kCopy16(const double2 *src, double2 *dst, long long n) {
	const long long stride = (long long) gridDim.x * kBlock;
	for (long long i = (long long) blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

inline int blocksFor(int n) { return (n + kBlock - 1) / kBlock; }

}// namespace

int stageBlocks(int nElem) { return blocksFor(nElem); }

// whether a plan runs the fixed-stride kernel (256-cell blocks, conservative-state layout)
bool stageFusedFixed(int cellsPerBlock, int maxSlots, int maxHalo, int layout) {
	return cellsPerBlock == 256 && layout == 2 && cellsPerBlock + maxHalo <= kFixCap && maxSlots <= kFixSlots;
}

size_t stageFusedLdsBytes(int cellsPerBlock, int maxSlots, int maxHalo, int layout) {
	if (stageFusedFixed(cellsPerBlock, maxSlots, maxHalo, layout)) maxHalo = kFixCap - cellsPerBlock, maxSlots = kFixSlots;
	// the flux region also stages the block's conservative states (5 per cell) before the fluxes
	const size_t cb = (size_t) cellsPerBlock, cap = cb + maxHalo;
	const size_t flux = 5 * std::max((size_t) maxSlots, cb);
	const size_t words = 6 * cap + (layout == 2 ? 2 * cap : 2 * cb) + flux;
	return sizeof(double) * words;
}

int stageFusedWavesPerCU(int cellsPerBlock, int maxSlots, int maxHalo, int scheme) {
	const size_t lds = std::max(stageFusedLdsBytes(cellsPerBlock, maxSlots, maxHalo, 1),
	                            stageFusedLdsBytes(cellsPerBlock, maxSlots, maxHalo, 2));
	if (lds > 160 * 1024) return 0;
	int n = 0;
	hipError_t e;
	if (cellsPerBlock == 128)
		e = scheme == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kStageFused<1, true, true, 0, 2, 128>, 128, lds)
		                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kStageFused<0, true, true, 0, 2, 128>, 128, lds);
	else
		e = scheme == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kStageFused<1, true, true, 0, 2, 256>, 256, lds)
		                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kStageFused<0, true, true, 0, 2, 256>, 256, lds);
	if (e != hipSuccess || n <= 0) n = 1;
	return n * cellsPerBlock / 64;
}

void launchHaloExchange(int n, const int *src, const int *dst, const int *slot, DevState s, const double *W,
                        double gamma, double *buf, hipStream_t st) {
	if (n <= 0) return;
	const int2 *sl = reinterpret_cast<const int2 *>(slot);
	hipLaunchKernelGGL(kHaloPack, dim3(blocksFor(n)), dim3(kBlock), 0, st, n, src, sl, s, W, gamma, buf);
	hipLaunchKernelGGL(kHaloUnpack, dim3(blocksFor(n)), dim3(kBlock), 0, st, n, dst, sl, buf, gamma, s);
}

void launchCopy16(const void *src, void *dst, long long chunks, hipStream_t st) {
	if (chunks <= 0) return;
	const long long want = (chunks + kBlock - 1) / kBlock;
	const unsigned grid = (unsigned) std::min<long long>(want, 65536);
	hipLaunchKernelGGL(kCopy16, dim3(grid), dim3(kBlock), 0, st, static_cast<const double2 *>(src),
	                   static_cast<double2 *>(dst), chunks);
}

void launchMathCheck(unsigned long long seed, long long n, unsigned long long *bad, hipStream_t st) {
	const long long perLaunch = (long long) kBlock * 65536;
	for (long long base = 0; base < n; base += perLaunch) {
		const long long cnt = std::min(perLaunch, n - base);
		hipLaunchKernelGGL(kMathCheck, dim3((unsigned) ((cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, (uint64_t) seed,
		                   (uint64_t) base, n, bad);
	}
}

void launchWToPrim(const DevMesh &m, const double *W, DevState s, const DevFree &q, hipStream_t st) {
	if (m.nElem == 0) return;
	hipLaunchKernelGGL(kWToPrim, dim3(blocksFor(m.nElem)), dim3(kBlock), 0, st, m, W, s, q);
}

void launchBc(const DevMesh &m, DevState s, const DevFree &q, int machDerived, hipStream_t st) {
	if (m.nGhost == 0) return;
	hipLaunchKernelGGL(kBc, dim3(blocksFor(m.nGhost)), dim3(kBlock), 0, st, m, s, q, machDerived);
}

void launchFlux(const DevMesh &m, DevState s, const DevFree &q, int scheme, double *fs, double *raw, hipStream_t st) {
	if (m.nFace == 0) return;
	hipLaunchKernelGGL(kFlux, dim3(blocksFor(m.nFace)), dim3(kBlock), 0, st, m, s, q, scheme, fs, raw);
}

void launchGather(const DevMesh &m, const double *fs, double *R, hipStream_t st) {
	if (m.nElem == 0) return;
	hipLaunchKernelGGL(kGather, dim3(blocksFor(m.nElem)), dim3(kBlock), 0, st, m, fs, R);
}

void launchTimestep(const DevMesh &m, DevState s, const DevFree &q, double *dt, hipStream_t st) {
	if (m.nElem == 0) return;
	hipLaunchKernelGGL(kTimestep, dim3(blocksFor(m.nElem)), dim3(kBlock), 0, st, m, s, q, dt);
}

void launchStage(const DevMesh &m, DevState s, const DevFree &q, const double *fs, const double *Rin, double *W0,
                 double alpha, int firstStage, int integ, double *rmsPartial, double *Rout, hipStream_t st) {
	if (m.nElem == 0) return;
	hipLaunchKernelGGL(kStage, dim3(blocksFor(m.nElem)), dim3(kBlock), 0, st, m, s, q, fs, Rin, W0, alpha, firstStage,
	                   integ, rmsPartial, Rout);
}

void launchSmooth(const double *R0, double *R1, double *R2, const int *nbr, const uint8_t *nface, int stride, int nElem,
                  double *blockMax, int *flag, const int *partStart, int nPart, double *cellMax, hipStream_t st) {
	if (nElem == 0) return;
	const int nb = blocksFor(nElem);
	if (nPart > 1) {
		hipLaunchKernelGGL(kSmooth1, dim3(nb), dim3(kBlock), 0, st, R0, R1, nbr, nface, stride, nElem, blockMax, cellMax);
		hipLaunchKernelGGL(kSmoothDecideParts, dim3(nPart), dim3(kBlock), 0, st, cellMax, partStart, nPart, flag);
	} else {
		hipLaunchKernelGGL(kSmooth1, dim3(nb), dim3(kBlock), 0, st, R0, R1, nbr, nface, stride, nElem, blockMax,
		                   (double *) nullptr);
		hipLaunchKernelGGL(kSmoothDecide, dim3(1), dim3(kBlock), 0, st, blockMax, nb, flag);
	}
	hipLaunchKernelGGL(kSmooth2, dim3(nb), dim3(kBlock), 0, st, R0, R1, R2, nbr, nface, stride, nElem, flag, partStart,
	                   nPart);
}

// The pipelined persistent stage: the shape a plan runs with (threads, halo cells and slot
// rounds held per thread), 0 when it runs the one-block-per-workgroup kernel instead
namespace {
struct PipeShape {
	int nth = 0, kh = 0;
	bool fixed = false;
};
PipeShape pipeShape(const DevPlan &p) {
	PipeShape s;
	if (p.cellsPerBlock != 256) return s;
	const size_t lds = stageFusedLdsBytes(256, p.maxSlots, p.maxHalo, 2);
	if (2 * lds > 160 * 1024) {
		if (p.maxHalo <= 2 * 512) s.nth = 512, s.kh = 2;
	} else if (p.maxHalo <= 2 * 256) {
		s.nth = 256, s.kh = p.maxHalo <= 256 ? 1 : 2;
	}
	s.fixed = stageFusedFixed(256, p.maxSlots, p.maxHalo, 2);
	return s;
}
// slot records held: 3 rounds of 256 (tetrahedral blocks), 2 of 512
template<int S, bool F, int NTH, int KH>
const void *pipeKernel() {
	return reinterpret_cast<const void *>(&kStagePipe<S, F, NTH, KH, NTH == 512 ? 2 : 3>);
}
const void *pipeKernelFor(int scheme, const PipeShape &s) {
	if (s.nth == 512) return scheme == 1 ? (s.fixed ? pipeKernel<1, true, 512, 2>() : pipeKernel<1, false, 512, 2>())
	                                      : (s.fixed ? pipeKernel<0, true, 512, 2>() : pipeKernel<0, false, 512, 2>());
	if (s.kh == 1) return scheme == 1 ? (s.fixed ? pipeKernel<1, true, 256, 1>() : pipeKernel<1, false, 256, 1>())
	                                   : (s.fixed ? pipeKernel<0, true, 256, 1>() : pipeKernel<0, false, 256, 1>());
	return scheme == 1 ? (s.fixed ? pipeKernel<1, true, 256, 2>() : pipeKernel<1, false, 256, 2>())
	                   : (s.fixed ? pipeKernel<0, true, 256, 2>() : pipeKernel<0, false, 256, 2>());
}
}// namespace

int stagePipeGrid(const DevPlan &p, int scheme, int device, int xcds) {
	const PipeShape s = pipeShape(p);
	if (!s.nth) return 0;
	const size_t lds = stageFusedLdsBytes(256, p.maxSlots, p.maxHalo, 2);
	int perCU = 0, cus = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, pipeKernelFor(scheme, s), s.nth, lds) != hipSuccess ||
	    perCU <= 0)
		return 0;
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) return 0;
	const int X = std::max(1, xcds);
	const int grid = perCU * cus / X * X;
	return grid >= X ? grid : 0;
}

void launchStageFused(const DevMesh &m, const DevPlan &p, DevState in, DevState out, const double *Win, double *Wout,
                      int layout, const DevFree &q, int scheme, int integ, double *W0, double alpha, int firstStage,
                      int machDerived, double *rmsPartial, int parity, const StageTuning &t, hipStream_t st) {
	if (p.nBlocks == 0) return;
	const int rev = t.alternate ? (parity & 1) : 0;
	const size_t lds = stageFusedLdsBytes(p.cellsPerBlock, p.maxSlots, p.maxHalo, integ == 1 ? 0 : layout);
	const bool fixed = integ != 1 && stageFusedFixed(p.cellsPerBlock, p.maxSlots, p.maxHalo, layout);
	// 256-cell blocks whose LDS leaves room for one workgroup per CU (hexahedral meshes with large
	// halos) run with 768 (or 512) threads, so that three (two) waves per SIMD share it
	const bool wide = integ != 1 && p.cellsPerBlock == 256 && 2 * lds > 160 * 1024;
	const int x = t.xcds;
	const int fixedShape = t.fixedShape >= 0 ? t.fixedShape : scheme == 1 ? 1 : 2;
	if (t.pipeGrid > 0 && integ != 1 && layout == 2 && t.nontemporal && t.sharedRecip) {
		const PipeShape ps = pipeShape(p);
		if (ps.nth) {
			const size_t lp = stageFusedLdsBytes(256, p.maxSlots, p.maxHalo, 2);
			const int g = t.pipeGrid;
#define E3D_PIPE(S, F, NTH, KH)                                                                                   \
	hipLaunchKernelGGL((kStagePipe<S, F, NTH, KH, NTH == 512 ? 2 : 3>), dim3(g), dim3(NTH), lp, st, m, p, in, out, Win, \
	                   Wout, q, W0, alpha, firstStage, rmsPartial, x, rev)
#define E3D_PIPE_F(S, NTH, KH)            \
	if (ps.fixed) E3D_PIPE(S, true, NTH, KH); \
	else E3D_PIPE(S, false, NTH, KH)
#define E3D_PIPE_S(S)                                    \
	if (ps.nth == 512) { E3D_PIPE_F(S, 512, 2); }         \
	else if (ps.kh == 1) { E3D_PIPE_F(S, 256, 1); }      \
	else { E3D_PIPE_F(S, 256, 2); }
			if (scheme == 1) { E3D_PIPE_S(1) }
			else { E3D_PIPE_S(0) }
#undef E3D_PIPE_S
#undef E3D_PIPE_F
#undef E3D_PIPE
			return;
		}
	}
#define E3D_LAUNCH_B(S, N, K, B)                                                                               \
	do {                                                                                                       \
		if (integ == 1)                                                                                        \
			hipLaunchKernelGGL((kStageFused<S, N, K, 1, 0, B>), dim3(p.nBlocks), dim3(B), lds, st, m, p, in, out, \
			                   Win, Wout, q, W0, alpha, 0, machDerived, rmsPartial, x, rev);                      \
		else if (layout == 2 && wide && t.wideThreads == 768)                                                  \
			hipLaunchKernelGGL((kStageFused<S, N, K, 0, 2, 256, false, 768>), dim3(p.nBlocks), dim3(768), lds, st, m, \
			                   p, in, out, Win, Wout, q, W0, alpha, firstStage, machDerived, rmsPartial, x, rev); \
		else if (layout == 2 && wide)                                                                          \
			hipLaunchKernelGGL((kStageFused<S, N, K, 0, 2, 256, false, 512>), dim3(p.nBlocks), dim3(512), lds, st, m, \
			                   p, in, out, Win, Wout, q, W0, alpha, firstStage, machDerived, rmsPartial, x, rev); \
		else if (layout == 2 && fixed && fixedShape == 1)                                                    \
			hipLaunchKernelGGL((kStageFused<S, N, K, 0, 2, 256, true, 320, true>), dim3(p.nBlocks), dim3(320), lds, st, \
			                   m, p, in, out, Win, Wout, q, W0, alpha, firstStage, machDerived, rmsPartial, x, rev); \
		else if (layout == 2 && fixed && fixedShape == 2)                                                    \
			hipLaunchKernelGGL((kStageFused<S, N, K, 0, 2, 256, true, 256, true>), dim3(p.nBlocks), dim3(256), lds, st, \
			                   m, p, in, out, Win, Wout, q, W0, alpha, firstStage, machDerived, rmsPartial, x, rev); \
		else if (layout == 2 && fixed)                                                                         \
			hipLaunchKernelGGL((kStageFused<S, N, K, 0, 2, 256, true>), dim3(p.nBlocks), dim3(256), lds, st, m, p, \
			                   in, out, Win, Wout, q, W0, alpha, firstStage, machDerived, rmsPartial, x, rev);     \
		else if (layout == 2)                                                                                  \
			hipLaunchKernelGGL((kStageFused<S, N, K, 0, 2, B>), dim3(p.nBlocks), dim3(B), lds, st, m, p, in, out, \
			                   Win, Wout, q, W0, alpha, firstStage, machDerived, rmsPartial, x, rev);             \
		else                                                                                                   \
			hipLaunchKernelGGL((kStageFused<S, N, K, 0, 1, B>), dim3(p.nBlocks), dim3(B), lds, st, m, p, in, out, \
			                   Win, Wout, q, W0, alpha, firstStage, machDerived, rmsPartial, x, rev);             \
	} while (0)
#define E3D_LAUNCH(S, N, K)                                  \
	do {                                                     \
		if (p.cellsPerBlock == 128) E3D_LAUNCH_B(S, N, K, 128); \
		else E3D_LAUNCH_B(S, N, K, 256);                        \
	} while (0)
#define E3D_LAUNCH_K(S, N)               \
	if (t.sharedRecip) E3D_LAUNCH(S, N, true); \
	else E3D_LAUNCH(S, N, false)
	if (scheme == 1) {
		if (t.nontemporal) E3D_LAUNCH_K(1, true);
		else E3D_LAUNCH_K(1, false);
	} else {
		if (t.nontemporal) E3D_LAUNCH_K(0, true);
		else E3D_LAUNCH_K(0, false);
	}
#undef E3D_LAUNCH_K
#undef E3D_LAUNCH
#undef E3D_LAUNCH_B
}

// Per-iteration RMS sums of the allocation-free iteration entry point: a static device array,
// one copy per device, shared by every call on that device. Host code never passes `scratch`
// itself to a kernel (in host code the name is only a stand-in); it uses the address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[kIterScratchMax];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScratchMu;
double *gScratch[kMaxDevices] = {};
}// namespace

hipError_t iterScratchInit(int device) {
	if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lk(gScratchMu);
	if (gScratch[device]) return hipSuccess;
	// the caller has made `device` current
	void *ptr = nullptr;
	hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
	if (e != hipSuccess) return e;
	gScratch[device] = static_cast<double *>(ptr);
	return hipSuccess;
}

double *iterScratch(int device) {
	if (device < 0 || device >= kMaxDevices) return nullptr;
	std::lock_guard<std::mutex> lk(gScratchMu);
	return gScratch[device];
}

int sumScratch(int n) { return std::max(1, (n + kSumChunk - 1) / kSumChunk); }

void launchSum(const double *in, int n, double *tmp, double *out, hipStream_t st) {
	const int chunks = sumScratch(n);
	hipLaunchKernelGGL(kSumChunks, dim3(chunks), dim3(kSumThreads), 0, st, in, n, tmp);
	hipLaunchKernelGGL(kSumTotal, dim3(1), dim3(kSumThreads), 0, st, tmp, chunks, out);
}

void launchMach(const DevMesh &m, DevState s, const DevFree &q, hipStream_t st) {
	if (m.nElem == 0) return;
	hipLaunchKernelGGL(kMach, dim3(blocksFor(m.nElem)), dim3(kBlock), 0, st, m, s, q);
}

int aeroScratch(int nWall) { return 6 * std::max(1, (nWall + kAeroChunk - 1) / kAeroChunk); }

void launchAero(int nWall, const int *wallCell, const double *Nx, const double *Ny, const double *Nz, const double *Cx,
                const double *Cy, const double *Cz, DevState s, const double *W, double gamma, double cpFactor,
                double pInf, double cref, double sref, double *tmp, double *out6, double *cpOut, hipStream_t st) {
	const int chunks = std::max(1, (nWall + kAeroChunk - 1) / kAeroChunk);
	hipLaunchKernelGGL(kAeroChunks, dim3(chunks), dim3(kBlock), 0, st, nWall, wallCell, Nx, Ny, Nz, Cx, Cy, Cz, s, W,
	                   gamma, cpFactor, pInf, cref, sref, tmp, cpOut);
	hipLaunchKernelGGL(kAeroTotal, dim3(1), dim3(kBlock), 0, st, tmp, chunks, out6);
}

}// namespace e3d
