// HIP kernels of the mixed-precision multi-case solver for gfx950 (wave64).
//
// One launch per RK stage walks the (case, block) table. A block of kMixBlock
// consecutive cells of one case is processed as in the FP64 fused stage: own and halo cells are
// staged in LDS, every face touching the block is evaluated once (FP32, perturbation form), then
// one thread per cell gathers its faces in ascending reference face id with a Kahan-compensated
// FP64 sum, forms the local time step and applies the FP64 RK update.
//
// FP32 quotients and square roots use the hardware approximations (v_rcp_f32, v_sqrt_f32, 1 ulp):
// their error is of the order of the FP32 rounding of the inputs.
//
// The library is compiled with -ffp-contract=off for the bit-exact FP64 path. The FP32 interior
// fluxes (physDelta, roeDelta, ausmDelta) turn it back on locally: a multiply-add becomes one
// FMA (about 1 % of the stage time). Every term of those fluxes carries a perturbation factor, so
// a freestream state still gives an exact zero. The ghost states, the time step and the FP64
// update keep separately rounded products.
#include "mixed.hpp"
#include "reduce.hpp"

namespace e3d {

namespace {

constexpr int kB = kMixBlock;
#define kConst __attribute__((address_space(4)))

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template<bool NT, typename T>
__device__ __forceinline__ T ldS(const T *p) {
	if constexpr (NT) return __builtin_nontemporal_load(p);
	else return *p;
}

// freestream constants of one case (FP32)
struct Free32 {
	float r, u, v, w, p, h, m, g, gr; // gr = gamma / (gamma - 1)
};

__device__ __forceinline__ Free32 free32(const kConst MixCase &C) {
	return {C.rInf, C.uInf, C.vInf, C.wInf, C.pInf, C.hInf, C.mInf, C.gamma, C.gamma / (C.gamma - 1.0f)};
}

// perturbation state of a cell: primitives minus freestream
struct D {
	float r, u, v, w, p, H;
};

// H - H_inf from the other perturbations, without cancellation:
// g/(g-1) (p/rho - p_inf/rho_inf) + du (u_inf + du/2) + dv (...) + dw (...)
__device__ __forceinline__ float deltaH(float dr, float du, float dv, float dw, float dp, const Free32 &f) {
	const float rho = f.r + dr;
	const float dpr = (dp * f.r - f.p * dr) * frcp(rho * f.r);
	return f.gr * dpr + du * (f.u + 0.5f * du) + dv * (f.v + 0.5f * dv) + dw * (f.w + 0.5f * dw);
}

// Physical flux through n of the state inf + d, minus the freestream flux through n. With
// m = rho V: F = (m, m u + p n, m H); F - F_inf = (dm, dm u + m_inf du + dp n, dm H + m_inf dH),
// dm = rho_inf dV + d_rho V.
__device__ __forceinline__ void physDelta(const D &d, float nx, float ny, float nz, float Vinf, float minf,
                                          const Free32 &f, float out[5]) {
#pragma clang fp contract(fast)
	const float dV = d.u * nx + d.v * ny + d.w * nz;
	const float V = Vinf + dV;
	const float dm = f.r * dV + d.r * V;
	out[0] = dm;
	out[1] = dm * (f.u + d.u) + minf * d.u + d.p * nx;
	out[2] = dm * (f.v + d.v) + minf * d.v + d.p * ny;
	out[3] = dm * (f.w + d.w) + minf * d.w + d.p * nz;
	out[4] = dm * (f.h + d.H) + minf * d.H;
}

// Roe flux (ConvectiveFluxesSchemes.cpp:10-165: Roe averages, Harten fix on the acoustic waves
// with delta = c_L / 15) minus the freestream flux. The central part is the average of the two
// physical-flux perturbations; the dissipation uses jumps of the perturbations, which equal the
// jumps of the primitives.
__device__ __forceinline__ void roeDelta(const D &L, const D &R, float nx, float ny, float nz, const Free32 &f,
                                         float out[5]) {
#pragma clang fp contract(fast)
	const float Vinf = f.u * nx + f.v * ny + f.w * nz;
	const float minf = f.r * Vinf;
	float FL[5], FR[5];
	physDelta(L, nx, ny, nz, Vinf, minf, f, FL);
	physDelta(R, nx, ny, nz, Vinf, minf, f, FR);
	const float rL = f.r + L.r, rR = f.r + R.r;
	const float uL = f.u + L.u, vL = f.v + L.v, wL = f.w + L.w, hL = f.h + L.H;
	const float uR = f.u + R.u, vR = f.v + R.v, wR = f.w + R.w, hR = f.h + R.H;
	const float sL = fsqrt(rL), sR = fsqrt(rR);
	const float rT = sL * sR;
	const float iden = frcp(sL + sR);
	const float uT = (uL * sL + uR * sR) * iden;
	const float vT = (vL * sL + vR * sR) * iden;
	const float wT = (wL * sL + wR * sR) * iden;
	const float hT = (hL * sL + hR * sR) * iden;
	const float q2 = uT * uT + vT * vT + wT * wT;
	const float c2 = (f.g - 1.0f) * (hT - 0.5f * q2);
	const float cT = fsqrt(c2);
	const float ic2 = frcp(c2);
	const float VT = uT * nx + vT * ny + wT * nz;
	const float dU = R.u - L.u, dV = R.v - L.v, dW = R.w - L.w, dP = R.p - L.p, dR = R.r - L.r;
	const float dVn = dU * nx + dV * ny + dW * nz;
	const float cL = fsqrt(f.g * (f.p + L.p) * frcp(rL));
	const float del = cL * (1.0f / 15.0f);
	const float a1 = VT - cT, a5 = VT + cT;
	const float l1 = fabsf(a1) > del ? fabsf(a1) : (a1 * a1 + del * del) * frcp(2.0f * del);
	const float l5 = fabsf(a5) > del ? fabsf(a5) : (a5 * a5 + del * del) * frcp(2.0f * del);
	const float k1 = dP - rT * cT * dVn;
	const float k5 = dP + rT * cT * dVn;
	const float kr = dR - dP * ic2;
	const float aV = fabsf(VT);
	// e1 h1 + e5 h5 with e1,5 = (1, u~ -+ c~ n, H~ -+ c~ V~): (h1 + h5) (1, u~, H~) + c~ (h5 - h1) (0, n, V~)
	const float h1 = 0.5f * l1 * k1 * ic2, h5 = 0.5f * l5 * k5 * ic2;
	const float hs = h1 + h5, hd = cT * (h5 - h1);
	const float t0 = kr * aV;
	out[0] = 0.5f * (FL[0] + FR[0]) - 0.5f * (hs + t0);
	out[1] = 0.5f * (FL[1] + FR[1]) - 0.5f * (hs * uT + hd * nx + (uT * kr + rT * (dU - dVn * nx)) * aV);
	out[2] = 0.5f * (FL[2] + FR[2]) - 0.5f * (hs * vT + hd * ny + (vT * kr + rT * (dV - dVn * ny)) * aV);
	out[3] = 0.5f * (FL[3] + FR[3]) - 0.5f * (hs * wT + hd * nz + (wT * kr + rT * (dW - dVn * nz)) * aV);
	out[4] = 0.5f * (FL[4] + FR[4]) -
	         0.5f * (hs * hT + hd * VT + (0.5f * q2 * kr + rT * (uT * dU + vT * dV + wT * dW - VT * dVn)) * aV);
}

// AUSM (Liou & Steffen, J. Comput. Phys. 107 (1993) 23-39) minus the freestream flux.
// F = m (Phi_L + Phi_R)/2 - |m| (Phi_R - Phi_L)/2 + p (0, n, 0), Phi = rho a (1, u, H),
// m = M+(M_L) + M-(M_R), p = P+(M_L) p_L + P-(M_R) p_R. For two freestream states m = M_inf and
// p = p_inf, so F_inf = M_inf Phi_inf + p_inf (0, n, 0) and
// F - F_inf = (m - M_inf) Phi_inf + m (dPhi_L + dPhi_R)/2 - |m| (dPhi_R - dPhi_L)/2 + (p - p_inf) (0, n, 0),
// where the split-Mach and split-pressure differences are factored by M - M_inf.
__device__ __forceinline__ float mPlus(float M) { return fabsf(M) <= 1.0f ? 0.25f * (M + 1.0f) * (M + 1.0f) : 0.5f * (M + fabsf(M)); }
__device__ __forceinline__ float mMinus(float M) { return fabsf(M) <= 1.0f ? -0.25f * (M - 1.0f) * (M - 1.0f) : 0.5f * (M - fabsf(M)); }
__device__ __forceinline__ float pPlus(float M) { return fabsf(M) <= 1.0f ? 0.25f * (M + 1.0f) * (M + 1.0f) * (2.0f - M) : (M > 0.0f ? 1.0f : 0.0f); }
__device__ __forceinline__ float pMinus(float M) { return fabsf(M) <= 1.0f ? 0.25f * (M - 1.0f) * (M - 1.0f) * (2.0f + M) : (M < 0.0f ? 1.0f : 0.0f); }
// M+(M) - M+(N), M-(M) - M-(N), P+(M) - P+(N), P-(M) - P-(N) given dM = M - N
__device__ __forceinline__ float dmPlus(float M, float N, float dM) {
	return (fabsf(M) <= 1.0f && fabsf(N) <= 1.0f) ? 0.25f * dM * (M + N + 2.0f) : mPlus(M) - mPlus(N);
}
__device__ __forceinline__ float dmMinus(float M, float N, float dM) {
	return (fabsf(M) <= 1.0f && fabsf(N) <= 1.0f) ? -0.25f * dM * (M + N - 2.0f) : mMinus(M) - mMinus(N);
}
__device__ __forceinline__ float dpPlus(float M, float N, float dM) {
	return (fabsf(M) <= 1.0f && fabsf(N) <= 1.0f) ? 0.25f * dM * (3.0f - (M * M + M * N + N * N)) : pPlus(M) - pPlus(N);
}
__device__ __forceinline__ float dpMinus(float M, float N, float dM) {
	return (fabsf(M) <= 1.0f && fabsf(N) <= 1.0f) ? 0.25f * dM * ((M * M + M * N + N * N) - 3.0f) : pMinus(M) - pMinus(N);
}

__device__ __forceinline__ void ausmDelta(const D &L, const D &R, float nx, float ny, float nz, const Free32 &f,
                                          float out[5]) {
#pragma clang fp contract(fast)
	const float aInf = fsqrt(f.g * f.p * frcp(f.r));
	const float Vinf = f.u * nx + f.v * ny + f.w * nz;
	const float Minf = Vinf * frcp(aInf);
	auto side = [&](const D &d, float &a, float &M, float &dM, float &dra) {
		const float rho = f.r + d.r;
		const float da = f.g * (d.p * f.r - f.p * d.r) * frcp(rho * f.r) ; // a^2 - a_inf^2
		a = fsqrt(f.g * (f.p + d.p) * frcp(rho));
		const float dA = da * frcp(a + aInf);                             // a - a_inf
		const float dVn = d.u * nx + d.v * ny + d.w * nz;
		M = (Vinf + dVn) * frcp(a);
		dM = (dVn * aInf - Vinf * dA) * frcp(a * aInf);
		dra = d.r * a + f.r * dA;                                          // rho a - rho_inf a_inf
	};
	float aL, ML, dML, draL, aR, MR, dMR, draR;
	side(L, aL, ML, dML, draL);
	side(R, aR, MR, dMR, draR);
	const float m = mPlus(ML) + mMinus(MR);
	const float dm = dmPlus(ML, Minf, dML) + dmMinus(MR, Minf, dMR);
	const float dp = pPlus(ML) * L.p + pMinus(MR) * R.p + f.p * (dpPlus(ML, Minf, dML) + dpMinus(MR, Minf, dMR));
	const float raInf = f.r * aInf;
	// dPhi = d(rho a) (1, u, H) + rho_inf a_inf (0, du, dH)
	const float uL = f.u + L.u, vL = f.v + L.v, wL = f.w + L.w, hL = f.h + L.H;
	const float uR = f.u + R.u, vR = f.v + R.v, wR = f.w + R.w, hR = f.h + R.H;
	const float PL[5] = {draL, draL * uL + raInf * L.u, draL * vL + raInf * L.v, draL * wL + raInf * L.w, draL * hL + raInf * L.H};
	const float PR[5] = {draR, draR * uR + raInf * R.u, draR * vR + raInf * R.v, draR * wR + raInf * R.w, draR * hR + raInf * R.H};
	const float Pinf[5] = {raInf, raInf * f.u, raInf * f.v, raInf * f.w, raInf * f.h};
	const float pn[5] = {0.0f, dp * nx, dp * ny, dp * nz, 0.0f};
	const float am = fabsf(m);
#pragma unroll
	for (int i = 0; i < 5; i++) out[i] = dm * Pinf[i] + 0.5f * m * (PL[i] + PR[i]) - 0.5f * am * (PR[i] - PL[i]) + pn[i];
}

__device__ __forceinline__ void fluxDelta(int scheme, const D &L, const D &R, float nx, float ny, float nz,
                                          const Free32 &f, float out[5]) {
	if (scheme == 1) ausmDelta(L, R, nx, ny, nz, f, out);
	else roeDelta(L, R, nx, ny, nz, f, out);
}

struct GhostD {
	D d;
	float M;
};

// Wall (PhysicalBC.cpp:124-162) and symmetry (:165-231) ghost: mirrored velocity, same rho and p.
// u_g - u_inf = du_i - 2 V_i n. The face "unit" normal of a non-planar quad is N / area with
// |n| < 1 (Metrics.cpp:139), so the mirror changes |u|: |u_g|^2 = |u_i|^2 + 4 V^2 (|n|^2 - 1),
// and with it E and H (H_g - H_i = 2 V^2 (|n|^2 - 1)).
__device__ __forceinline__ GhostD mirrorD(const D &I, float nx, float ny, float nz, const Free32 &f) {
	const float rho = f.r + I.r, p = f.p + I.p;
	const float ui = f.u + I.u, vi = f.v + I.v, wi = f.w + I.w;
	const float V = ui * nx + vi * ny + wi * nz;
	const float c = fsqrt(f.g * p * frcp(rho));
	const float dn = 2.0f * V * V * ((nx * nx + ny * ny + nz * nz) - 1.0f);
	const float q2 = fmaxf(ui * ui + vi * vi + wi * wi + 2.0f * dn, 0.0f);
	const float M = fsqrt(q2) * frcp(c);
	return {{I.r, I.u - 2.0f * V * nx, I.v - 2.0f * V * ny, I.w - 2.0f * V * nz, I.p, I.H + dn}, M};
}

// Farfield ghost (EulerSolver.cpp:205-226 branch on the ghost's previous state, then
// PhysicalBC.cpp:11-121) written for perturbations. prev: the ghost's previous record; Mi: the
// interior cell's stored Mach number.
__device__ __forceinline__ GhostD farfieldD(const D &I, float Mi, const float *prev, float nx, float ny, float nz,
                                            const Free32 &f) {
	const float Vg = (f.u + prev[R_DU]) * nx + (f.v + prev[R_DV]) * ny + (f.w + prev[R_DW]) * nz;
	const float Mg = prev[R_M];
	if (Mg >= 1.0f && Vg > 0.0f) return {I, Mi};
	if (Mg >= 1.0f) return {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, f.m};
	const float ri = f.r + I.r, pi = f.p + I.p;
	const float c = fsqrt(f.g * pi * frcp(ri));
	const float irc = frcp(ri * c), ic2 = frcp(c * c);
	D g;
	if (Vg > 0.0f) {
		// subsonic outflow: p = p_inf
		g.p = 0.0f;
		g.r = I.r - I.p * ic2;
		g.u = I.u + nx * I.p * irc;
		g.v = I.v + ny * I.p * irc;
		g.w = I.w + nz * I.p * irc;
	} else {
		// subsonic inflow: p = (p_inf + p_i - rho_i c n.(u_inf - u_i)) / 2
		g.p = 0.5f * (I.p + ri * c * (nx * I.u + ny * I.v + nz * I.w));
		g.r = g.p * ic2;
		g.u = nx * g.p * irc;
		g.v = ny * g.p * irc;
		g.w = nz * g.p * irc;
	}
	g.H = deltaH(g.r, g.u, g.v, g.w, g.p, f);
	const float u = f.u + g.u, v = f.v + g.v, w = f.w + g.w;
	return {g, fsqrt(u * u + v * v + w * w) * frcp(c)};
}

__device__ __forceinline__ const float *recPtr(const float *rec, int i) { return rec + (size_t) i * kRecWidth; }

__device__ __forceinline__ void storeRec(float *rec, int i, const D &d, float M) {
	float4 *p = reinterpret_cast<float4 *>(rec + (size_t) i * kRecWidth);
	p[0] = make_float4(d.r, d.u, d.v, d.w);
	p[1] = make_float4(d.p, d.H, M, 0.0f);
}

// (case, block) pair of walk index w: each XCD takes a contiguous range of pairs; with reverse it
// walks its range from the end, so that a stage starts on the cells the previous one wrote last
__device__ __forceinline__ int logicalBlock(int w, int n, int xcds, bool reverse) {
	if (xcds <= 1) return reverse ? n - 1 - w : w;
	const int x = w % xcds, k = w / xcds;
	const int q = n / xcds, r = n % xcds;
	const int len = q + (x < r ? 1 : 0);
	return x * q + min(x, r) + (reverse ? len - 1 - k : k);
}

// deterministic workgroup sum (reduce.hpp); the trailing barrier frees lds for the caller
__device__ __forceinline__ double blockSumD(double x, double *lds) {
	const double s = groupSum<kB>(x, lds);
	__syncthreads();
	return s;
}

template<int SCHEME, bool NT>
__global__ void __launch_bounds__(kB) kMixStage(MixLaunch L, int stage, double alpha, int cap, int ms, int xcds) {
	extern __shared__ float lds[];
	float *ls = lds;            // [6][cap] perturbations of own cells, then halo cells
	float *lf = lds + 6 * cap;  // [5][ms] (F - F_inf) * S per interior slot
	const int t = threadIdx.x;
	const int nb = *(const kConst int *) L.nBlocks;
	const int per = (nb + (int) gridDim.x - 1) / (int) gridDim.x;
	for (int it = 0; it < per; it++) {
		// blocks are dealt round-robin, so each workgroup takes the same number of passes
		const int gb = it * (int) gridDim.x + (int) blockIdx.x;
		if (gb >= nb) break;
		// the tables are read-only during the launch: read them through the constant address
		// space so that the per-block descriptor lands in scalar registers
		const int lb = logicalBlock(gb, nb, xcds, (stage & 1) != 0);
		const int cslot = ((const kConst int *) L.blocks)[2 * lb], b = ((const kConst int *) L.blocks)[2 * lb + 1];
		const kConst MixCase &C = ((const kConst MixCase *) L.cases)[cslot];
		const kConst MixStatus &stt = ((const kConst MixStatus *) L.status)[cslot];
		if (!stt.active || stt.done || C.scheme != SCHEME) continue;
		const int par = (stage + stt.parity) & 1;
		const float *in = C.rec[par];
		float *out = C.rec[par ^ 1];
		const Free32 f = free32(C);
		const int c0 = b * kB;
		const int nc = min(kB, C.nElem - c0);
		const int s0 = C.blockSlot[b], ni = C.blockInt[b]; // slots, interior ones first
		const int h0 = C.blockHalo[b], nh = C.blockHalo[b + 1] - h0;
		const size_t n = (size_t) C.nElem;

		// ---- loads
		const int e = c0 + t;
		const bool own = t < nc;
		float4 ra = make_float4(0, 0, 0, 0), rb = ra;
		if (own) {
			const float4 *rp = reinterpret_cast<const float4 *>(recPtr(in, e));
			ra = rp[0], rb = rp[1];
		}
		for (int h = t; h < nh; h += kB) {
			const float4 *hp = reinterpret_cast<const float4 *>(recPtr(in, C.haloCell[h0 + h]));
			const float4 a = hp[0], bb = hp[1];
			const int l = kB + h;
			ls[0 * cap + l] = a.x, ls[1 * cap + l] = a.y, ls[2 * cap + l] = a.z, ls[3 * cap + l] = a.w;
			ls[4 * cap + l] = bb.x, ls[5 * cap + l] = bb.y;
		}
		if (own) {
			ls[0 * cap + t] = ra.x, ls[1 * cap + t] = ra.y, ls[2 * cap + t] = ra.z, ls[3 * cap + t] = ra.w;
			ls[4 * cap + t] = rb.x, ls[5 * cap + t] = rb.y;
		}
		uint32_t pkN = 0;
		float nxN = 0, nyN = 0, nzN = 0, SN = 0;
		if (t < ni) {
			pkN = ldS<NT>(C.pack + s0 + t);
			nxN = ldS<NT>(C.nx + s0 + t), nyN = ldS<NT>(C.ny + s0 + t), nzN = ldS<NT>(C.nz + s0 + t);
			SN = ldS<NT>(C.S + s0 + t);
		}
		__syncthreads();

		// ---- fluxes of the block's shared faces into LDS. The block's slots list first the faces
		// evaluated here (interior faces, and symmetry faces without a closed form); the other
		// boundary faces involve one own cell only and are formed by its thread in the gather.
		auto ld = [&](int l) -> D {
			return {ls[l], ls[cap + l], ls[2 * cap + l], ls[3 * cap + l], ls[4 * cap + l], ls[5 * cap + l]};
		};
		for (int j = t; j < ni; j += kB) {
			const int s = s0 + j;
			const uint32_t pk = pkN;
			const float nx = nxN, ny = nyN, nz = nzN, S = SN;
			if (j + kB < ni) {
				pkN = ldS<NT>(C.pack + s + kB);
				nxN = ldS<NT>(C.nx + s + kB), nyN = ldS<NT>(C.ny + s + kB), nzN = ldS<NT>(C.nz + s + kB);
				SN = ldS<NT>(C.S + s + kB);
			}
			const int l0 = (int) (pk & 0x7FFu), l1 = (int) ((pk >> 11) & 0x7FFu);
			float F[5];
			if ((pk >> 22) == 0) {
				fluxDelta(SCHEME, ld(l0), ld(l1), nx, ny, nz, f, F);
			} else {
				// symmetry face without a closed form (AUSM, or a non-planar quad): mirror ghost
				const D Lc = ld(l0);
				const int g = C.ghostOf[s];
				const GhostD G = mirrorD(Lc, nx, ny, nz, f);
				fluxDelta(SCHEME, Lc, G.d, nx, ny, nz, f, F);
			}
#pragma unroll
			for (int k = 0; k < 5; k++) lf[k * ms + j] = F[k] * S;
		}
		__syncthreads();

		// ---- gather (Kahan, FP64), local time step, FP64 RK update. The update's inputs are loaded
		// here rather than with the block's state so that they hold no registers during the
		// flux loop (other workgroups on the CU cover the latency).
		double r0 = 0.0;
		if (own) {
			uint16_t code[6];
			const uint16_t *cl = C.csr + (size_t) e * C.csrStride;
			if (C.csrStride == 4) {
				// tetrahedra: the four codes as one 8 B load (rows are 8 B aligned)
				const uint64_t c4 = ldS<NT>(reinterpret_cast<const uint64_t *>(cl));
#pragma unroll
				for (int i = 0; i < 4; i++) code[i] = (uint16_t) (c4 >> (16 * i));
				code[4] = code[5] = 0xFFFF;
			} else {
#pragma unroll
				for (int i = 0; i < 6; i++) code[i] = i < C.csrStride ? ldS<NT>(cl + i) : (uint16_t) 0xFFFF;
			}
			const D Lc = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y};
			double r[5], cmp[5];
#pragma unroll
			for (int k = 0; k < 5; k++) r[k] = 0.0, cmp[k] = 0.0;
#pragma unroll
			for (int i = 0; i < 6; i++) {
				const unsigned cd = code[i];
				if (cd == 0xFFFFu) break;
				const int j = (int) (cd & 0x7FFFu);
				const bool neg = (cd & 0x8000u) != 0;
				float x[5];
				if (j < ni) {
#pragma unroll
					for (int k = 0; k < 5; k++) x[k] = lf[k * ms + j];
				} else {
					// boundary face of this cell: ghost state (written for the next stage) and flux
					const int s = s0 + j;
					const int kind = (int) (C.pack[s] >> 22);
					const float nx = C.nx[s], ny = C.ny[s], nz = C.nz[s], S = C.S[s];
					const int g = C.ghostOf[s];
					GhostD G = kind == 3 ? farfieldD(Lc, rb.z, recPtr(in, g), nx, ny, nz, f) : mirrorD(Lc, nx, ny, nz, f);
					// a farfield ghost's state is read by the next stage's classifier; wall and
					// symmetry ghosts are rebuilt from their cell every stage and nothing reads
					// them back while iterating, so kMixMirrorGhosts forms them for read-back only
					if (kind == 3) storeRec(out, g, G.d, G.M);
					const float Vinf = f.u * nx + f.v * ny + f.w * nz;
					const float minf = f.r * Vinf;
					float F[5];
					if (kind == 1) {
						// wall: F = (0, p_g n, 0) (EulerSolver.cpp:331-339)
						F[0] = -minf;
						F[1] = G.d.p * nx - minf * f.u;
						F[2] = G.d.p * ny - minf * f.v;
						F[3] = G.d.p * nz - minf * f.w;
						F[4] = -minf * f.h;
					} else if (kind == 2) {
						// Roe flux between a state and its mirror image through a unit normal, in
						// closed form (the other symmetry faces are in the flux loop): the
						// Roe-averaged normal velocity vanishes, so only the two acoustic waves
						// remain and F = (0, (p + rho V^2 + rho c~ V) n, 0) with
						// c~^2 = c^2 + (gamma - 1) V^2 / 2 (c~ > c_L / 15: no entropy fix)
						const float rho = f.r + Lc.r;
						const float V = (f.u + Lc.u) * nx + (f.v + Lc.v) * ny + (f.w + Lc.w) * nz;
						const float ct = fsqrt(f.g * (f.p + Lc.p) * frcp(rho) + 0.5f * (f.g - 1.0f) * V * V);
						const float q = Lc.p + rho * V * (V + ct);
						F[0] = -minf;
						F[1] = q * nx - minf * f.u;
						F[2] = q * ny - minf * f.v;
						F[3] = q * nz - minf * f.w;
						F[4] = -minf * f.h;
					} else {
						// farfield: physical flux of the ghost state (Solver::Fc)
						physDelta(G.d, nx, ny, nz, Vinf, minf, f, F);
					}
#pragma unroll
					for (int k = 0; k < 5; k++) x[k] = F[k] * S;
				}
#pragma unroll
				for (int k = 0; k < 5; k++) {
					const double xd = (double) (neg ? -x[k] : x[k]);
					if (i == 0) {
						// the first term starts the sum exactly, with no compensation
						r[k] = xd;
						continue;
					}
					const double y = xd - cmp[k];
					const double tt = r[k] + y;
					cmp[k] = (tt - r[k]) - y;
					r[k] = tt;
				}
			}
			// the iteration's start state: FP64 in the last stage, which writes the new W; the
			// stages before it only form their FP32 records from the result
			double w0[5];
			if (stage == 4) {
#pragma unroll
				for (int k = 0; k < 5; k++) w0[k] = ldS<NT>(C.W + k * n + e);
			} else {
#pragma unroll
				for (int k = 0; k < 5; k++) w0[k] = C.consInf[k] + (double) ldS<NT>(C.dW0 + k * n + e);
			}
			const float sx = ldS<NT>(C.sx + e), sy = ldS<NT>(C.sy + e), sz = ldS<NT>(C.sz + e);
			if (C.defect) {
				// F_inf through the cell's unclosed remainder (exact freestream flux, FP64)
				const double cx = C.defect[3 * (size_t) e], cy = C.defect[3 * (size_t) e + 1], cz = C.defect[3 * (size_t) e + 2];
				const double V = C.uInfD * cx + C.vInfD * cy + C.wInfD * cz;
				const double m = C.rInfD * V;
				r[0] += m;
				r[1] += m * C.uInfD + C.pInfD * cx;
				r[2] += m * C.vInfD + C.pInfD * cy;
				r[3] += m * C.wInfD + C.pInfD * cz;
				r[4] += m * C.hInfD;
			}
			// TimeStepMethod2 (TimeIntegration.cpp:8-40) on the FP32 state
			const float rho = f.r + ra.x, u = f.u + ra.y, v = f.v + ra.z, w = f.w + ra.w, p = f.p + rb.x;
			const float c = fsqrt(f.g * p * frcp(rho));
			const float lam = 0.5f * (fabsf(u) + c) * sx + 0.5f * (fabsf(v) + c) * sy + 0.5f * (fabsf(w) + c) * sz;
			// W_k = W0 - ((alpha dt) R) / V (EulerSolver.cpp:441-449) with dt = CFL V / lambda: the
			// volume cancels, so the factor is alpha CFL / lambda (FP32 to FP64), and the update FP64
			const double ad = alpha * (double) (C.cfl * frcp(lam));
			double nw[5];
#pragma unroll
			for (int k = 0; k < 5; k++) nw[k] = w0[k] - ad * r[k];
			// 1 / rho for the FP32 record only: the hardware reciprocal and one Newton step (about
			// 2^-50 relative), not the IEEE division
			double ir = __builtin_amdgcn_rcp(nw[0]);
			ir = fma(ir, fma(-nw[0], ir, 1.0), ir);
			const double nu = nw[1] * ir, nv = nw[2] * ir, nwv = nw[3] * ir, nE = nw[4] * ir;
			const double gd = C.gammaD;
			const double q2 = nu * nu + nv * nv + nwv * nwv;
			const double np = (gd - 1.0) * nw[0] * (nE - 0.5 * q2);
			const double nH = nE + np * ir;
			const D d = {(float) (nw[0] - C.rInfD), (float) (nu - C.uInfD), (float) (nv - C.vInfD), (float) (nwv - C.wInfD),
			             (float) (np - C.pInfD), (float) (nH - C.hInfD)};
			// stored Mach number c / |u| as FlowField::updateWRungeKutta keeps it (FlowField.cpp:157)
			const float M = fsqrt(f.g * (float) np * frcp((float) nw[0])) * __builtin_amdgcn_rsqf((float) q2);
			storeRec(out, e, d, M);
			if (stage == 4) {
#pragma unroll
				for (int k = 0; k < 5; k++) {
					C.W[k * n + e] = nw[k];
					C.dW0[k * n + e] = (float) (nw[k] - C.consInf[k]);
				}
			}
			r0 = r[0];
		}
		if (stage == 4) {
			const double tot = blockSumD(r0 * r0, reinterpret_cast<double *>(lds + 6 * cap + 5 * ms));
			if (t == 0) C.partial[b] = tot;
		}
		__syncthreads();
	}
}

// one workgroup per case slot: RMS of the last stage (fixed-order sum of the block partials),
// residual history, convergence, the stall test and buffer parity.
// Stall test (stallWindow > 0): iterations are cut into windows of stallWindow from the first; at
// the end of a window its lowest rms is compared with the previous window's. A case whose lowest rms
// has stopped falling (not below stallRatio x the previous) while below stallLevel x the first rms
// retires as stalled: the FP32 fluxes' rounding floor, which the FP64 path can take further.
constexpr int kEndThreads = 1024;
__global__ void __launch_bounds__(kEndThreads) kMixIterEnd(const MixCase *cases, MixStatus *status) {
	__shared__ double red[kEndThreads / 64];
	const int slot = blockIdx.x;
	MixStatus &st = status[slot];
	if (!st.active || st.done) return;
	const MixCase &C = cases[slot];
	double acc = 0.0;
	for (int i = threadIdx.x; i < C.nBlocks; i += kEndThreads) acc += C.partial[i];
	const double sum = groupSum<kEndThreads>(acc, red);
	if (threadIdx.x == 0) {
		const double rms = sqrt(sum / C.nElem);
		const int it = st.iter;
		if (it < C.histCap) C.hist[it] = rms;
		if (it == 0) st.rms0 = rms;
		const double r0 = it == 0 ? rms : st.rms0;
		st.rmsLast = rms;
		st.iter = it + 1;
		st.parity ^= 1;
		// a NaN or infinite residual retires the case as failed; it never counts as converged
		const bool bad = !(rms - rms == 0.0);
		const bool conv = (C.drop > 0.0 && rms < C.drop * r0) || rms < C.minResidual;
		bool stalled = false;
		if (C.stallWindow > 0) {
			const double wm = it == st.winStart ? rms : fmin(st.winMin, rms);
			st.winMin = wm;
			if (it + 1 - st.winStart == C.stallWindow) {
				stalled = st.prevMin > 0.0 && !(wm < C.stallRatio * st.prevMin) && wm < C.stallLevel * r0;
				st.prevMin = wm;
				st.winStart = it + 1;
			}
		}
		if (bad) st.done = 3;
		else if (conv) st.done = 1;
		else if (it + 1 >= C.maxIter) st.done = 2;
		else if (stalled) st.done = 4;
	}
}

// Wall and symmetry ghosts as the last stage built them, for read-back (the stage kernel does not
// store them): the mirror of the cell state that stage started from, rec[par ^ 1], written into
// the current buffer rec[par]. One workgroup per block; a boundary face's cell is an own cell.
__global__ void __launch_bounds__(kB) kMixMirrorGhosts(MixCase C, int par) {
	const int b = blockIdx.x;
	const int s0 = C.blockSlot[b], ns = C.blockSlot[b + 1] - s0;
	const Free32 f = {C.rInf, C.uInf, C.vInf, C.wInf, C.pInf, C.hInf, C.mInf, C.gamma, C.gamma / (C.gamma - 1.0f)};
	const float *in = C.rec[par ^ 1];
	float *out = C.rec[par];
	for (int j = threadIdx.x; j < ns; j += kB) {
		const int s = s0 + j;
		const uint32_t pk = C.pack[s];
		const int kind = (int) (pk >> 22);
		if (kind != 1 && kind != 2) continue;
		const float *r = recPtr(in, b * kB + (int) (pk & 0x7FFu));
		const D Lc = {r[R_DRHO], r[R_DU], r[R_DV], r[R_DW], r[R_DP], r[R_DH]};
		const GhostD G = mirrorD(Lc, C.nx[s], C.ny[s], C.nz[s], f);
		storeRec(out, C.ghostOf[s], G.d, G.M);
	}
}

// FP64 primitives of the interior cells from W
__global__ void __launch_bounds__(256) kMixPrimitives(MixCase C, double *out) {
	const int e = blockIdx.x * 256 + threadIdx.x;
	if (e >= C.nElem) return;
	const size_t n = (size_t) C.nElem;
	const double r = C.W[e], u = C.W[n + e] / r, v = C.W[2 * n + e] / r, w = C.W[3 * n + e] / r, E = C.W[4 * n + e] / r;
	const double g = C.gammaD;
	const double q2 = u * u + v * v + w * w;
	const double p = (g - 1.0) * r * (E - 0.5 * q2);
	out[e] = r, out[n + e] = u, out[2 * n + e] = v, out[3 * n + e] = w;
	out[4 * n + e] = p, out[5 * n + e] = E, out[6 * n + e] = E + p / r;
	out[7 * n + e] = sqrt(g * (p / r)) / sqrt(q2);
}

// AeroCoefficients::Update (AeroCoefficients.cpp:38-52) with p of the wall cell from the FP64 W
__global__ void __launch_bounds__(256) kMixAero(int nWall, const int *wallCell, const double *Nx, const double *Ny,
                                                const double *Nz, const double *Cx, const double *Cy, const double *Cz,
                                                const MixCase *dc, double cpFactor, double pInf, double cref,
                                                double sref, double *out6) {
	__shared__ double red[256 / 64];
	const MixCase &C = *dc;
	const size_t n = (size_t) C.nElem;
	double acc[6] = {0, 0, 0, 0, 0, 0};
	const int per = (nWall + 255) / 256;
	const int b = threadIdx.x * per, en = min(nWall, b + per);
	for (int i = b; i < en; i++) {
		const int e = wallCell[i];
		const double r = C.W[e], u = C.W[n + e] / r, v = C.W[2 * n + e] / r, w = C.W[3 * n + e] / r;
		const double E = C.W[4 * n + e] / r;
		const double p = (C.gammaD - 1.0) * r * (E - 0.5 * (u * u + v * v + w * w));
		const double cp = cpFactor * (p - pInf);
		const double cx = Nx[i] * cp * cref / sref, cy = Ny[i] * cp * cref / sref, cz = Nz[i] * cp * cref / sref;
		acc[0] += cx, acc[1] += cy, acc[2] += cz;
		acc[3] += cx * Cx[i], acc[4] += cy * Cy[i], acc[5] += cz * Cz[i];
	}
	for (int k = 0; k < 6; k++) {
		const double t = groupSum<256>(acc[k], red);
		if (threadIdx.x == 0) out6[k] = t;
		__syncthreads();
	}
}

}// namespace

size_t mixLdsBytes(int maxSlots, int maxHalo) {
	size_t words = (size_t) 6 * (kB + maxHalo) + (size_t) 5 * maxSlots;
	return sizeof(float) * words + sizeof(double) * kB; // + the block reduction
}

void launchMixStage(const MixLaunch &L, int stage, const MixTuning &t, hipStream_t st) {
	static const double alphas[5] = {0.0533, 0.1263, 0.2375, 0.4414, 1.0};
	if (L.grid <= 0) return;
	const size_t lds = mixLdsBytes(L.ms, L.cap - kB);
#define E3D_MIX(S, N) \
	hipLaunchKernelGGL((kMixStage<S, N>), dim3(L.grid), dim3(kB), lds, st, L, stage, alphas[stage], L.cap, L.ms, t.xcds)
	for (int scheme = 0; scheme < 2; scheme++) {
		if (!(L.schemes & (1 << scheme))) continue;
		if (scheme == 1) {
			if (t.nontemporal) E3D_MIX(1, true);
			else E3D_MIX(1, false);
		} else {
			if (t.nontemporal) E3D_MIX(0, true);
			else E3D_MIX(0, false);
		}
	}
#undef E3D_MIX
}

int mixResidentPerCU(int ms, int cap, int schemes, bool nontemporal) {
	const size_t lds = mixLdsBytes(ms, cap - kB);
	int best = 1 << 30;
	for (int scheme = 0; scheme < 2; scheme++) {
		if (!(schemes & (1 << scheme))) continue;
		int n = 0;
		hipError_t e;
		if (scheme == 1)
			e = nontemporal ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kMixStage<1, true>, kB, lds)
			                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kMixStage<1, false>, kB, lds);
		else
			e = nontemporal ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kMixStage<0, true>, kB, lds)
			                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kMixStage<0, false>, kB, lds);
		if (e != hipSuccess || n <= 0) n = 1;
		best = n < best ? n : best;
	}
	return best == (1 << 30) ? 1 : best;
}

void launchMixIterEnd(const MixLaunch &L, hipStream_t st) {
	if (L.nSlots <= 0) return;
	hipLaunchKernelGGL(kMixIterEnd, dim3(L.nSlots), dim3(kEndThreads), 0, st, L.cases, L.status);
}

void launchMixMirrorGhosts(const MixCase &hc, int parity, hipStream_t st) {
	if (hc.nBlocks <= 0 || hc.nGhost <= 0) return;
	hipLaunchKernelGGL(kMixMirrorGhosts, dim3(hc.nBlocks), dim3(kB), 0, st, hc, parity);
}

void launchMixPrimitives(const MixCase &hc, double *out, hipStream_t st) {
	if (hc.nElem <= 0) return;
	hipLaunchKernelGGL(kMixPrimitives, dim3((hc.nElem + 255) / 256), dim3(256), 0, st, hc, out);
}

void launchMixAero(int nWall, const int *wallCell, const double *Nx, const double *Ny, const double *Nz,
                   const double *Cx, const double *Cy, const double *Cz, const MixCase *dc, double cpFactor,
                   double pInf, double cref, double sref, double *out6, hipStream_t st) {
	hipLaunchKernelGGL(kMixAero, dim3(1), dim3(256), 0, st, nWall, wallCell, Nx, Ny, Nz, Cx, Cy, Cz, dc, cpFactor,
	                   pInf, cref, sref, out6);
}

}// namespace e3d
