// Mixed-precision multi-case engine behind the e3d_batch_* entry points: device memory of the
// resident cases, the case and block tables the stage kernel walks, admission against
// the HBM budget, stream-ordered retirement, and the captured HIP graph of one iteration.
#include "abi_util.hpp"
#include "mixed.hpp"
#include "plan.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

using e3d_abi::fail;
using e3d_abi::fillFree;
using e3d_abi::guarded;
using e3d_abi::hipCheck;

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t) 255; }

constexpr int kMaxClasses = 32; // (scheme, workgroups per CU) pairs
constexpr int kGraphIters = 8;  // iterations of the longer captured graph

// sizes of one case's device arrays, in the order they are carved from its allocation
struct CaseLayout {
	size_t nB, nH, nS, nE, nG, stride, hist, nW, nDefect = 0;
	size_t off[25];
	size_t total = 0;
	void build() {
		const size_t sz[] = {
		    (nB + 1) * 4, (nB + 1) * 4, std::max<size_t>(nH, 1) * 4, nS * 4, nS * 4,  // blockSlot blockHalo haloCell ghostOf pack
		    nS * 4, nS * 4, nS * 4, nS * 4,                                           // nx ny nz S
		    nE * stride * 2,                                                          // csr
		    0, nE * 4, nE * 4, nE * 4,                                                // (unused) sx sy sz
		    5 * nE * 8,                                                               // W
		    (nE + nG) * e3d::kRecWidth * 4, (nE + nG) * e3d::kRecWidth * 4,           // rec[2]
		    std::max<size_t>(nB, 1) * 8, std::max<size_t>(hist, 1) * 8,               // partial hist
		    std::max<size_t>(nW, 1) * 4, std::max<size_t>(nW, 1) * 8 * 6,             // wallCell wN wC
		    8 * 8,                                                                    // scalars
		    std::max<size_t>(nDefect, 1) * 12,                                        // defect
		    std::max<size_t>(nB, 1) * 4,                                              // blockInt
		    5 * nE * 4};                                                              // dW0
		total = 0;
		for (size_t i = 0; i < sizeof(sz) / sizeof(sz[0]); i++) {
			off[i] = total;
			total += align256(sz[i]);
		}
	}
};

struct CaseHost {
	bool used = false;    // slot taken (the case may still be uploading)
	bool ready = false;   // uploaded; part of the device tables from the next activation on
	int nElem = 0, nGhost = 0, nBlocks = 0, maxSlots = 0, maxHalo = 0;
	e3d_params prm{};
	e3d::DevFree fr{};
	std::vector<int> new2old;
	char *mem = nullptr;
	size_t bytes = 0;
	int nWall = 0;
	int *wallCell = nullptr;
	double *wN[3] = {nullptr, nullptr, nullptr}, *wC[3] = {nullptr, nullptr, nullptr};
	double *scalars = nullptr;
	double cpFactor = 0.0;
	e3d::MixCase desc{};
};

}// namespace

struct e3d_batch {
	int device = 0;
	hipStream_t stream = nullptr;
	e3d::MixTuning tune{};
	bool useGraph = true;
	int nSlots = 0, nCU = 0;
	double budget = 0, used = 0;
	// device tables
	e3d::MixCase *dCases = nullptr;
	e3d::MixStatus *dStatus = nullptr;
	int2 *dBlocks = nullptr;
	int *dNBlocks = nullptr; // one count per launch class
	size_t blocksCap = 0;
	// pinned host staging
	e3d::MixCase *hCases = nullptr;
	e3d::MixStatus *hStatusInit = nullptr, *hStatus = nullptr;
	int2 *hBlocks = nullptr;
	size_t hBlocksCap = 0;
	int *hN = nullptr;       // kMaxClasses counts
	std::vector<CaseHost> cs;
	std::vector<int> statusInit; // slots whose status must be (re)written at the next activation
	bool dirty = false;
	// guards the host-side tables: admissions may run on other threads while iterations run. No
	// call waits for queued iterations while holding it, so an admission is never held up by them.
	std::mutex mu;
	// serializes the calls that read device status through the pinned hStatus (taken before mu)
	std::mutex pollMu;
	// recorded after the last table copies from the pinned staging (activate) and after a status
	// copy (poll, status, wait)
	hipEvent_t staged = nullptr, polled = nullptr;
	// launch description and the captured iteration
	// one launch per class of cases: cases are grouped by flux scheme and by how many
	// of their workgroups fit on a CU (their LDS need), so that a case with large blocks does not
	// lower the occupancy of every other case
	std::vector<e3d::MixLaunch> Ls;
	// one iteration, and kGraphIters iterations in a row (one graph launch instead of several:
	// each launch costs about 8 us between graphs, a sixth of a small case's iteration)
	hipGraph_t graph = nullptr, graphK = nullptr;
	hipGraphExec_t exec = nullptr, execK = nullptr;
	std::vector<e3d::MixLaunch> graphLs;
	bool haveGraph = false;
	// timing
	bool timing = false;
	std::vector<hipEvent_t> evPool;
	size_t evUsed = 0;

	void sync() { hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
};

namespace {

void freeGraph(e3d_batch *b) {
	if (b->exec) (void) hipGraphExecDestroy(b->exec);
	if (b->graph) (void) hipGraphDestroy(b->graph);
	if (b->execK) (void) hipGraphExecDestroy(b->execK);
	if (b->graphK) (void) hipGraphDestroy(b->graphK);
	b->exec = b->execK = nullptr;
	b->graph = b->graphK = nullptr;
	b->haveGraph = false;
}

void enqueueIteration(e3d_batch *b) {
	for (int s = 0; s < 5; s++)
		for (const auto &L : b->Ls) e3d::launchMixStage(L, s, b->tune, b->stream);
	e3d::launchMixIterEnd(b->Ls.front(), b->stream);
}

void capture(e3d_batch *b) {
	freeGraph(b);
	hipCheck(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
	enqueueIteration(b);
	hipCheck(hipStreamEndCapture(b->stream, &b->graph), "hipStreamEndCapture");
	hipCheck(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
	hipCheck(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
	for (int i = 0; i < kGraphIters; i++) enqueueIteration(b);
	hipCheck(hipStreamEndCapture(b->stream, &b->graphK), "hipStreamEndCapture");
	hipCheck(hipGraphInstantiate(&b->execK, b->graphK, nullptr, nullptr, 0), "hipGraphInstantiate");
	b->graphLs = b->Ls;
	b->haveGraph = true;
}

bool sameLaunch(const e3d::MixLaunch &a, const e3d::MixLaunch &c) {
	return a.cases == c.cases && a.status == c.status && a.blocks == c.blocks && a.nBlocks == c.nBlocks &&
	       a.nSlots == c.nSlots && a.grid == c.grid && a.cap == c.cap && a.ms == c.ms && a.schemes == c.schemes;
}

// Brings the device tables in line with the resident cases (after admissions or removals).
void activate(e3d_batch *b) {
	if (!b->dirty) return;
	// the pinned staging written below may still be read by the previous activation's copies
	if (b->staged) hipCheck(hipEventSynchronize(b->staged), "hipEventSynchronize");
	// classes: (scheme, workgroups of the case's own layout that fit on a CU)
	struct Cls {
		int scheme, perCU, maxSlots = 2, maxHalo = 0;
		size_t blocks = 0;
		std::vector<int> slots;
	};
	std::vector<Cls> cls;
	size_t total = 0;
	for (int s = 0; s < b->nSlots; s++) {
		const CaseHost &c = b->cs[s];
		if (!c.ready) continue;
		const int ms = (c.maxSlots + 1) & ~1, cap = e3d::kMixBlock + c.maxHalo;
		const int fit = e3d::mixResidentPerCU(ms, cap, 1 << c.desc.scheme, b->tune.nontemporal);
		Cls *k = nullptr;
		for (auto &x : cls)
			if (x.scheme == c.desc.scheme && x.perCU == fit) k = &x;
		if (!k) {
			if ((int) cls.size() == kMaxClasses) k = &cls.back(); // (cannot happen: 2 schemes x few sizes)
			else {
				cls.push_back(Cls{c.desc.scheme, fit});
				k = &cls.back();
			}
		}
		k->maxSlots = std::max(k->maxSlots, c.maxSlots);
		k->maxHalo = std::max(k->maxHalo, c.maxHalo);
		k->blocks += (size_t) c.nBlocks;
		k->slots.push_back(s);
		total += (size_t) c.nBlocks;
	}
	// One launch per class keeps a large-LDS case from lowering the others' occupancy, but every
	// launch costs about one round of workgroups. When a scheme's cases are few and small, the
	// stage is launch-bound and one launch with the largest layout is cheaper: the classes of a
	// scheme are merged when 1 + rounds(all blocks at the merged occupancy) is below the sum of
	// 1 + rounds(blocks) over the classes. E3D_MIX_CLASSES=merge|split forces either. The launch
	// shape changes no result.
	{
		const char *env = std::getenv("E3D_MIX_CLASSES");
		const int force = env ? (std::string(env) == "merge" ? 1 : std::string(env) == "split" ? 0 : -1) : -1;
		const size_t cu = (size_t) std::max(1, b->nCU);
		auto rounds = [&](size_t blocks, int perCU) { return (blocks + cu * perCU - 1) / (cu * perCU); };
		for (int scheme = 0; scheme < 2 && force != 0; scheme++) {
			std::vector<size_t> idx;
			for (size_t i = 0; i < cls.size(); i++)
				if (cls[i].scheme == scheme) idx.push_back(i);
			if (idx.size() < 2) continue;
			Cls m{scheme, 1};
			size_t sep = 0;
			for (size_t i : idx) {
				const Cls &k = cls[i];
				sep += 1 + rounds(k.blocks, k.perCU);
				m.maxSlots = std::max(m.maxSlots, k.maxSlots);
				m.maxHalo = std::max(m.maxHalo, k.maxHalo);
				m.blocks += k.blocks;
				m.slots.insert(m.slots.end(), k.slots.begin(), k.slots.end());
			}
			m.perCU = e3d::mixResidentPerCU((m.maxSlots + 1) & ~1, e3d::kMixBlock + m.maxHalo, 1 << scheme,
			                                b->tune.nontemporal);
			if (force != 1 && 1 + rounds(m.blocks, m.perCU) >= sep) continue;
			std::sort(m.slots.begin(), m.slots.end());
			for (size_t j = idx.size(); j-- > 0;) cls.erase(cls.begin() + (long) idx[j]);
			cls.push_back(std::move(m));
		}
	}
	if (total > b->hBlocksCap) {
		if (b->hBlocks) (void) hipHostFree(b->hBlocks);
		b->hBlocksCap = std::max<size_t>(total, 2 * b->hBlocksCap);
		hipCheck(hipHostMalloc((void **) &b->hBlocks, b->hBlocksCap * sizeof(int2)), "hipHostMalloc");
	}
	if (total > b->blocksCap) {
		if (b->dBlocks) hipCheck(hipFreeAsync(b->dBlocks, b->stream), "hipFreeAsync");
		b->blocksCap = std::max<size_t>(total, 2 * b->blocksCap);
		hipCheck(hipMallocAsync((void **) &b->dBlocks, b->blocksCap * sizeof(int2), b->stream), "hipMallocAsync");
	}
	for (int s = 0; s < b->nSlots; s++) b->hCases[s] = b->cs[s].ready ? b->cs[s].desc : e3d::MixCase{};
	std::vector<e3d::MixLaunch> Ls;
	size_t n = 0;
	const int xc = std::max(1, b->tune.xcds);
	for (size_t ci = 0; ci < cls.size(); ci++) {
		const Cls &k = cls[ci];
		const size_t first = n;
		for (int s : k.slots)
			for (int q = 0; q < b->cs[s].nBlocks; q++) b->hBlocks[n++] = make_int2(s, q);
		b->hN[ci] = (int) (n - first);
		e3d::MixLaunch L{};
		L.cases = b->dCases;
		L.status = b->dStatus;
		L.blocks = b->dBlocks + first;
		L.nBlocks = b->dNBlocks + ci;
		L.nSlots = b->nSlots;
		L.cap = e3d::kMixBlock + k.maxHalo;
		L.ms = (k.maxSlots + 1) & ~1;
		L.schemes = 1 << k.scheme;
		// wg_per_cu 0: one workgroup per (case, block) pair, left to the hardware scheduler (on the
		// 6.3M-tet bump 251.6 us per stage against 262.1 us for the best persistent grid); the grid
		// is rounded up to a multiple of 2048 (surplus workgroups exit at once) so that admissions
		// and retirements rarely change it and force a new graph. Else a persistent grid, never
		// more workgroups than are resident at once.
		const size_t gran = (size_t) std::max(xc, 2048);
		L.grid = (int) ((k.blocks + gran - 1) / gran * gran);
		if (b->tune.wgPerCU > 0) {
			const int perCU = std::min(b->tune.wgPerCU, e3d::mixResidentPerCU(L.ms, L.cap, L.schemes, b->tune.nontemporal));
			const int maxGrid = std::max(xc, b->nCU * perCU / xc * xc);
			L.grid = std::min(L.grid, maxGrid);
		}
		Ls.push_back(L);
	}
	if (n) hipCheck(hipMemcpyAsync(b->dBlocks, b->hBlocks, n * sizeof(int2), hipMemcpyHostToDevice, b->stream), "H2D blocks");
	if (!cls.empty())
		hipCheck(hipMemcpyAsync(b->dNBlocks, b->hN, sizeof(int) * cls.size(), hipMemcpyHostToDevice, b->stream),
		         "H2D nBlocks");
	hipCheck(hipMemcpyAsync(b->dCases, b->hCases, sizeof(e3d::MixCase) * b->nSlots, hipMemcpyHostToDevice, b->stream),
	         "H2D cases");
	for (int s : b->statusInit) {
		e3d::MixStatus &st = b->hStatusInit[s];
		st = e3d::MixStatus{};
		st.active = b->cs[s].ready ? 1 : 0;
		hipCheck(hipMemcpyAsync(b->dStatus + s, &st, sizeof(st), hipMemcpyHostToDevice, b->stream), "H2D status");
	}
	b->statusInit.clear();
	hipCheck(hipEventRecord(b->staged, b->stream), "hipEventRecord");
	b->Ls = std::move(Ls);
	b->dirty = false;
}

// Lists each block's interior faces first (the stage kernel evaluates only those in its block-wide
// flux loop and keeps their fluxes in LDS; a boundary face touches one own cell and is formed by
// that cell's thread). Rewrites the slot arrays and the CSR codes; returns the most interior
// slots of any block.
int interiorFirst(e3d::HostPlan &P, std::vector<int> &blockInt, int scheme) {
	// symmetry faces stay in the flux loop unless the closed form applies (Roe, unit normal as
	// the kernel sees it in FP32)
	auto inLoop = [&](size_t s) {
		if (P.skind[s] == 0) return true;
		if (P.skind[s] != 2) return false;
		const float nx = (float) P.snx[s], ny = (float) P.sny[s], nz = (float) P.snz[s];
		return scheme != 0 || !(std::fabs(nx * nx + ny * ny + nz * nz - 1.0f) < 1e-5f);
	};
	blockInt.assign(P.nBlocks, 0);
	int maxInt = 0;
	std::vector<int> newIdx;
	for (int b = 0; b < P.nBlocks; b++) {
		const int s0 = P.blockSlot[b], s1 = P.blockSlot[b + 1], ns = s1 - s0;
		newIdx.assign(ns, 0);
		int k = 0;
		for (int j = 0; j < ns; j++)
			if (inLoop(s0 + j)) newIdx[j] = k++;
		blockInt[b] = k;
		maxInt = std::max(maxInt, k);
		for (int j = 0; j < ns; j++)
			if (!inLoop(s0 + j)) newIdx[j] = k++;
		auto permute = [&](auto &v) {
			using T = typename std::decay_t<decltype(v)>::value_type;
			std::vector<T> tmp(v.begin() + s0, v.begin() + s1);
			for (int j = 0; j < ns; j++) v[s0 + newIdx[j]] = tmp[j];
		};
		permute(P.se0), permute(P.se1), permute(P.skind), permute(P.snx), permute(P.sny), permute(P.snz);
		permute(P.sS), permute(P.spack);
		const int c0 = b * P.cellsPerBlock, c1 = std::min(P.nElem, c0 + P.cellsPerBlock);
		for (size_t i = (size_t) c0 * P.csrStride; i < (size_t) c1 * P.csrStride; i++) {
			const uint16_t code = P.csr[i];
			if (code == 0xFFFF) continue;
			P.csr[i] = (uint16_t) (newIdx[code & 0x7FFF] | (code & 0x8000));
		}
	}
	return maxInt;
}

void initCaseHost(const e3d::HostPlan &P, const e3d::SolverMesh &sm, const e3d_params &prm, const e3d_case_opts &co,
                  CaseLayout &lay) {
	lay.nB = (size_t) P.nBlocks;
	lay.nH = P.haloCell.size();
	lay.nS = P.se0.size();
	lay.nE = (size_t) P.nElem;
	lay.nG = (size_t) P.nGhost;
	lay.stride = (size_t) P.csrStride;
	lay.hist = (size_t) std::max(1, co.hist_cap);
	lay.nW = P.wallCell.size();
	lay.build();
	(void) sm, (void) prm;
}

}// namespace

extern "C" {

int e3d_batch_opts_default(e3d_batch_opts *o) {
	if (!o) return fail(E3D_ERR_ARG, "null options");
	std::memset(o, 0, sizeof(*o));
	o->hbm_budget = 0.0;
	o->max_cases = 256;
	o->wg_per_cu = 0; // one workgroup per (case, block) pair
	o->xcds = 8;
	o->nontemporal = 1;
	o->use_graph = 1;
	return E3D_OK;
}

int e3d_batch_create(int device, const e3d_batch_opts *opts, e3d_batch **out) {
	if (!out) return fail(E3D_ERR_ARG, "null argument");
	e3d_batch_opts o;
	e3d_batch_opts_default(&o);
	if (opts) o = *opts;
	if (o.max_cases <= 0 || o.max_cases > 65536) return fail(E3D_ERR_ARG, "max_cases must be in 1..65536");
	if (o.wg_per_cu < 0 || o.wg_per_cu > 32) return fail(E3D_ERR_ARG, "wg_per_cu must be in 0..32");
	if (o.xcds <= 0 || o.xcds > 64) return fail(E3D_ERR_ARG, "xcds must be in 1..64");
	return guarded([&] {
		auto b = std::make_unique<e3d_batch>();
		int nDev = 0;
		hipCheck(hipGetDeviceCount(&nDev), "hipGetDeviceCount");
		if (device < 0 || device >= nDev) return fail(E3D_ERR_DEVICE, "no HIP device " + std::to_string(device));
		b->device = device;
		hipCheck(hipSetDevice(device), "hipSetDevice");
		hipCheck(hipDeviceGetAttribute(&b->nCU, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
		hipCheck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
		hipCheck(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming), "hipEventCreate");
		hipCheck(hipEventCreateWithFlags(&b->polled, hipEventDisableTiming), "hipEventCreate");
		b->tune.wgPerCU = o.wg_per_cu;
		b->tune.xcds = o.xcds;
		b->tune.nontemporal = o.nontemporal != 0;
		b->useGraph = o.use_graph != 0;
		b->nSlots = o.max_cases;
		if (o.hbm_budget > 0) {
			b->budget = o.hbm_budget;
		} else {
			size_t fr = 0, tot = 0;
			hipCheck(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
			b->budget = 0.9 * (double) fr;
		}
		b->cs.resize(b->nSlots);
		hipCheck(hipMalloc((void **) &b->dCases, sizeof(e3d::MixCase) * b->nSlots), "hipMalloc");
		hipCheck(hipMalloc((void **) &b->dStatus, sizeof(e3d::MixStatus) * b->nSlots), "hipMalloc");
		hipCheck(hipMalloc((void **) &b->dNBlocks, sizeof(int) * kMaxClasses), "hipMalloc");
		hipCheck(hipHostMalloc((void **) &b->hCases, sizeof(e3d::MixCase) * b->nSlots), "hipHostMalloc");
		hipCheck(hipHostMalloc((void **) &b->hStatusInit, sizeof(e3d::MixStatus) * b->nSlots), "hipHostMalloc");
		hipCheck(hipHostMalloc((void **) &b->hStatus, sizeof(e3d::MixStatus) * b->nSlots), "hipHostMalloc");
		hipCheck(hipHostMalloc((void **) &b->hN, sizeof(int) * kMaxClasses), "hipHostMalloc");
		std::memset(b->hCases, 0, sizeof(e3d::MixCase) * b->nSlots);
		std::memset(b->hStatusInit, 0, sizeof(e3d::MixStatus) * b->nSlots);
		hipCheck(hipMemcpy(b->dStatus, b->hStatusInit, sizeof(e3d::MixStatus) * b->nSlots, hipMemcpyHostToDevice), "H2D");
		std::fill(b->hN, b->hN + kMaxClasses, 0);
		hipCheck(hipMemcpy(b->dNBlocks, b->hN, sizeof(int) * kMaxClasses, hipMemcpyHostToDevice), "H2D");
		*out = b.release();
		return E3D_OK;
	});
}

int e3d_batch_case_bytes(const e3d_mesh *m, const e3d_case_opts *co, double *bytes) {
	if (!m || !bytes) return fail(E3D_ERR_ARG, "null argument");
	if (!m->preprocessed) return fail(E3D_ERR_ARG, "mesh is not preprocessed (call e3d_mesh_preprocess)");
	// upper bound: every face in at most two blocks, a halo no larger than the slot count
	const e3d::SolverMesh &sm = m->sm;
	int stride = 0;
	for (int e = 0; e < sm.nElem; e++) stride = std::max(stride, sm.element2faceStart[e + 1] - sm.element2faceStart[e]);
	CaseLayout lay;
	lay.nB = (size_t) (sm.nElem + e3d::kMixBlock - 1) / e3d::kMixBlock;
	lay.nS = 2 * (size_t) sm.nFace;
	lay.nH = lay.nS;
	lay.nE = (size_t) sm.nElem;
	lay.nG = (size_t) sm.nGhost;
	lay.stride = (size_t) stride;
	lay.hist = (size_t) std::max(1, co ? co->hist_cap : 1);
	lay.nW = sm.wallGhost.size();
	lay.build();
	*bytes = (double) lay.total;
	return E3D_OK;
}

int e3d_batch_add(e3d_batch *b, const e3d_mesh *m, const e3d_params *p, const e3d_case_opts *co, int *case_id) {
	if (!b || !m || !p || !co || !case_id) return fail(E3D_ERR_ARG, "null argument");
	if (!m->preprocessed) return fail(E3D_ERR_ARG, "mesh is not preprocessed (call e3d_mesh_preprocess)");
	if (p->scheme != 0 && p->scheme != 1) return fail(E3D_ERR_ARG, "SCHEME must be 0 (Roe) or 1 (AUSM)");
	if (!(p->cfl > 0) || !(p->gamma > 1)) return fail(E3D_ERR_ARG, "invalid CFL or GAMMA");
	if (p->time_integration != 2 || p->residual_smoothing)
		return fail(E3D_ERR_UNSUPPORTED, "the mixed-precision engine runs RK5 (TIME_INTEGRATION= 2) without residual "
		                                 "smoothing; use e3d_create for the other integrators");
	if (co->max_iter <= 0 || co->hist_cap < 0 || co->drop < 0) return fail(E3D_ERR_ARG, "invalid case options");
	if (co->stall_window < 0 ||
	    (co->stall_window > 0 && !(co->stall_ratio > 0 && co->stall_ratio <= 1 && co->stall_level > 0)))
		return fail(E3D_ERR_ARG, "invalid stall test (stall_window >= 0, 0 < stall_ratio <= 1, stall_level > 0)");
	if (!m->parts.empty())
		return fail(E3D_ERR_UNSUPPORTED, "the mixed-precision engine takes single meshes; use e3d_create for partitioned input");
	const e3d::SolverMesh &sm = m->sm;
	if (sm.nElem <= 0) return fail(E3D_ERR_ARG, "empty mesh");
	if ((long long) sm.nElem + sm.nGhost >= (1LL << 31) / e3d::kRecWidth) return fail(E3D_ERR_ARG, "mesh too large");
	// reserve a slot; the plan and the upload run without the lock so that iterations of the
	// resident cases continue meanwhile
	int slot = -1;
	{
		std::lock_guard<std::mutex> lk(b->mu);
		for (int s = 0; s < b->nSlots; s++)
			if (!b->cs[s].used) {
				slot = s;
				break;
			}
		if (slot < 0) return fail(E3D_ERR_BUSY, "no free case slot");
		b->cs[slot].used = true;
		b->cs[slot].ready = false;
	}
	double reserved = 0.0;
	hipStream_t ust = nullptr;
	char *mem = nullptr;
	int rc = guarded([&] {
		hipCheck(hipSetDevice(b->device), "hipSetDevice");
		int order = e3d::ORDER_BISECT; // box-shaped blocks cut fewer faces than Morton runs
		if (const char *v = std::getenv("E3D_CELL_ORDER")) {
			if (std::string(v) == "morton") order = e3d::ORDER_MORTON;
			if (std::string(v) == "reference") order = e3d::ORDER_REFERENCE;
		}
		e3d::HostPlan P = e3d::buildPlan(sm, order, e3d::kMixBlock);
		if (P.csrStride > 6) return fail(E3D_ERR_UNSUPPORTED, "cells with more than 6 faces are not on the device path");
		std::vector<int> blockInt;
		const int maxInt = interiorFirst(P, blockInt, p->scheme);
		if (e3d::mixLdsBytes((maxInt + 1) & ~1, P.maxHalo) > 160 * 1024)
			return fail(E3D_ERR_UNSUPPORTED, "a block of this mesh needs more LDS than a CU has");
		// closure of every cell: sum over its faces of +-n S, FP64, in reference face order
		std::vector<float> defect;
		{
			double worst = 0.0;
			std::vector<double> cd(3 * (size_t) P.nElem);
			for (int c = 0; c < P.nElem; c++) {
				double a[3] = {0, 0, 0}, area = 0;
				for (int j = P.fstart[c]; j < P.fstart[c + 1]; j++) {
					const unsigned code = (unsigned) P.flist[j];
					const int f = (int) (code & 0x7fffffffu);
					const double sg = (code >> 31) ? -1.0 : 1.0;
					a[0] += sg * P.fnx[f] * P.fS[f];
					a[1] += sg * P.fny[f] * P.fS[f];
					a[2] += sg * P.fnz[f] * P.fS[f];
					area += P.fS[f];
				}
				for (int k = 0; k < 3; k++) cd[3 * (size_t) c + k] = a[k];
				worst = std::max(worst, std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) / area);
			}
			if (worst > 1e-12) {
				defect.resize(cd.size());
				for (size_t k = 0; k < cd.size(); k++) defect[k] = (float) cd[k];
			}
		}
		CaseLayout lay;
		lay.nDefect = defect.size() / 3;
		initCaseHost(P, sm, *p, *co, lay);
		{
			std::lock_guard<std::mutex> lk(b->mu);
			if ((double) lay.total > b->budget)
				return fail(E3D_ERR_ARG, "the case needs " + std::to_string(lay.total) +
				                             " B, more than the whole HBM budget; it can never be admitted");
			if (b->used + (double) lay.total > b->budget)
				return fail(E3D_ERR_BUSY, "HBM budget exhausted (" + std::to_string(lay.total) + " B requested)");
			b->used += (double) lay.total;
			reserved = (double) lay.total;
		}
		hipCheck(hipStreamCreateWithFlags(&ust, hipStreamNonBlocking), "hipStreamCreate");
		CaseHost c;
		c.used = true;
		c.nElem = P.nElem;
		c.nGhost = P.nGhost;
		c.nBlocks = P.nBlocks;
		c.maxSlots = maxInt; // LDS holds the interior slots only
		c.maxHalo = P.maxHalo;
		c.prm = *p;
		fillFree(*p, c.fr);
		c.new2old = P.new2old;
		c.bytes = lay.total;
		hipCheck(hipMallocAsync((void **) &c.mem, lay.total, ust), "hipMallocAsync");
		mem = c.mem;
		char *base = c.mem;
		auto at = [&](int i) { return base + lay.off[i]; };
		auto up = [&](int i, const void *h, size_t bytes) {
			if (bytes) hipCheck(hipMemcpyAsync(at(i), h, bytes, hipMemcpyHostToDevice, ust), "H2D case");
		};
		const size_t nS = lay.nS, nE = lay.nE, nG = lay.nG;
		// FP32 copies of the face geometry and cell metrics
		std::vector<float> f32(std::max(nS, nE));
		auto upF = [&](int i, const std::vector<double> &v) {
			for (size_t k = 0; k < v.size(); k++) f32[k] = (float) v[k];
			up(i, f32.data(), v.size() * 4);
			hipCheck(hipStreamSynchronize(ust), "hipStreamSynchronize(upload)"); // f32 is reused
		};
		up(0, P.blockSlot.data(), P.blockSlot.size() * 4);
		up(1, P.blockHalo.data(), P.blockHalo.size() * 4);
		up(2, P.haloCell.data(), P.haloCell.size() * 4);
		up(3, P.se1.data(), nS * 4);
		up(4, P.spack.data(), nS * 4);
		upF(5, P.snx);
		upF(6, P.sny);
		upF(7, P.snz);
		upF(8, P.sS);
		up(9, P.csr.data(), P.csr.size() * 2);
		upF(11, P.sx);
		upF(12, P.sy);
		upF(13, P.sz);
		// freestream start (FlowField::Initialize): W and the FP32 records (zero perturbation,
		// stored Mach number M_inf)
		const e3d::DevFree &q = c.fr;
		const double wInf[5] = {q.rho_inf, q.rho_inf * q.u_inf, q.rho_inf * q.v_inf, q.rho_inf * q.w_inf, q.rho_inf * q.E_inf};
		{
			std::vector<double> W(5 * nE);
			for (int k = 0; k < 5; k++) std::fill(W.begin() + k * nE, W.begin() + (k + 1) * nE, wInf[k]);
			up(14, W.data(), W.size() * 8);
			const std::vector<float> zero(5 * nE, 0.0f);
			up(24, zero.data(), zero.size() * 4);
			std::vector<float> rec((nE + nG) * e3d::kRecWidth, 0.0f);
			for (size_t i = 0; i < nE + nG; i++) rec[i * e3d::kRecWidth + e3d::R_M] = (float) q.M_inf;
			up(15, rec.data(), rec.size() * 4);
			up(16, rec.data(), rec.size() * 4);
			hipCheck(hipStreamSynchronize(ust), "hipStreamSynchronize(upload)");
		}
		c.nWall = (int) lay.nW;
		c.wallCell = (int *) at(19);
		for (int k = 0; k < 3; k++) {
			c.wN[k] = (double *) at(20) + k * std::max<size_t>(lay.nW, 1);
			c.wC[k] = (double *) at(20) + (3 + k) * std::max<size_t>(lay.nW, 1);
		}
		c.scalars = (double *) at(21);
		up(23, blockInt.data(), blockInt.size() * 4);
		if (!defect.empty()) up(22, defect.data(), defect.size() * 4);
		if (lay.nW) {
			up(19, P.wallCell.data(), lay.nW * 4);
			for (int k = 0; k < 3; k++) {
				hipCheck(hipMemcpyAsync(c.wN[k], P.wallN[k].data(), lay.nW * 8, hipMemcpyHostToDevice, ust), "H2D");
				hipCheck(hipMemcpyAsync(c.wC[k], P.wallC[k].data(), lay.nW * 8, hipMemcpyHostToDevice, ust), "H2D");
			}
		}
		hipCheck(hipStreamSynchronize(ust), "hipStreamSynchronize(upload)");
		double uinf2 = q.u_inf * q.u_inf + q.v_inf * q.v_inf + q.w_inf * q.w_inf;
		c.cpFactor = 1 / (0.5 * q.rho_inf * uinf2);

		e3d::MixCase &d = c.desc;
		d = e3d::MixCase{};
		d.nElem = P.nElem;
		d.nGhost = P.nGhost;
		d.nBlocks = P.nBlocks;
		d.csrStride = P.csrStride;
		d.scheme = p->scheme;
		d.histCap = (int) lay.hist;
		d.maxIter = co->max_iter;
		d.drop = co->drop;
		d.minResidual = p->min_residual;
		d.stallWindow = co->stall_window;
		d.stallRatio = co->stall_ratio;
		d.stallLevel = co->stall_level;
		d.blockSlot = (const int *) at(0);
		d.blockInt = (const int *) at(23);
		d.blockHalo = (const int *) at(1);
		d.haloCell = (const int *) at(2);
		d.ghostOf = (const int *) at(3);
		d.pack = (const uint32_t *) at(4);
		d.nx = (const float *) at(5);
		d.ny = (const float *) at(6);
		d.nz = (const float *) at(7);
		d.S = (const float *) at(8);
		d.csr = (const uint16_t *) at(9);
		d.sx = (const float *) at(11);
		d.sy = (const float *) at(12);
		d.sz = (const float *) at(13);
		d.defect = defect.empty() ? nullptr : (const float *) at(22);
		d.W = (double *) at(14);
		d.dW0 = (float *) at(24);
		for (int k = 0; k < 5; k++) d.consInf[k] = wInf[k];
		d.rec[0] = (float *) at(15);
		d.rec[1] = (float *) at(16);
		d.partial = (double *) at(17);
		d.hist = (double *) at(18);
		d.gamma = (float) q.gamma;
		d.cfl = (float) q.cfl;
		d.rInf = (float) q.rho_inf;
		d.uInf = (float) q.u_inf;
		d.vInf = (float) q.v_inf;
		d.wInf = (float) q.w_inf;
		d.pInf = (float) q.p_inf;
		d.hInf = (float) q.H_inf;
		d.mInf = (float) q.M_inf;
		d.gammaD = q.gamma;
		d.rInfD = q.rho_inf;
		d.uInfD = q.u_inf;
		d.vInfD = q.v_inf;
		d.wInfD = q.w_inf;
		d.pInfD = q.p_inf;
		d.hInfD = q.H_inf;
		c.ready = true;
		{
			std::lock_guard<std::mutex> lk(b->mu);
			b->cs[slot] = std::move(c);
			b->statusInit.push_back(slot);
			b->dirty = true;
		}
		mem = nullptr;
		*case_id = slot;
		return E3D_OK;
	});
	if (ust) {
		if (mem) {
			(void) hipFreeAsync(mem, ust);
			(void) hipStreamSynchronize(ust);
		}
		(void) hipStreamDestroy(ust);
	}
	if (rc != E3D_OK) {
		std::lock_guard<std::mutex> lk(b->mu);
		b->used -= reserved;
		b->cs[slot] = CaseHost{};
	}
	return rc;
}

int e3d_batch_run(e3d_batch *b, int n) {
	if (!b || n < 0) return fail(E3D_ERR_ARG, "bad argument");
	return guarded([&] {
		std::lock_guard<std::mutex> lk(b->mu);
		hipCheck(hipSetDevice(b->device), "hipSetDevice");
		activate(b);
		if (b->Ls.empty() || n == 0) return E3D_OK;
		bool same = b->haveGraph && b->graphLs.size() == b->Ls.size();
		for (size_t i = 0; same && i < b->Ls.size(); i++) same = sameLaunch(b->graphLs[i], b->Ls[i]);
		if (b->useGraph && !same) capture(b);
		auto ev = [&]() {
			if (b->evUsed >= b->evPool.size()) {
				for (int i = 0; i < 64; i++) {
					hipEvent_t e;
					hipCheck(hipEventCreate(&e), "hipEventCreate");
					b->evPool.push_back(e);
				}
			}
			return b->evPool[b->evUsed++];
		};
		int i = 0;
		// whole groups of iterations as one graph launch (not when each iteration is timed)
		if (b->useGraph && !b->timing)
			for (; i + kGraphIters <= n; i += kGraphIters) hipCheck(hipGraphLaunch(b->execK, b->stream), "hipGraphLaunch");
		for (; i < n; i++) {
			if (b->useGraph) {
				hipEvent_t a = b->timing ? ev() : nullptr, z = b->timing ? ev() : nullptr;
				if (a) hipCheck(hipEventRecord(a, b->stream), "hipEventRecord");
				hipCheck(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
				if (z) hipCheck(hipEventRecord(z, b->stream), "hipEventRecord");
			} else {
				for (int s = 0; s < 5; s++) {
					hipEvent_t a = b->timing ? ev() : nullptr, z = b->timing ? ev() : nullptr;
					if (a) hipCheck(hipEventRecord(a, b->stream), "hipEventRecord");
					for (const auto &L : b->Ls) e3d::launchMixStage(L, s, b->tune, b->stream);
					if (z) hipCheck(hipEventRecord(z, b->stream), "hipEventRecord");
				}
				e3d::launchMixIterEnd(b->Ls.front(), b->stream);
			}
		}
		hipCheck(hipGetLastError(), "kernel launch");
		return E3D_OK;
	});
}

int e3d_batch_wait(e3d_batch *b) {
	if (!b) return fail(E3D_ERR_ARG, "null batch");
	return guarded([&] {
		std::lock_guard<std::mutex> pk(b->pollMu);
		{
			std::lock_guard<std::mutex> lk(b->mu);
			hipCheck(hipSetDevice(b->device), "hipSetDevice");
			hipCheck(hipEventRecord(b->polled, b->stream), "hipEventRecord");
		}
		hipCheck(hipEventSynchronize(b->polled), "hipEventSynchronize");
		return E3D_OK;
	});
}

int e3d_batch_status(e3d_batch *b, int id, int *done, int *iters, double *rms0, double *rmsLast) {
	if (!b || id < 0 || id >= b->nSlots) return fail(E3D_ERR_ARG, "no such case");
	return guarded([&] {
		std::lock_guard<std::mutex> pk(b->pollMu);
		e3d::MixStatus *h = b->hStatus + id;
		{
			std::lock_guard<std::mutex> lk(b->mu);
			if (!b->cs[id].ready) return fail(E3D_ERR_ARG, "no such case");
			hipCheck(hipSetDevice(b->device), "hipSetDevice");
			activate(b);
			hipCheck(hipMemcpyAsync(h, b->dStatus + id, sizeof(*h), hipMemcpyDeviceToHost, b->stream), "D2H status");
			hipCheck(hipEventRecord(b->polled, b->stream), "hipEventRecord");
		}
		hipCheck(hipEventSynchronize(b->polled), "hipEventSynchronize");
		if (done) *done = h->done;
		if (iters) *iters = h->iter;
		if (rms0) *rms0 = h->rms0;
		if (rmsLast) *rmsLast = h->rmsLast;
		if (h->done == E3D_CASE_FAILED)
			return fail(E3D_ERR_NUMERIC, "case " + std::to_string(id) + ": non-finite residual at iteration " +
			                                 std::to_string(h->iter));
		return E3D_OK;
	});
}

int e3d_batch_poll(e3d_batch *b, int *ids, int *done, int *iters, double *rmsLast, int cap, int *n) {
	if (!b || !n) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		std::lock_guard<std::mutex> pk(b->pollMu);
		{
			std::lock_guard<std::mutex> lk(b->mu);
			hipCheck(hipSetDevice(b->device), "hipSetDevice");
			activate(b);
			hipCheck(hipMemcpyAsync(b->hStatus, b->dStatus, sizeof(e3d::MixStatus) * b->nSlots, hipMemcpyDeviceToHost,
			                        b->stream),
			         "D2H status");
			hipCheck(hipEventRecord(b->polled, b->stream), "hipEventRecord");
		}
		// the copy waits for every iteration queued so far; admissions go on meanwhile
		hipCheck(hipEventSynchronize(b->polled), "hipEventSynchronize");
		std::lock_guard<std::mutex> lk(b->mu);
		int k = 0;
		for (int s = 0; s < b->nSlots; s++) {
			if (!b->cs[s].ready) continue;
			if (k < cap) {
				if (ids) ids[k] = s;
				if (done) done[k] = b->hStatus[s].done;
				if (iters) iters[k] = b->hStatus[s].iter;
				if (rmsLast) rmsLast[k] = b->hStatus[s].rmsLast;
			}
			k++;
		}
		*n = k;
		return E3D_OK;
	});
}

int e3d_batch_result(e3d_batch *b, int id, double *hist, int cap, int *nHist, double coeffs[3], double cfOut[3],
                     double cmOut[3]) {
	if (!b || id < 0 || id >= b->nSlots) return fail(E3D_ERR_ARG, "no such case");
	return guarded([&] {
		std::lock_guard<std::mutex> pk(b->pollMu);
		std::lock_guard<std::mutex> lk(b->mu);
		if (!b->cs[id].ready) return fail(E3D_ERR_ARG, "no such case");
		hipCheck(hipSetDevice(b->device), "hipSetDevice");
		activate(b);
		CaseHost &c = b->cs[id];
		e3d::MixStatus *h = b->hStatus + id;
		hipCheck(hipMemcpyAsync(h, b->dStatus + id, sizeof(*h), hipMemcpyDeviceToHost, b->stream), "D2H status");
		b->sync();
		const int nh = std::min(h->iter, c.desc.histCap);
		if (nHist) *nHist = nh;
		if (hist && cap > 0 && nh > 0)
			hipCheck(hipMemcpy(hist, c.desc.hist, sizeof(double) * std::min(nh, cap), hipMemcpyDeviceToHost), "D2H hist");
		double v[6] = {0, 0, 0, 0, 0, 0};
		if (c.nWall) {
			e3d::launchMixAero(c.nWall, c.wallCell, c.wN[0], c.wN[1], c.wN[2], c.wC[0], c.wC[1], c.wC[2], b->dCases + id,
			                   c.cpFactor, c.fr.p_inf, c.prm.cref, c.prm.sref, c.scalars, b->stream);
			hipCheck(hipGetLastError(), "kernel launch");
			hipCheck(hipMemcpyAsync(v, c.scalars, sizeof(v), hipMemcpyDeviceToHost, b->stream), "D2H aero");
			b->sync();
		}
		// AeroCoefficients.cpp:49-51
		const e3d_params &q = c.prm;
		double lever[3] = {q.ref_point[0] - v[3], q.ref_point[1] - v[4], q.ref_point[2] - v[5]};
		double M[3] = {v[1] * lever[2] - v[2] * lever[1], v[2] * lever[0] - v[0] * lever[2], v[0] * lever[1] - v[1] * lever[0]};
		double cf[3], cm[3];
		for (int k = 0; k < 3; k++) cf[k] = v[k], cm[k] = M[k] / q.cref;
		if (coeffs) {
			coeffs[0] = cf[q.orient_cl[0]] * q.orient_cl[1];
			coeffs[1] = cf[q.orient_cd[0]] * q.orient_cd[1];
			coeffs[2] = cm[q.orient_cm[0]] * q.orient_cm[1];
		}
		for (int k = 0; k < 3; k++) {
			if (cfOut) cfOut[k] = cf[k];
			if (cmOut) cmOut[k] = cm[k];
		}
		if (h->done == E3D_CASE_FAILED)
			return fail(E3D_ERR_NUMERIC, "case " + std::to_string(id) + ": non-finite residual at iteration " +
			                                 std::to_string(h->iter));
		return E3D_OK;
	});
}

int e3d_batch_get_state(e3d_batch *b, int id, double *const fields[8]) {
	if (!b || !fields || id < 0 || id >= b->nSlots) return fail(E3D_ERR_ARG, "no such case");
	return guarded([&] {
		std::lock_guard<std::mutex> pk(b->pollMu);
		std::lock_guard<std::mutex> lk(b->mu);
		if (!b->cs[id].ready) return fail(E3D_ERR_ARG, "no such case");
		hipCheck(hipSetDevice(b->device), "hipSetDevice");
		activate(b);
		CaseHost &c = b->cs[id];
		e3d::MixStatus *h = b->hStatus + id;
		hipCheck(hipMemcpyAsync(h, b->dStatus + id, sizeof(*h), hipMemcpyDeviceToHost, b->stream), "D2H status");
		b->sync();
		const size_t nE = (size_t) c.nElem, nG = (size_t) c.nGhost;
		double *tmp = nullptr;
		hipCheck(hipMallocAsync((void **) &tmp, 8 * nE * sizeof(double), b->stream), "hipMallocAsync");
		if (h->iter > 0) e3d::launchMixMirrorGhosts(c.desc, h->parity, b->stream);
		e3d::launchMixPrimitives(c.desc, tmp, b->stream);
		hipCheck(hipGetLastError(), "kernel launch");
		std::vector<double> prim(8 * nE);
		std::vector<float> rec(nG * e3d::kRecWidth);
		hipCheck(hipMemcpyAsync(prim.data(), tmp, prim.size() * 8, hipMemcpyDeviceToHost, b->stream), "D2H");
		if (nG)
			hipCheck(hipMemcpyAsync(rec.data(), c.desc.rec[h->parity] + nE * e3d::kRecWidth, rec.size() * 4,
			                        hipMemcpyDeviceToHost, b->stream),
			         "D2H");
		hipCheck(hipFreeAsync(tmp, b->stream), "hipFreeAsync");
		b->sync();
		for (int k = 0; k < 8; k++)
			if (fields[k])
				for (size_t i = 0; i < nE; i++) fields[k][c.new2old[i]] = prim[k * nE + i];
		const e3d::DevFree &q = c.fr;
		for (size_t g = 0; g < nG; g++) {
			const float *r = rec.data() + g * e3d::kRecWidth;
			const double rho = q.rho_inf + r[e3d::R_DRHO], u = q.u_inf + r[e3d::R_DU], v = q.v_inf + r[e3d::R_DV];
			const double w = q.w_inf + r[e3d::R_DW], p = q.p_inf + r[e3d::R_DP], H = q.H_inf + r[e3d::R_DH];
			const double vals[8] = {rho, u, v, w, p, H - p / rho, H, (double) r[e3d::R_M]};
			for (int k = 0; k < 8; k++)
				if (fields[k]) fields[k][nE + g] = vals[k];
		}
		return E3D_OK;
	});
}

int e3d_batch_remove(e3d_batch *b, int id) {
	if (!b || id < 0 || id >= b->nSlots) return fail(E3D_ERR_ARG, "no such case");
	return guarded([&] {
		std::lock_guard<std::mutex> lk(b->mu);
		if (!b->cs[id].ready) return fail(E3D_ERR_ARG, "no such case");
		hipCheck(hipSetDevice(b->device), "hipSetDevice");
		CaseHost &c = b->cs[id];
		// kernels already queued may still read the case; the free is ordered after them
		hipCheck(hipFreeAsync(c.mem, b->stream), "hipFreeAsync");
		b->used -= (double) c.bytes;
		c = CaseHost{};
		b->statusInit.push_back(id);
		b->dirty = true;
		return E3D_OK;
	});
}

int e3d_batch_set_tuning(e3d_batch *b, int wgPerCU, int xcds, int nontemporal) {
	if (!b) return fail(E3D_ERR_ARG, "null batch");
	if (wgPerCU < 0 || wgPerCU > 32 || xcds <= 0 || xcds > 64) return fail(E3D_ERR_ARG, "invalid launch shape");
	b->tune.wgPerCU = wgPerCU;
	b->tune.xcds = xcds;
	b->tune.nontemporal = nontemporal != 0;
	b->dirty = true; // recompute the grid; the graph is captured again when the launch changes
	freeGraph(b);
	return E3D_OK;
}

int e3d_batch_info(e3d_batch *b, int *nCases, double *bytes, double *budget) {
	if (!b) return fail(E3D_ERR_ARG, "null batch");
	int n = 0;
	for (auto &c : b->cs) n += c.used ? 1 : 0;
	if (nCases) *nCases = n;
	if (bytes) *bytes = b->used;
	if (budget) *budget = b->budget;
	return E3D_OK;
}

int e3d_batch_set_timing(e3d_batch *b, int enable) {
	if (!b) return fail(E3D_ERR_ARG, "null batch");
	b->timing = enable != 0;
	b->evUsed = 0;
	return E3D_OK;
}

int e3d_batch_timing(e3d_batch *b, double *avgMs, int *launches, int *perIteration) {
	if (!b) return fail(E3D_ERR_ARG, "null batch");
	return guarded([&] {
		std::lock_guard<std::mutex> lk(b->mu);
		b->sync();
		double total = 0;
		int n = (int) (b->evUsed / 2);
		for (int i = 0; i < n; i++) {
			float ms = 0;
			hipCheck(hipEventElapsedTime(&ms, b->evPool[2 * i], b->evPool[2 * i + 1]), "hipEventElapsedTime");
			total += ms;
		}
		if (avgMs) *avgMs = n ? total / n : 0.0;
		if (launches) *launches = n;
		if (perIteration) *perIteration = b->useGraph ? 1 : 0;
		b->evUsed = 0;
		return E3D_OK;
	});
}

void e3d_batch_destroy(e3d_batch *b) {
	if (!b) return;
	(void) hipSetDevice(b->device);
	(void) hipStreamSynchronize(b->stream);
	freeGraph(b);
	for (auto &c : b->cs)
		if (c.used && c.mem) (void) hipFreeAsync(c.mem, b->stream);
	if (b->dBlocks) (void) hipFreeAsync(b->dBlocks, b->stream);
	(void) hipStreamSynchronize(b->stream);
	(void) hipFree(b->dCases);
	(void) hipFree(b->dStatus);
	(void) hipFree(b->dNBlocks);
	(void) hipHostFree(b->hCases);
	(void) hipHostFree(b->hStatusInit);
	(void) hipHostFree(b->hStatus);
	(void) hipHostFree(b->hN);
	if (b->hBlocks) (void) hipHostFree(b->hBlocks);
	for (auto e : b->evPool) (void) hipEventDestroy(e);
	if (b->staged) (void) hipEventDestroy(b->staged);
	if (b->polled) (void) hipEventDestroy(b->polled);
	(void) hipStreamDestroy(b->stream);
	delete b;
}

}// extern "C"
