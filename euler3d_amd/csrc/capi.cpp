// C ABI of libe3d_hip.so: mesh and configuration handling on the host, device context that keeps
// one case resident in HBM and advances it with the kernels of kernels.hip.
#include "../../include/e3d.h"
#include "abi_util.hpp"
#include "device.hpp"
#include "e3d_host.hpp"
#include "plan.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

std::string &e3d_abi::lastError() {
	thread_local std::string msg;
	return msg;
}

namespace {

using e3d_abi::fail;
using e3d_abi::guarded;
using e3d_abi::hipCheck;
using e3d_abi::HipError;
using e3d_abi::fillFree;

template<typename T>
struct DBuf {
	T *p = nullptr;
	size_t n = 0;
	void alloc(size_t count) {
		free();
		n = count;
		if (count) hipCheck(hipMalloc((void **) &p, count * sizeof(T)), "hipMalloc");
	}
	void upload(const T *h, size_t count, hipStream_t s) {
		if (count) hipCheck(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
	}
	void free() {
		if (p) (void) hipFree(p);
		p = nullptr;
		n = 0;
	}
	~DBuf() { free(); }
};

}// namespace

struct e3d_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	e3d_params prm{};
	e3d::DevFree fr{};
	e3d::DevMesh dm{};
	int nElem = 0, nFace = 0, nFaceInt = 0, nGhost = 0, nTot = 0;
	int machDerived = 0;
	double lastMs = 0.0;
	// per-launch timing of the stage kernel(s)
	bool timing = false;
	std::vector<hipEvent_t> evPool;
	size_t evUsed = 0;
	bool fused = true;       // fused stage kernel (off: the three-pass kBc/kFlux/kStage form)
	e3d::StageTuning tune{};
	std::vector<int> new2old; // device cell id -> reference cell id (interior)
	// mesh, device numbering
	DBuf<int> e0, e1, fstart, flist;
	DBuf<double> nx, ny, nz, S, vol, sx, sy, sz;
	DBuf<int8_t> gkind;
	// fused-stage plan
	e3d::DevPlan plan{};
	DBuf<int> blockSlot, se0, se1;
	DBuf<int8_t> skind;
	DBuf<double> snx, sny, snz, sS;
	DBuf<uint16_t> csr;
	DBuf<uint32_t> spack;
	DBuf<int> blockHalo, haloCell, blockMeta;
	// state: two buffers of (nElem + nGhost) x 8 doubles, st[cur] is the current one
	DBuf<double> st[2];
	int cur = 0;
	// interior state as conservative W (5 doubles per cell) in wst[cur] instead of primitive
	// records in st[cur]: the fused RK stage keeps it so, readers of primitives convert first
	DBuf<double> wst[2];
	bool wMode = false;
	DBuf<double> W0, fs, partial, sumTmp, scalars; // sumTmp: chunk sums of the RMS partials
	int partialCount = 0; // block partials the last RMS-forming stage wrote
	// residual smoothing (RESIDUAL_SMOOTHING= 1): neighbour lists in device numbering, residual
	// before/during/after the sweeps, per-block largest change, decision and NaN flags
	int nbrStride = 0;
	DBuf<int> nbr;
	DBuf<uint8_t> nface;
	DBuf<double> R0, R1, R2, blockMax, cellMax;
	DBuf<int> smoothFlag, partStart;
	// aero (wall list order)
	int nWall = 0;
	DBuf<int> wallCell;
	DBuf<double> wN[3], wC[3];
	double cpFactor = 0.0;
	// partitions of a composite mesh (one entry otherwise): wall ranges for the per-partition
	// coefficients, interface ghosts (reference ids) and the cells they copy
	int nPart = 1;
	std::vector<int> wall0, nWallPart;
	std::vector<int> mpiGhost, mpiCell;
	DBuf<double> aeroOut, aeroTmp; // aeroTmp: chunk sums of the force integration
	DBuf<double> wallCp;  // cp of every wall face at the last force update (AeroCoefficients::_cp)
	bool wallCpSet = false;
	// partitioned input with the interface exchange (E3D_EXCHANGE=halo): interface faces read their
	// ghosts, which launchHaloExchange fills from the send cells before every stage and in
	// e3d_update_bc; exchanged = the ghosts hold exchanged values (after e3d_update_bc) rather
	// than the copy of their own cell that an update leaves (FlowField.cpp:129-138)
	bool halo = false, exchanged = false;
	int nX = 0;
	DBuf<int> xSrc, xDst, xSlot;
	DBuf<double> xBuf;

	e3d::DevState state(int which) { return {st[which].p}; }
	e3d::DevState state() { return state(cur); }
	// primitive records of the interior cells in st[cur] (from wst[cur] when the RK stages left
	// the state there)
	void ensurePrims() {
		if (!wMode) return;
		e3d::launchWToPrim(dm, wst[cur].p, state(), fr, stream);
		check();
		wMode = false;
	}
	void sync() { hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
	void check() { hipCheck(hipGetLastError(), "kernel launch"); }
	// MPIHandler::updateFlowField on one device: send cells -> message buffer -> ghost records of
	// the state the next stage reads
	void exchange() {
		if (!halo) return;
		e3d::launchHaloExchange(nX, xSrc.p, xDst.p, xSlot.p, state(), wMode ? wst[cur].p : nullptr, fr.gamma, xBuf.p,
		                        stream);
		check();
	}
};

namespace {

void initFreestream(e3d_ctx *c) {
	std::vector<double> h(c->nTot);
	const double vals[8] = {c->fr.rho_inf, c->fr.u_inf, c->fr.v_inf, c->fr.w_inf, c->fr.p_inf, c->fr.E_inf, c->fr.H_inf, c->fr.M_inf};
	c->cur = 0;
	h.resize((size_t) c->nTot * e3d::kStateWidth);
	for (size_t i = 0; i < (size_t) c->nTot; i++)
		for (int k = 0; k < 8; k++) h[i * e3d::kStateWidth + k] = vals[k];
	c->st[0].upload(h.data(), h.size(), c->stream);
	// FlowField::Initialize (FlowField.cpp:83-95): rho u, rho v, rho w, rho E at rho_inf * (.)_inf,
	// the explicit Euler step's conservative state
	std::vector<double> w(5 * (size_t) c->nElem);
	const double wi[5] = {c->fr.rho_inf, c->fr.rho_inf * c->fr.u_inf, c->fr.rho_inf * c->fr.v_inf,
	                      c->fr.rho_inf * c->fr.w_inf, c->fr.rho_inf * c->fr.E_inf};
	for (int k = 0; k < 5; k++) std::fill(w.begin() + k * (size_t) c->nElem, w.begin() + (k + 1) * (size_t) c->nElem, wi[k]);
	c->W0.upload(w.data(), w.size(), c->stream);
	c->sync();
	c->machDerived = 0;
	c->wMode = false;
}

void runStage(e3d_ctx *c, int stage, double *rmsPartial, bool timed) {
	static const double alphas[5] = {0.0533, 0.1263, 0.2375, 0.4414, 1.0};
	auto s = c->state();
	hipEvent_t evA = nullptr, evB = nullptr;
	auto record = [&](hipEvent_t e) {
		if (e) hipCheck(hipEventRecord(e, c->stream), "hipEventRecord");
	};
	if (c->timing && timed) {
		if (c->evUsed + 2 > c->evPool.size()) {
			for (int i = 0; i < 64; i++) {
				hipEvent_t e;
				hipCheck(hipEventCreate(&e), "hipEventCreate");
				c->evPool.push_back(e);
			}
		}
		evA = c->evPool[c->evUsed++];
		evB = c->evPool[c->evUsed++];
	}
	const int ti = c->prm.time_integration;
	const bool smooth = c->prm.residual_smoothing != 0;
	c->exchange(); // updateBC's interface exchange (EulerSolver.cpp:248-250), before the stage reads
	if (c->fused && !smooth && ti != 1) {
		// one fused launch: ghosts, fluxes, gather, dt, update (RK stage or explicit Euler step);
		// RK stages leave the interior state as W, the explicit Euler step as primitives
		const int integ = ti == 0 ? 1 : 0;
		if (integ == 1) c->ensurePrims();
		const int layout = integ == 1 ? 0 : c->wMode ? 2 : 1;
		auto o = c->state(c->cur ^ 1);
		record(evA);
		e3d::launchStageFused(c->dm, c->plan, s, o, c->wst[c->cur].p, c->wst[c->cur ^ 1].p, layout, c->fr,
		                      c->prm.scheme, integ, c->W0.p, alphas[stage], stage == 0, c->machDerived, rmsPartial,
		                      c->cur, c->tune, c->stream);
		record(evB);
		c->cur ^= 1;
		c->wMode = layout != 0;
		if (rmsPartial) c->partialCount = c->plan.nBlocks;
	} else {
		// EulerSolver::Run / RungeKutta body: updateBC, computeResidual, smoothResiduals, update in place
		c->ensurePrims();
		record(evA);
		e3d::launchBc(c->dm, s, c->fr, c->machDerived, c->stream);
		e3d::launchFlux(c->dm, s, c->fr, c->prm.scheme, c->fs.p, nullptr, c->stream);
		const double *Rin = nullptr;
		if (smooth) {
			e3d::launchGather(c->dm, c->fs.p, c->R0.p, c->stream);
			e3d::launchSmooth(c->R0.p, c->R1.p, c->R2.p, c->nbr.p, c->nface.p, c->nbrStride, c->nElem, c->blockMax.p,
			                  c->smoothFlag.p, c->partStart.p, c->nPart, c->cellMax.p, c->stream);
			Rin = c->R2.p;
		}
		const int integ = ti == 2 ? 0 : ti == 0 ? 1 : 2;
		e3d::launchStage(c->dm, s, c->fr, c->fs.p, Rin, c->W0.p, alphas[stage], stage == 0, integ, rmsPartial, nullptr,
		                 c->stream);
		record(evB);
		if (rmsPartial) c->partialCount = e3d::stageBlocks(c->nElem);
	}
	// the stored interior Mach number becomes c/|u| with the first update; TIME_INTEGRATION= 1
	// never updates the state
	if (ti != 1) c->machDerived = 1;
}

const char *stageKernelName(const e3d_ctx *c) {
	if (c->fused && !c->prm.residual_smoothing && c->prm.time_integration != 1)
		return c->tune.pipeGrid > 0 && c->tune.nontemporal && c->tune.sharedRecip ? "kStagePipe" : "kStageFused";
	return c->prm.residual_smoothing ? "kBc+kFlux+kGather+kSmooth+kStage" : "kBc+kFlux+kStage";
}

// NaN in the residual smoothing (the reference exits the process, ResidualVar.hpp:96-104)
void checkSmoothing(e3d_ctx *c) {
	if (!c->prm.residual_smoothing) return;
	int nan = 0;
	hipCheck(hipMemcpyAsync(&nan, c->smoothFlag.p + c->nPart, sizeof(int), hipMemcpyDeviceToHost, c->stream), "D2H flag");
	c->sync();
	if (nan) throw std::domain_error("NaN residual in the residual smoothing");
}

// host reference-order field k <-> device state (interior permuted, ghosts as they are, 8 doubles
// per cell)
void toDevice(const e3d_ctx *c, const double *ref, int k, std::vector<double> &dev) {
	for (int i = 0; i < c->nElem; i++) dev[(size_t) i * e3d::kStateWidth + k] = ref[c->new2old[i]];
	for (int i = c->nElem; i < c->nTot; i++) dev[(size_t) i * e3d::kStateWidth + k] = ref[i];
}
void toReference(const e3d_ctx *c, const std::vector<double> &dev, int k, double *ref) {
	for (int i = 0; i < c->nElem; i++) ref[c->new2old[i]] = dev[(size_t) i * e3d::kStateWidth + k];
	for (int i = c->nElem; i < c->nTot; i++) ref[i] = dev[(size_t) i * e3d::kStateWidth + k];
}

// one iteration of Run's body; the sum of R_rho^2 of the last stage lands in *dsum (device).
// timed = false leaves out the per-launch timing events (they are created on demand).
void runIteration(e3d_ctx *c, double *dsum, bool timed = true) {
	if (c->prm.time_integration == 2) {
		for (int stage = 0; stage < 5; stage++) runStage(c, stage, stage == 4 ? c->partial.p : nullptr, timed);
	} else {
		runStage(c, 0, c->partial.p, timed); // explicit Euler step, or (implicit) residual only
	}
	e3d::launchSum(c->partial.p, c->partialCount, c->sumTmp.p, dsum, c->stream);
}

}// namespace

extern "C" {

const char *e3d_last_error(void) { return e3d_abi::lastError().c_str(); }
const char *e3d_version(void) { return "e3d-mi355x 0.1 (gfx950, FP64 plain path)"; }

int e3d_device_count(int *n) {
	if (!n) return fail(E3D_ERR_ARG, "null argument");
	int k = 0;
	// no device (or no driver) is a count of zero, not an error
	if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
	*n = k;
	return E3D_OK;
}

int e3d_params_default(e3d_params *out) {
	if (!out) return fail(E3D_ERR_ARG, "null params");
	e3d::Params p;
	std::memset(out, 0, sizeof(*out));
	out->mach = 0.5;
	out->aoa_deg = 0.0;
	out->gamma = p.gamma;
	out->cfl = 1.5;
	out->sref = 1.0;
	out->cref = 1.0;
	out->orient_cl[0] = 1, out->orient_cl[1] = 1;
	out->orient_cd[0] = 0, out->orient_cd[1] = 1;
	out->orient_cm[0] = 2, out->orient_cm[1] = 1;
	out->scheme = 0;
	out->time_integration = 2;
	out->residual_smoothing = 0;
	out->max_iter = 1;
	out->sampling = 1;
	out->min_residual = 0.0;
	out->min_aerocoeff_error = 0.0;
	return E3D_OK;
}

int e3d_config_read(const char *path, e3d_params *out, char *mesh_path, int mesh_path_len, char *out_dir,
                    int out_dir_len) {
	if (!path || !out) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		e3d::Params p = e3d::readConfig(path);
		std::memset(out, 0, sizeof(*out));
		out->mach = p.mach;
		out->aoa_deg = p.aoa;
		out->gamma = p.gamma;
		out->cfl = p.cfl;
		out->sref = p.sref;
		out->cref = p.cref;
		for (int k = 0; k < 3; k++) out->ref_point[k] = p.refPoint[k];
		for (int k = 0; k < 2; k++) {
			out->orient_cl[k] = p.orientCL[k];
			out->orient_cd[k] = p.orientCD[k];
			out->orient_cm[k] = p.orientCM[k];
		}
		out->scheme = p.scheme;
		out->time_integration = p.timeIntegration;
		out->residual_smoothing = p.residualSmoothing;
		out->max_iter = p.maxIter;
		out->sampling = p.sampling;
		out->min_residual = p.minResidual;
		out->min_aerocoeff_error = p.minAeroCoeffError;
		auto copy = [](const std::string &s, char *dst, int len) {
			if (!dst || len <= 0) return;
			std::strncpy(dst, s.c_str(), (size_t) len - 1);
			dst[len - 1] = 0;
		};
		copy(p.partitionFiles.empty() ? std::string() : p.partitionFiles[0], mesh_path, mesh_path_len);
		copy(p.outputDir, out_dir, out_dir_len);
		return E3D_OK;
	});
}

int e3d_config_partition(const char *path, int index, char *buf, int len, int *n_part) {
	if (!path) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		e3d::Params p = e3d::readConfig(path);
		if (n_part) *n_part = (int) p.partitionFiles.size();
		if (buf && len > 0) {
			if (index < 0 || index >= (int) p.partitionFiles.size()) return fail(E3D_ERR_ARG, "partition index out of range");
			std::strncpy(buf, p.partitionFiles[index].c_str(), (size_t) len - 1);
			buf[len - 1] = 0;
		}
		return E3D_OK;
	});
}

int e3d_config_initial_mesh(const char *path, char *buf, int len) {
	if (!path || !buf || len <= 0) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		e3d::Params p = e3d::readConfig(path);
		std::strncpy(buf, p.initialMesh.c_str(), (size_t) len - 1);
		buf[len - 1] = 0;
		return E3D_OK;
	});
}

int e3d_mesh_read(const char *path, int require_par, e3d_mesh **out) {
	if (!path || !out) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		auto m = std::make_unique<e3d_mesh>();
		try {
			m->file = e3d::readSu2(path, require_par ? e3d::ParBlock::Required : e3d::ParBlock::Optional);
		} catch (const std::exception &e) {
			return fail(E3D_ERR_IO, e.what());
		}
		*out = m.release();
		return E3D_OK;
	});
}

int e3d_mesh_read_partitions(int n_part, const char *const *paths, e3d_mesh **out) {
	if (n_part < 1 || !paths || !out) return fail(E3D_ERR_ARG, "null argument or no partition");
	for (int r = 0; r < n_part; r++)
		if (!paths[r]) return fail(E3D_ERR_ARG, "null partition path");
	return guarded([&] {
		auto m = std::make_unique<e3d_mesh>();
		for (int r = 0; r < n_part; r++) {
			try {
				m->parts.push_back(e3d::readSu2(paths[r], e3d::ParBlock::Partition));
			} catch (const std::exception &e) {
				return fail(E3D_ERR_IO, e.what());
			}
		}
		// until preprocessing, the merged copy holds the raw node and element counts
		for (auto &pf : m->parts) {
			const int node0 = m->file.nNode();
			m->file.xyz.insert(m->file.xyz.end(), pf.xyz.begin(), pf.xyz.end());
			if (m->file.elemStart.empty()) m->file.elemStart.push_back(0);
			for (int e = 0; e < pf.nElem(); e++) {
				m->file.elemVtk.push_back(pf.elemVtk[e]);
				for (int j = pf.elemStart[e]; j < pf.elemStart[e + 1]; j++) m->file.elemNodes.push_back(node0 + pf.elemNodes[j]);
				m->file.elemStart.push_back((int) m->file.elemNodes.size());
			}
		}
		*out = m.release();
		return E3D_OK;
	});
}

int e3d_mesh_partition_count(const e3d_mesh *m, int *n) {
	if (!m || !n) return fail(E3D_ERR_ARG, "null argument");
	*n = m->parts.empty() ? 1 : (int) m->parts.size();
	return E3D_OK;
}

int e3d_mesh_partition_info(const e3d_mesh *m, int r, e3d_partition_info *out) {
	if (!m || !out) return fail(E3D_ERR_ARG, "null argument");
	if (!m->preprocessed) return fail(E3D_ERR_ARG, "mesh is not preprocessed (call e3d_mesh_preprocess)");
	const int n = m->parts.empty() ? 1 : (int) m->parts.size();
	if (r < 0 || r >= n) return fail(E3D_ERR_ARG, "partition index out of range");
	e3d::PartRange R;
	if (m->parts.empty()) {
		R.nElem = m->sm.nElem, R.nNode = m->file.nNode(), R.nFaceInt = m->sm.nFaceInt, R.nGhost = m->sm.nGhost;
		R.nWall = (int) m->sm.wallGhost.size();
	} else {
		R = m->ranges[r];
	}
	*out = {R.elem0, R.nElem, R.node0, R.nNode, R.faceInt0, R.nFaceInt, R.ghost0, R.nGhost, R.wall0, R.nWall};
	return E3D_OK;
}

int e3d_mesh_partition_assign(const e3d_mesh *m, int n_part, int *elem2part) {
	if (!m || !elem2part) return fail(E3D_ERR_ARG, "null argument");
	if (!m->preprocessed) return fail(E3D_ERR_ARG, "mesh is not preprocessed (call e3d_mesh_preprocess)");
	if (!m->parts.empty()) return fail(E3D_ERR_ARG, "the mesh is already partitioned");
	return guarded([&] {
		std::vector<int> p = e3d::partitionRcb(m->sm, n_part);
		std::copy(p.begin(), p.end(), elem2part);
		return E3D_OK;
	});
}

int e3d_mesh_partition_write(const e3d_mesh *m, int n_part, const int *elem2part, const char *const *paths,
                             int precision) {
	if (!m || !paths || n_part < 1) return fail(E3D_ERR_ARG, "null argument");
	for (int r = 0; r < n_part; r++)
		if (!paths[r]) return fail(E3D_ERR_ARG, "null partition path");
	if (!m->preprocessed) return fail(E3D_ERR_ARG, "mesh is not preprocessed (call e3d_mesh_preprocess)");
	if (!m->parts.empty()) return fail(E3D_ERR_ARG, "the mesh is already partitioned");
	return guarded([&] {
		std::vector<int> assign;
		if (elem2part) assign.assign(elem2part, elem2part + m->sm.nElem);
		else assign = e3d::partitionRcb(m->sm, n_part);
		std::vector<e3d::MeshFile> files = e3d::splitMesh(m->file, m->sm, assign, n_part);
		for (int r = 0; r < n_part; r++) {
			try {
				e3d::writeSu2(files[r], paths[r], precision, true);
			} catch (const std::exception &e) {
				return fail(E3D_ERR_IO, e.what());
			}
		}
		return E3D_OK;
	});
}

int e3d_mesh_write(const e3d_mesh *m, const char *path, int precision, int par_block) {
	if (!m || !path) return fail(E3D_ERR_ARG, "null argument");
	if (!m->parts.empty()) return fail(E3D_ERR_ARG, "a mesh of several partitions is written by the partitioner");
	return guarded([&] {
		e3d::writeSu2(m->file, path, precision, par_block != 0);
		return E3D_OK;
	});
}

int e3d_post_tecplot_files(const char *out_path, int n_part, const char *const *par_paths,
                           const char *const *sol_paths) {
	if (!out_path || n_part < 0 || (n_part > 0 && (!par_paths || !sol_paths))) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		std::vector<std::string> par, sol;
		for (int i = 0; i < n_part; i++) {
			if (!par_paths[i] || !sol_paths[i]) return fail(E3D_ERR_ARG, "null path");
			par.emplace_back(par_paths[i]);
			sol.emplace_back(sol_paths[i]);
		}
		try {
			e3d::writeTecplot(out_path, par, sol);
		} catch (const std::invalid_argument &) {
			throw;
		} catch (const std::exception &e) {
			return fail(E3D_ERR_IO, e.what());
		}
		return E3D_OK;
	});
}

int e3d_post_tecplot(const char *config_path) {
	if (!config_path) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		e3d::Params p = e3d::readConfig(config_path);
		std::vector<std::string> sol = e3d::postSolutionPaths(p);
		std::vector<const char *> pp, sp;
		for (size_t i = 0; i < sol.size(); i++) {
			pp.push_back(p.partitionFiles[i].c_str());
			sp.push_back(sol[i].c_str());
		}
		return e3d_post_tecplot_files(p.outputFile.c_str(), (int) pp.size(), pp.data(), sp.data());
	});
}

int e3d_mesh_gen_box(int nx, int ny, int nz, double lx, double ly, double lz, const char *tags6, e3d_mesh **out) {
	if (!tags6 || !out) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		std::vector<std::string> t;
		std::string s(tags6), cur;
		for (char ch : s) {
			if (ch == ',') {
				t.push_back(cur);
				cur.clear();
			} else cur += ch;
		}
		t.push_back(cur);
		if (t.size() != 6) return fail(E3D_ERR_ARG, "tags6 needs six comma-separated tags");
		const char *tp[6];
		for (int i = 0; i < 6; i++) tp[i] = t[i].c_str();
		auto m = std::make_unique<e3d_mesh>();
		m->file = e3d::genBox(nx, ny, nz, lx, ly, lz, tp);
		*out = m.release();
		return E3D_OK;
	});
}

int e3d_mesh_gen_naca(int ni, int nj, int nk, double thickness, double camber, double camber_pos, double wall_spacing,
                      double farfield_radius, double dz, e3d_mesh **out) {
	if (!out) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		auto m = std::make_unique<e3d_mesh>();
		m->file = e3d::genNacaOgrid(ni, nj, nk, thickness, camber, camber_pos, wall_spacing, farfield_radius, dz, 0);
		*out = m.release();
		return E3D_OK;
	});
}

int e3d_mesh_gen_bump(int nx, int ny, int nz, double jitter, unsigned long long seed, int permute, e3d_mesh **out) {
	if (!out) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		auto m = std::make_unique<e3d_mesh>();
		m->file = e3d::genBumpTets(nx, ny, nz, jitter, seed, permute != 0);
		*out = m.release();
		return E3D_OK;
	});
}

int e3d_mesh_gen_wing(int ni, int nj, int nk, double thickness, double sweep_deg, double taper, double semispan,
                      double farfield, e3d_mesh **out) {
	if (!out) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		auto m = std::make_unique<e3d_mesh>();
		m->file = e3d::genWing(ni, nj, nk, thickness, sweep_deg, taper, semispan, farfield);
		*out = m.release();
		return E3D_OK;
	});
}

int e3d_mesh_preprocess(e3d_mesh *m, double cref) {
	if (!m) return fail(E3D_ERR_ARG, "null mesh");
	if (!(cref > 0)) return fail(E3D_ERR_ARG, "CREF must be positive");
	return guarded([&] {
		if (m->parts.empty()) {
			m->sm = e3d::preprocess(m->file, cref);
		} else {
			std::vector<e3d::SolverMesh> sms;
			for (auto &pf : m->parts) sms.push_back(e3d::preprocess(pf, cref));
			m->sm = e3d::mergePartitions(m->parts, sms, m->ranges, m->file);
		}
		m->preprocessed = true;
		return E3D_OK;
	});
}

int e3d_mesh_get_info(const e3d_mesh *m, e3d_mesh_info *out) {
	if (!m || !out) return fail(E3D_ERR_ARG, "null argument");
	out->n_node = m->file.nNode();
	out->n_elem = m->file.nElem();
	out->n_face = m->sm.nFace;
	out->n_face_int = m->sm.nFaceInt;
	out->n_ghost = m->sm.nGhost;
	out->n_wall = (int) m->sm.wallGhost.size();
	out->n_far = (int) m->sm.farGhost.size();
	out->n_sym = (int) m->sm.symGhost.size();
	return E3D_OK;
}

int e3d_mesh_get_i32(const e3d_mesh *m, const char *name, int *out, long long capacity, long long *n) {
	if (!m || !name) return fail(E3D_ERR_ARG, "null argument");
	if (out && capacity < 0) return fail(E3D_ERR_ARG, "negative capacity");
	const std::vector<int> *v = nullptr;
	std::string s(name);
	const auto &sm = m->sm;
	std::vector<int> order;
	if (s == "deviceOrder") {
		// bisection blocks of the default size, as e3d_create plans them
		order = e3d::cellOrder(sm, e3d::ORDER_BISECT, e3d::kCellsPerBlock);
		v = &order;
	} else if (s == "face2element") v = &sm.face2element;
	else if (s == "face2nodeStart") v = &sm.face2nodeStart;
	else if (s == "face2node") v = &sm.face2node;
	else if (s == "element2faceStart") v = &sm.element2faceStart;
	else if (s == "element2face") v = &sm.element2face;
	else if (s == "element2element") v = &sm.element2element;
	else if (s == "wallGhost") v = &sm.wallGhost;
	else if (s == "wallCell") v = &sm.wallCell;
	else if (s == "wallFace") v = &sm.wallFace;
	else if (s == "farGhost") v = &sm.farGhost;
	else if (s == "farCell") v = &sm.farCell;
	else if (s == "farFace") v = &sm.farFace;
	else if (s == "symGhost") v = &sm.symGhost;
	else if (s == "symCell") v = &sm.symCell;
	else if (s == "symFace") v = &sm.symFace;
	else if (s == "ghostKind") v = &sm.ghostKind;
	else if (s == "mpiGhost") v = &sm.mpiGhost;
	else if (s == "mpiCell") v = &sm.mpiCell;
	else if (s == "partStart") v = &sm.partStart;
	else if (s == "haloStart") v = &sm.haloStart;
	else if (s == "haloFrom") v = &sm.haloFrom;
	else if (s == "haloTo") v = &sm.haloTo;
	else if (s == "haloSrc") v = &sm.haloSrc;
	else if (s == "haloDst") v = &sm.haloDst;
	else if (s == "elemVtk") v = &m->file.elemVtk;
	else if (s == "elemStart") v = &m->file.elemStart;
	else if (s == "elemNodes") v = &m->file.elemNodes;
	else return fail(E3D_ERR_ARG, "unknown int32 array " + s);
	if (n) *n = (long long) v->size();
	if (out && !v->empty())
		std::memcpy(out, v->data(), sizeof(int) * (size_t) std::min<long long>(capacity, (long long) v->size()));
	return E3D_OK;
}

int e3d_mesh_get_f64(const e3d_mesh *m, const char *name, double *out, long long capacity, long long *n) {
	if (!m || !name) return fail(E3D_ERR_ARG, "null argument");
	if (out && capacity < 0) return fail(E3D_ERR_ARG, "negative capacity");
	const std::vector<double> *v = nullptr;
	std::string s(name);
	const auto &sm = m->sm;
	if (s == "faceNormals") v = &sm.faceNormal;
	else if (s == "faceNormalsUnit") v = &sm.faceUnitNormal;
	else if (s == "faceSurfaces") v = &sm.faceArea;
	else if (s == "faceCenters") v = &sm.faceCenter;
	else if (s == "cellVolumes") v = &sm.cellVolume;
	else if (s == "cellCentroids") v = &sm.cellCentroid;
	else if (s == "xyz") v = &m->file.xyz;
	else return fail(E3D_ERR_ARG, "unknown f64 array " + s);
	if (n) *n = (long long) v->size();
	if (out && !v->empty())
		std::memcpy(out, v->data(), sizeof(double) * (size_t) std::min<long long>(capacity, (long long) v->size()));
	return E3D_OK;
}

void e3d_mesh_free(e3d_mesh *m) { delete m; }

int e3d_create(const e3d_mesh *m, const e3d_params *p, int device, e3d_ctx **out) {
	if (!m || !p || !out) return fail(E3D_ERR_ARG, "null argument");
	if (!m->preprocessed) return fail(E3D_ERR_ARG, "mesh is not preprocessed (call e3d_mesh_preprocess)");
	if (p->time_integration < 0 || p->time_integration > 2)
		return fail(E3D_ERR_ARG, "TIME_INTEGRATION must be 0 (explicit Euler), 1 (implicit Euler) or 2 (RK5)");
	if (p->scheme != 0 && p->scheme != 1) return fail(E3D_ERR_ARG, "SCHEME must be 0 (Roe) or 1 (AUSM)");
	if (!(p->cfl > 0) || !(p->sampling > 0) || p->max_iter < 0) return fail(E3D_ERR_ARG, "invalid CFL/SAMPLING/MAX_ITER");
	const e3d::SolverMesh &sm = m->sm;
	if ((long long) sm.nElem + sm.nGhost >= (1LL << 31) || (long long) sm.nFace >= (1LL << 31))
		return fail(E3D_ERR_ARG, "mesh too large for 32-bit indexing");
	return guarded([&] {
		auto c = std::make_unique<e3d_ctx>();
		int nDev = 0;
		hipCheck(hipGetDeviceCount(&nDev), "hipGetDeviceCount");
		if (device < 0 || device >= nDev) return fail(E3D_ERR_DEVICE, "no HIP device " + std::to_string(device));
		c->device = device;
		hipCheck(hipSetDevice(device), "hipSetDevice");
		// address of the allocation-free iteration's static scratch on this device (once per device)
		hipCheck(e3d::iterScratchInit(device), "hipGetSymbolAddress(scratch)");
		hipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
		hipCheck(hipEventCreate(&c->ev0), "hipEventCreate");
		hipCheck(hipEventCreate(&c->ev1), "hipEventCreate");
		c->prm = *p;
		fillFree(*p, c->fr);
		c->nElem = sm.nElem;
		c->nFace = sm.nFace;
		c->nFaceInt = sm.nFaceInt;
		c->nGhost = sm.nGhost;
		c->nTot = sm.nElem + sm.nGhost;
		const size_t nF = sm.nFace, nE = sm.nElem;
		hipStream_t s = c->stream;
		const char *orderEnv = std::getenv("E3D_CELL_ORDER");
		const char *pathEnv = std::getenv("E3D_STAGE_PATH");
		int order = e3d::ORDER_BISECT;
		if (orderEnv && std::string(orderEnv) == "reference") order = e3d::ORDER_REFERENCE;
		if (orderEnv && std::string(orderEnv) == "morton") order = e3d::ORDER_MORTON;
		c->fused = !(pathEnv && std::string(pathEnv) == "unfused");
		// launch-shape overrides for measurement (results are identical for every choice)
		if (const char *v = std::getenv("E3D_NONTEMPORAL")) c->tune.nontemporal = std::atoi(v) != 0;
		if (const char *v = std::getenv("E3D_XCDS")) c->tune.xcds = std::max(1, std::atoi(v));
		if (const char *v = std::getenv("E3D_SHARED_RECIP")) c->tune.sharedRecip = std::atoi(v) != 0;
		if (const char *v = std::getenv("E3D_ALTERNATE")) c->tune.alternate = std::atoi(v) != 0;
		if (const char *v = std::getenv("E3D_WIDE_THREADS")) c->tune.wideThreads = std::atoi(v) == 512 ? 512 : 768;
		if (const char *v = std::getenv("E3D_FIXED_SHAPE")) c->tune.fixedShape = std::max(-1, std::min(2, std::atoi(v)));
		// blocks of 256 cells, or of 128 where a 256-cell block does not fit in a CU's LDS;
		// E3D_BLOCK_CELLS=128|256 forces one. (Smaller blocks keep more waves resident on
		// hexahedral meshes with large halos, but the extra halo cells and duplicated faces cost
		// more: the 1.18M-hex wing ran 2.48 s against 2.27 s per 2,000 iterations.)
		int cells = 0;
		if (const char *v = std::getenv("E3D_BLOCK_CELLS")) cells = std::atoi(v) == 128 ? 128 : 256;
		// partitioned input: interface faces read the other partition's cell (default), or with
		// E3D_EXCHANGE=halo their ghosts, filled by the exchange of MPIHandler::updateFlowField
		if (const char *v = std::getenv("E3D_EXCHANGE")) c->halo = std::string(v) == "halo" && !sm.haloSrc.empty();
		e3d::HostPlan P = e3d::buildPlan(sm, order, cells ? cells : e3d::kCellsPerBlock, c->halo);
		if (!cells && e3d::stageFusedWavesPerCU(256, P.maxSlots, P.maxHalo, p->scheme) == 0)
			P = e3d::buildPlan(sm, order, 128, c->halo);
		if (c->halo) {
			c->nX = (int) P.xSrc.size();
			c->xSrc.alloc(c->nX), c->xSrc.upload(P.xSrc.data(), c->nX, s);
			c->xDst.alloc(c->nX), c->xDst.upload(P.xDst.data(), c->nX, s);
			c->xSlot.alloc(2 * (size_t) c->nX), c->xSlot.upload(P.xSlot.data(), 2 * (size_t) c->nX, s);
			c->xBuf.alloc(std::max(1, P.xBufLen));
		}
		if (e3d::stageFusedWavesPerCU(P.cellsPerBlock, P.maxSlots, P.maxHalo, p->scheme) == 0)
			return fail(E3D_ERR_UNSUPPORTED, "a block of the fused stage needs more LDS than a CU has");
		c->new2old = P.new2old;

		// faces in reference order, device cell ids
		c->e0.alloc(nF), c->e0.upload(P.fe0.data(), nF, s);
		c->e1.alloc(nF), c->e1.upload(P.fe1.data(), nF, s);
		c->nx.alloc(nF), c->nx.upload(P.fnx.data(), nF, s);
		c->ny.alloc(nF), c->ny.upload(P.fny.data(), nF, s);
		c->nz.alloc(nF), c->nz.upload(P.fnz.data(), nF, s);
		c->S.alloc(nF), c->S.upload(P.fS.data(), nF, s);
		c->gkind.alloc(sm.nGhost), c->gkind.upload(P.gkind.data(), sm.nGhost, s);
		// cells in device order
		c->vol.alloc(nE), c->vol.upload(P.vol.data(), nE, s);
		c->sx.alloc(nE), c->sx.upload(P.sx.data(), nE, s);
		c->sy.alloc(nE), c->sy.upload(P.sy.data(), nE, s);
		c->sz.alloc(nE), c->sz.upload(P.sz.data(), nE, s);
		c->fstart.alloc(nE + 1), c->fstart.upload(P.fstart.data(), nE + 1, s);
		c->flist.alloc(P.flist.size()), c->flist.upload(P.flist.data(), P.flist.size(), s);
		// fused-stage plan
		size_t nS = P.se0.size();
		c->blockSlot.alloc(P.blockSlot.size()), c->blockSlot.upload(P.blockSlot.data(), P.blockSlot.size(), s);
		c->se0.alloc(nS), c->se0.upload(P.se0.data(), nS, s);
		c->se1.alloc(nS), c->se1.upload(P.se1.data(), nS, s);
		c->skind.alloc(nS), c->skind.upload(P.skind.data(), nS, s);
		c->snx.alloc(nS), c->snx.upload(P.snx.data(), nS, s);
		c->sny.alloc(nS), c->sny.upload(P.sny.data(), nS, s);
		c->snz.alloc(nS), c->snz.upload(P.snz.data(), nS, s);
		c->sS.alloc(nS), c->sS.upload(P.sS.data(), nS, s);
		c->csr.alloc(P.csr.size()), c->csr.upload(P.csr.data(), P.csr.size(), s);
		c->spack.alloc(nS), c->spack.upload(P.spack.data(), nS, s);
		c->blockHalo.alloc(P.blockHalo.size()), c->blockHalo.upload(P.blockHalo.data(), P.blockHalo.size(), s);
		c->haloCell.alloc(std::max<size_t>(1, P.haloCell.size())), c->haloCell.upload(P.haloCell.data(), P.haloCell.size(), s);
		c->plan = {P.nBlocks, P.maxSlots, P.csrStride, c->blockSlot.p, c->se0.p, c->se1.p, c->skind.p,
		           c->snx.p, c->sny.p, c->snz.p, c->sS.p, c->csr.p, c->spack.p, c->blockHalo.p, c->haloCell.p,
		           P.maxHalo, P.cellsPerBlock, nullptr};
		{
			std::vector<int> meta(4 * (size_t) P.nBlocks);
			for (int b = 0; b < P.nBlocks; b++) {
				int nInt = 0;
				while (P.blockSlot[b] + nInt < P.blockSlot[b + 1] && P.skind[P.blockSlot[b] + nInt] == 0) nInt++;
				meta[4 * (size_t) b] = P.blockSlot[b];
				meta[4 * (size_t) b + 1] = (P.blockSlot[b + 1] - P.blockSlot[b]) | nInt << 16;
				meta[4 * (size_t) b + 2] = P.blockHalo[b];
				meta[4 * (size_t) b + 3] = P.blockHalo[b + 1] - P.blockHalo[b];
			}
			c->blockMeta.alloc(std::max<size_t>(4, meta.size()));
			c->blockMeta.upload(meta.data(), meta.size(), s);
			c->plan.blockMeta = reinterpret_cast<const int4 *>(c->blockMeta.p);
		}
		// The persistent, software-pipelined stage where a block's LDS leaves room for one
		// workgroup per CU (there no other workgroup hides a block's load phase; on the 1.18M-hex
		// wing 135 against 157 us per stage). Where several workgroups fit, they hide each other's
		// loads better than the pipeline's two (6.3M-tet bump: 381 against 437 us). E3D_PIPE=0|1
		// turns it off or on for every plan it fits.
		{
			const char *v = std::getenv("E3D_PIPE");
			const bool wideLds = 2 * e3d::stageFusedLdsBytes(P.cellsPerBlock, P.maxSlots, P.maxHalo, 2) > 160 * 1024;
			const bool want = v ? std::atoi(v) != 0 : wideLds;
			if (want) c->tune.pipeGrid = e3d::stagePipeGrid(c->plan, p->scheme, c->device, c->tune.xcds);
		}
		// partitions: wall ranges, interface ghosts
		c->nPart = sm.partStart.size() > 1 ? (int) sm.partStart.size() - 1 : 1;
		if (!m->ranges.empty()) {
			for (auto &R : m->ranges) c->wall0.push_back(R.wall0), c->nWallPart.push_back(R.nWall);
		} else {
			c->wall0 = {0}, c->nWallPart = {(int) P.wallCell.size()};
		}
		c->mpiGhost = sm.mpiGhost, c->mpiCell = sm.mpiCell;
		c->aeroOut.alloc(6 * (size_t) c->nPart);
		// wall list in the reference's order for the force integration
		c->nWall = (int) P.wallCell.size();
		if (c->nWall) {
			c->wallCp.alloc(c->nWall);
			int most = 0;
			for (int w : c->nWallPart) most = std::max(most, w);
			c->aeroTmp.alloc(e3d::aeroScratch(most));
			c->wallCell.alloc(c->nWall), c->wallCell.upload(P.wallCell.data(), c->nWall, s);
			for (int k = 0; k < 3; k++) {
				c->wN[k].alloc(c->nWall), c->wN[k].upload(P.wallN[k].data(), c->nWall, s);
				c->wC[k].alloc(c->nWall), c->wC[k].upload(P.wallC[k].data(), c->nWall, s);
			}
		}
		c->sync();
		double uinf2 = c->fr.u_inf * c->fr.u_inf + c->fr.v_inf * c->fr.v_inf + c->fr.w_inf * c->fr.w_inf;
		c->cpFactor = 1 / (0.5 * c->fr.rho_inf * uinf2);

		for (auto &b : c->st) b.alloc((size_t) c->nTot * e3d::kStateWidth);
		for (auto &b : c->wst) b.alloc((size_t) 5 * nE + 2);
		c->W0.alloc(5 * nE);
		if (p->residual_smoothing) {
			// neighbour of each local face in device numbering (EulerSolver.cpp:495-503 walks
			// GetElement2ElementID in this order); -2 marks a ghost, -1 pads
			int stride = 0;
			for (int e = 0; e < sm.nElem; e++)
				stride = std::max(stride, sm.element2faceStart[e + 1] - sm.element2faceStart[e]);
			std::vector<int> nb((size_t) nE * stride, -1);
			std::vector<uint8_t> nf(nE);
			for (int d = 0; d < sm.nElem; d++) {
				int old = P.new2old[d];
				int b = sm.element2faceStart[old], en = sm.element2faceStart[old + 1];
				nf[d] = (uint8_t) (en - b);
				for (int j = b; j < en; j++) {
					int s2 = sm.element2element[j];
					nb[(size_t) d * stride + (j - b)] = s2 < sm.nElem ? P.old2new[s2] : -2;
				}
			}
			c->nbrStride = stride;
			c->nbr.alloc(nb.size()), c->nbr.upload(nb.data(), nb.size(), s);
			c->nface.alloc(nE), c->nface.upload(nf.data(), nE, s);
			for (auto *b : {&c->R0, &c->R1, &c->R2}) b->alloc(5 * nE);
			c->blockMax.alloc(std::max(1, e3d::stageBlocks(c->nElem)));
			std::vector<int> zero(c->nPart + 1, 0);
			c->smoothFlag.alloc(zero.size());
			c->smoothFlag.upload(zero.data(), zero.size(), s);
			if (c->nPart > 1) {
				// partitions keep their reference ranges in device order (the cell order groups them)
				c->cellMax.alloc(nE);
				c->partStart.alloc(sm.partStart.size()), c->partStart.upload(sm.partStart.data(), sm.partStart.size(), s);
			}
			c->sync();
		}
		c->fs.alloc(5 * nF);
		c->partial.alloc(std::max({1, e3d::stageBlocks(c->nElem), P.nBlocks}));
		c->sumTmp.alloc(e3d::sumScratch((int) c->partial.n));
		c->scalars.alloc(16);

		c->dm.nElem = c->nElem;
		c->dm.nFace = c->nFace;
		c->dm.nFaceInt = c->nFaceInt;
		c->dm.nGhost = c->nGhost;
		c->dm.f = {c->e0.p, c->e1.p, c->nx.p, c->ny.p, c->nz.p, c->S.p};
		c->dm.c = {c->vol.p, c->sx.p, c->sy.p, c->sz.p, c->fstart.p, c->flist.p};
		c->dm.gkind = c->gkind.p;
		initFreestream(c.get());
		*out = c.release();
		return E3D_OK;
	});
}

int e3d_set_state(e3d_ctx *c, const double *const fields[8]) {
	if (!c || !fields) return fail(E3D_ERR_ARG, "null argument");
	for (int k = 0; k < 8; k++)
		if (!fields[k]) return fail(E3D_ERR_ARG, "null state field");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		std::vector<double> h((size_t) c->nTot * e3d::kStateWidth);
		for (int k = 0; k < 8; k++) toDevice(c, fields[k], k, h);
		c->st[c->cur].upload(h.data(), h.size(), c->stream);
		std::vector<double> w(5 * (size_t) c->nElem);
		const size_t n = (size_t) c->nElem;
		for (size_t i = 0; i < n; i++) {
			const double *cell = &h[i * e3d::kStateWidth];
			const double rho = cell[e3d::S_RHO];
			w[i] = rho, w[n + i] = cell[e3d::S_U] * rho, w[2 * n + i] = cell[e3d::S_V] * rho;
			w[3 * n + i] = cell[e3d::S_W] * rho, w[4 * n + i] = cell[e3d::S_E] * rho;
		}
		c->W0.upload(w.data(), w.size(), c->stream);
		c->sync();
		c->machDerived = 0;
		c->wMode = false;
		c->exchanged = false;
		return E3D_OK;
	});
}

int e3d_get_state(e3d_ctx *c, double *const fields[8]) {
	if (!c || !fields) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		c->ensurePrims();
		if (c->machDerived) {
			e3d::launchMach(c->dm, c->state(), c->fr, c->stream);
			c->check();
		}
		std::vector<double> h((size_t) c->nTot * e3d::kStateWidth);
		hipCheck(hipMemcpyAsync(h.data(), c->st[c->cur].p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream),
		         "hipMemcpyAsync D2H");
		c->sync();
		for (int k = 0; k < 8; k++) {
			if (!fields[k]) continue;
			toReference(c, h, k, fields[k]);
			// an interface ghost holds the copy of its own cell that the last update left in it
			// (FlowField.cpp:129-138, 159-169), unless an exchange has filled it since
			if (c->halo && c->exchanged) continue;
			for (size_t i = 0; i < c->mpiGhost.size(); i++) fields[k][c->mpiGhost[i]] = fields[k][c->mpiCell[i]];
		}
		return E3D_OK;
	});
}

int e3d_update_bc(e3d_ctx *c) {
	if (!c) return fail(E3D_ERR_ARG, "null context");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		c->ensurePrims();
		e3d::launchBc(c->dm, c->state(), c->fr, c->machDerived, c->stream);
		c->check();
		c->exchange();
		c->exchanged = c->halo;
		c->sync();
		return E3D_OK;
	});
}

namespace {
void toComponentMinor(const std::vector<double> &soa, size_t n, double *out) {
	for (size_t i = 0; i < n; i++)
		for (int k = 0; k < 5; k++) out[5 * i + k] = soa[k * n + i];
}
}// namespace

int e3d_face_flux(e3d_ctx *c, double *F, double *FS) {
	if (!c) return fail(E3D_ERR_ARG, "null context");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		DBuf<double> raw;
		raw.alloc(5 * (size_t) c->nFace);
		c->ensurePrims();
		e3d::launchFlux(c->dm, c->state(), c->fr, c->prm.scheme, c->fs.p, raw.p, c->stream);
		c->check();
		std::vector<double> h(5 * (size_t) c->nFace);
		if (F) {
			hipCheck(hipMemcpyAsync(h.data(), raw.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream), "D2H");
			c->sync();
			toComponentMinor(h, c->nFace, F);
		}
		if (FS) {
			hipCheck(hipMemcpyAsync(h.data(), c->fs.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream), "D2H");
			c->sync();
			toComponentMinor(h, c->nFace, FS);
		}
		c->sync();
		return E3D_OK;
	});
}

int e3d_residual(e3d_ctx *c, double *R) {
	if (!c || !R) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		DBuf<double> dR;
		dR.alloc(5 * (size_t) c->nElem);
		c->ensurePrims();
		e3d::launchFlux(c->dm, c->state(), c->fr, c->prm.scheme, c->fs.p, nullptr, c->stream);
		e3d::launchGather(c->dm, c->fs.p, dR.p, c->stream);
		c->check();
		std::vector<double> h(5 * (size_t) c->nElem);
		hipCheck(hipMemcpyAsync(h.data(), dR.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream), "D2H");
		c->sync();
		std::vector<double> dev(5 * (size_t) c->nElem);
		toComponentMinor(h, c->nElem, dev.data());
		for (int i = 0; i < c->nElem; i++)
			for (int k = 0; k < 5; k++) R[5 * (size_t) c->new2old[i] + k] = dev[5 * (size_t) i + k];
		return E3D_OK;
	});
}

int e3d_timestep(e3d_ctx *c, double *dt) {
	if (!c || !dt) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		DBuf<double> d;
		d.alloc(c->nElem);
		c->ensurePrims();
		e3d::launchTimestep(c->dm, c->state(), c->fr, d.p, c->stream);
		c->check();
		std::vector<double> h(c->nElem);
		hipCheck(hipMemcpyAsync(h.data(), d.p, sizeof(double) * c->nElem, hipMemcpyDeviceToHost, c->stream), "D2H");
		c->sync();
		for (int i = 0; i < c->nElem; i++) dt[c->new2old[i]] = h[i];
		return E3D_OK;
	});
}

int e3d_iterate(e3d_ctx *c, int n, double *rms_sum) {
	if (!c || n < 0) return fail(E3D_ERR_ARG, "bad argument");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		c->exchanged = false;
		DBuf<double> sums;
		sums.alloc(std::max(1, n));
		hipCheck(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
		for (int it = 0; it < n; it++) runIteration(c, sums.p + it);
		hipCheck(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
		c->check();
		c->sync();
		checkSmoothing(c);
		float ms = 0;
		hipCheck(hipEventElapsedTime(&ms, c->ev0, c->ev1), "hipEventElapsedTime");
		c->lastMs = ms;
		if (rms_sum && n > 0)
			hipCheck(hipMemcpy(rms_sum, sums.p, sizeof(double) * n, hipMemcpyDeviceToHost), "D2H");
		return E3D_OK;
	});
}

int e3d_iterate_noalloc(e3d_ctx *c, int n, double *rms_sum) {
	if (!c || n < 0) return fail(E3D_ERR_ARG, "bad argument");
	if (n > e3d::kIterScratchMax)
		return fail(E3D_ERR_ARG, "e3d_iterate_noalloc: " + std::to_string(n) + " iterations exceed the " +
		                             std::to_string(e3d::kIterScratchMax) + " its static scratch holds");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		c->exchanged = false;
		double *sums = e3d::iterScratch(c->device);
		if (!sums) return fail(E3D_ERR_DEVICE, "static scratch not set up on this device");
		hipCheck(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
		// every iteration writes its own sum: nothing in the scratch needs clearing first
		for (int it = 0; it < n; it++) runIteration(c, sums + it, false);
		hipCheck(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
		c->check();
		c->sync();
		checkSmoothing(c);
		float ms = 0;
		hipCheck(hipEventElapsedTime(&ms, c->ev0, c->ev1), "hipEventElapsedTime");
		c->lastMs = ms;
		if (rms_sum && n > 0)
			hipCheck(hipMemcpy(rms_sum, sums, sizeof(double) * n, hipMemcpyDeviceToHost), "D2H");
		return E3D_OK;
	});
}

int e3d_aero(e3d_ctx *c, double cf[3], double cm[3]) {
	if (!c || !cf || !cm) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		// AeroCoefficients::Update per partition (AeroCoefficients.cpp:38-52), then the partitions'
		// force and moment vectors summed in rank order (MPIHandler::UpdateAerodynamicCoefficients,
		// MPIHandler.cpp:268-313)
		std::vector<double> h(6 * (size_t) c->nPart, 0.0);
		bool any = false;
		for (int r = 0; r < c->nPart; r++) {
			if (!c->nWallPart[r]) continue;
			const int w0 = c->wall0[r];
			e3d::launchAero(c->nWallPart[r], c->wallCell.p + w0, c->wN[0].p + w0, c->wN[1].p + w0, c->wN[2].p + w0,
			                c->wC[0].p + w0, c->wC[1].p + w0, c->wC[2].p + w0, c->state(),
			                c->wMode ? c->wst[c->cur].p : nullptr, c->fr.gamma, c->cpFactor, c->fr.p_inf,
			                c->prm.cref, c->prm.sref, c->aeroTmp.p, c->aeroOut.p + 6 * (size_t) r, c->wallCp.p + w0,
			                c->stream);
			c->check();
			any = true;
		}
		if (any) {
			hipCheck(hipMemcpyAsync(h.data(), c->aeroOut.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream), "D2H");
			c->sync();
			c->wallCpSet = true;
		}
		for (int r = 0; r < c->nPart; r++)
			if (!c->nWallPart[r]) std::fill(h.begin() + 6 * r, h.begin() + 6 * r + 6, 0.0); // no wall face: zero sums
		for (int r = 0; r < c->nPart; r++) {
			const double *q = &h[6 * (size_t) r];
			// AeroCoefficients.cpp:49-51: lever = ref - CP (CP never divided), M = F x lever / Cref
			double lever[3] = {c->prm.ref_point[0] - q[3], c->prm.ref_point[1] - q[4], c->prm.ref_point[2] - q[5]};
			double M[3] = {q[1] * lever[2] - q[2] * lever[1], q[2] * lever[0] - q[0] * lever[2], q[0] * lever[1] - q[1] * lever[0]};
			for (int k = 0; k < 3; k++) {
				const double mk = M[k] / c->prm.cref;
				cf[k] = r ? cf[k] + q[k] : q[k];
				cm[k] = r ? cm[k] + mk : mk;
			}
		}
		return E3D_OK;
	});
}

int e3d_wall_cp(e3d_ctx *c, double *cp, long long capacity, long long *n) {
	if (!c) return fail(E3D_ERR_ARG, "null context");
	return guarded([&] {
		if (n) *n = c->nWall;
		if (!cp || !c->nWall) return E3D_OK;
		if (capacity < c->nWall) return fail(E3D_ERR_ARG, "wall cp: capacity below the wall face count");
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		if (!c->wallCpSet) {
			// AeroCoefficients' cp vector before its first Update: zeros (AeroCoefficients.cpp:27)
			std::fill(cp, cp + c->nWall, 0.0);
			return E3D_OK;
		}
		hipCheck(hipMemcpyAsync(cp, c->wallCp.p, sizeof(double) * c->nWall, hipMemcpyDeviceToHost, c->stream), "D2H");
		c->sync();
		return E3D_OK;
	});
}

int e3d_run(e3d_ctx *c, double *rows, int max_rows, int *n_rows, int *iters, double coeffs[3], double cfOut[3],
            double cmOut[3]) {
	if (!c) return fail(E3D_ERR_ARG, "null context");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		c->exchanged = false;
		const e3d_params &q = c->prm;
		int nb = 0, counter = 0, nrows = 0;
		double CL = 0, CD = 0, CM = 0, cf[3] = {0, 0, 0}, cm[3] = {0, 0, 0};
		auto row = [&](int it, double rms) {
			CL = cf[q.orient_cl[0]] * q.orient_cl[1];
			CD = cf[q.orient_cd[0]] * q.orient_cd[1];
			CM = cm[q.orient_cm[0]] * q.orient_cm[1];
			if (rows && nrows < max_rows) {
				double *r = rows + 8 * (size_t) nrows;
				r[0] = it, r[1] = rms;
				for (int k = 0; k < 3; k++) r[2 + k] = cf[k], r[5 + k] = cm[k];
			}
			nrows++;
		};
		double *dsum = c->scalars.p + 8;
		hipCheck(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
		while (nb < q.max_iter) {
			runIteration(c, dsum);
			c->check();
			checkSmoothing(c);
			double sum = 0;
			hipCheck(hipMemcpyAsync(&sum, dsum, sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
			c->sync();
			double rms = std::sqrt(sum / c->nElem);
			if (rms < q.min_residual) {
				row(nb, rms);
				break;
			}
			nb += 1;
			if (nb % q.sampling == 0) {
				double oCL = CL, oCD = CD, oCM = CM;
				int rc = e3d_aero(c, cf, cm);
				if (rc != E3D_OK) return rc;
				row(nb, rms);
				double tol = q.min_aerocoeff_error;
				bool crit = std::fabs(oCL - CL) < tol && std::fabs(oCD - CD) < tol && std::fabs(oCM - CM) < tol;
				counter = crit ? counter + 1 : 0;
				if (counter >= 2) break;
			}
		}
		hipCheck(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
		c->sync();
		float ms = 0;
		hipCheck(hipEventElapsedTime(&ms, c->ev0, c->ev1), "hipEventElapsedTime");
		c->lastMs = ms;
		if (n_rows) *n_rows = nrows;
		if (iters) *iters = nb;
		if (coeffs) coeffs[0] = CL, coeffs[1] = CD, coeffs[2] = CM;
		for (int k = 0; k < 3; k++) {
			if (cfOut) cfOut[k] = cf[k];
			if (cmOut) cmOut[k] = cm[k];
		}
		return E3D_OK;
	});
}

namespace {
constexpr char kCkptMagic[8] = {'E', '3', 'D', 'C', 'K', 'P', 'T', '1'};
struct CkptHeader {
	char magic[8];
	int32_t nElem, nGhost, nFace, machDerived;
	int32_t timeIntegration, scheme, pad[2];
};
}// namespace

int e3d_checkpoint_write(e3d_ctx *c, const char *path) {
	if (!c || !path) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		const size_t nT = (size_t) c->nTot, nE = (size_t) c->nElem;
		std::vector<double> st(nT * e3d::kStateWidth), w(5 * nE);
		c->ensurePrims();
		hipCheck(hipMemcpyAsync(st.data(), c->st[c->cur].p, st.size() * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
		hipCheck(hipMemcpyAsync(w.data(), c->W0.p, w.size() * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
		c->sync();
		CkptHeader h{};
		std::memcpy(h.magic, kCkptMagic, 8);
		h.nElem = c->nElem, h.nGhost = c->nGhost, h.nFace = c->nFace, h.machDerived = c->machDerived;
		h.timeIntegration = c->prm.time_integration, h.scheme = c->prm.scheme;
		std::FILE *f = std::fopen(path, "wb");
		if (!f) return fail(E3D_ERR_IO, std::string("cannot open ") + path);
		bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
		std::vector<double> col(nT);
		for (int k = 0; k < e3d::kStateWidth && ok; k++) {
			toReference(c, st, k, col.data());
			ok = std::fwrite(col.data(), 8, nT, f) == nT;
		}
		std::vector<double> wr(nE);
		for (int k = 0; k < 5 && ok; k++) {
			for (size_t i = 0; i < nE; i++) wr[c->new2old[i]] = w[k * nE + i];
			ok = std::fwrite(wr.data(), 8, nE, f) == nE;
		}
		ok = (std::fclose(f) == 0) && ok;
		if (!ok) return fail(E3D_ERR_IO, std::string("writing ") + path);
		return E3D_OK;
	});
}

int e3d_checkpoint_read(e3d_ctx *c, const char *path) {
	if (!c || !path) return fail(E3D_ERR_ARG, "null argument");
	return guarded([&] {
		hipCheck(hipSetDevice(c->device), "hipSetDevice");
		c->exchanged = false;
		std::FILE *f = std::fopen(path, "rb");
		if (!f) return fail(E3D_ERR_IO, std::string("cannot open ") + path);
		CkptHeader h{};
		const size_t nT = (size_t) c->nTot, nE = (size_t) c->nElem;
		bool ok = std::fread(&h, sizeof(h), 1, f) == 1;
		if (ok && (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.nElem != c->nElem || h.nGhost != c->nGhost ||
		           h.nFace != c->nFace)) {
			std::fclose(f);
			return fail(E3D_ERR_ARG, std::string(path) + " is not a checkpoint of this mesh");
		}
		std::vector<double> st(nT * e3d::kStateWidth), col(nT), w(5 * nE), wr(nE);
		for (int k = 0; k < e3d::kStateWidth && ok; k++) {
			ok = std::fread(col.data(), 8, nT, f) == nT;
			if (ok) toDevice(c, col.data(), k, st);
		}
		for (int k = 0; k < 5 && ok; k++) {
			ok = std::fread(wr.data(), 8, nE, f) == nE;
			if (ok)
				for (size_t i = 0; i < nE; i++) w[k * nE + i] = wr[c->new2old[i]];
		}
		std::fclose(f);
		if (!ok) return fail(E3D_ERR_IO, std::string("truncated checkpoint ") + path);
		c->st[c->cur].upload(st.data(), st.size(), c->stream);
		c->W0.upload(w.data(), w.size(), c->stream);
		c->sync();
		c->machDerived = h.machDerived;
		c->wMode = false;
		return E3D_OK;
	});
}

int e3d_math_selfcheck(int device, long long n, unsigned long long seed, unsigned long long mismatches[3]) {
	if (!mismatches || n < 0) return fail(E3D_ERR_ARG, "bad argument");
	return guarded([&] {
		hipCheck(hipSetDevice(device), "hipSetDevice");
		unsigned long long *bad = nullptr;
		hipCheck(hipMalloc((void **) &bad, 3 * sizeof(unsigned long long)), "hipMalloc");
		hipError_t e = hipMemset(bad, 0, 3 * sizeof(unsigned long long));
		if (e == hipSuccess) {
			e3d::launchMathCheck(seed, n, bad, nullptr);
			e = hipGetLastError();
		}
		if (e == hipSuccess) e = hipMemcpy(mismatches, bad, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
		(void) hipFree(bad);
		hipCheck(e, "math self-check");
		return E3D_OK;
	});
}

int e3d_copy_probe(int device, long long bytes, double *ms) {
	if (bytes <= 0 || bytes % 16) return fail(E3D_ERR_ARG, "copy probe: bytes must be a positive multiple of 16");
	return guarded([&] {
		hipCheck(hipSetDevice(device), "hipSetDevice");
		void *src = nullptr, *dst = nullptr;
		hipEvent_t a = nullptr, z = nullptr;
		hipError_t e = hipMalloc(&src, (size_t) bytes);
		if (e == hipSuccess) e = hipMalloc(&dst, (size_t) bytes);
		if (e == hipSuccess) e = hipMemset(src, 1, (size_t) bytes);
		if (e == hipSuccess) e = hipEventCreate(&a);
		if (e == hipSuccess) e = hipEventCreate(&z);
		if (e == hipSuccess) e = hipDeviceSynchronize();
		if (e == hipSuccess) e = hipEventRecord(a, nullptr);
		if (e == hipSuccess) {
			e3d::launchCopy16(src, dst, bytes / 16, nullptr);
			e = hipGetLastError();
		}
		if (e == hipSuccess) e = hipEventRecord(z, nullptr);
		if (e == hipSuccess) e = hipEventSynchronize(z);
		float t = 0;
		if (e == hipSuccess) e = hipEventElapsedTime(&t, a, z);
		if (a) (void) hipEventDestroy(a);
		if (z) (void) hipEventDestroy(z);
		if (src) (void) hipFree(src);
		if (dst) (void) hipFree(dst);
		hipCheck(e, "copy probe");
		if (ms) *ms = t;
		return E3D_OK;
	});
}

int e3d_last_elapsed_ms(e3d_ctx *c, double *ms) {
	if (!c || !ms) return fail(E3D_ERR_ARG, "null argument");
	*ms = c->lastMs;
	return E3D_OK;
}

int e3d_set_kernel_timing(e3d_ctx *c, int enable) {
	if (!c) return fail(E3D_ERR_ARG, "null context");
	c->timing = enable != 0;
	c->evUsed = 0;
	return E3D_OK;
}

int e3d_kernel_timing(e3d_ctx *c, double *avg_ms, int *launches, char *name, int name_len) {
	if (!c) return fail(E3D_ERR_ARG, "null context");
	return guarded([&] {
		c->sync();
		double total = 0.0;
		int n = (int) (c->evUsed / 2);
		for (int i = 0; i < n; i++) {
			float ms = 0;
			hipCheck(hipEventElapsedTime(&ms, c->evPool[2 * i], c->evPool[2 * i + 1]), "hipEventElapsedTime");
			total += ms;
		}
		if (avg_ms) *avg_ms = n ? total / n : 0.0;
		if (launches) *launches = n;
		if (name && name_len > 0) {
			std::strncpy(name, stageKernelName(c), (size_t) name_len - 1);
			name[name_len - 1] = 0;
		}
		c->evUsed = 0;
		return E3D_OK;
	});
}

void e3d_destroy(e3d_ctx *c) {
	if (!c) return;
	(void) hipSetDevice(c->device);
	if (c->stream) (void) hipStreamSynchronize(c->stream);
	for (auto &b : c->st) b.free();
	c->W0.free(), c->fs.free(), c->partial.free(), c->scalars.free();
	if (c->ev0) (void) hipEventDestroy(c->ev0);
	if (c->ev1) (void) hipEventDestroy(c->ev1);
	for (auto e : c->evPool) (void) hipEventDestroy(e);
	if (c->stream) (void) hipStreamDestroy(c->stream);
	delete c;
}

}// extern "C"
