// Golden-vector harness around the reference Euler3D solver.
//
// TEST INFRASTRUCTURE ONLY. This file is compiled (by oracle/ref.mk) together with the
// reference sources where they lie, into oracle/_ref/, and is run only in the container that
// holds the reference, to write the fixtures under tests/golden/. Nothing in the product
// links or loads it.
//
// It drives the reference's own objects (SimConfig, Mesh<MeshPartition>, Metrics, FlowField,
// EulerSolver, Roe, Fc, TimeStepMethod2) and prints every array it is asked for as C99 hex
// floats, so the fixtures carry exact FP64 bits.
//
// Modes:
//   mesh <cfg.e3d> <out.txt>            connectivity, metrics and ghost lists
//                                        (src/connectivity/Connectivity.cpp:70-262,
//                                         include/mesh/Mesh.hpp:471-707, src/mesh/Metrics.cpp:9-309)
//   flux <cfg.e3d> <out.txt> <seed>     seeded non-uniform state -> updateBC -> per-face flux,
//                                        computeResidual, TimeStepMethod2, one RungeKutta step
//                                        (src/solver/EulerSolver.cpp:201-359, 425-472)
//   run  <cfg.e3d> <out.txt>            EulerSolver::Run() for MAX_ITER iterations, final state,
//                                        last residual and coefficients (EulerSolver.cpp:60-167)
// Under mpiexec -np N (partitioned input) every rank writes <out.txt>.r<rank> for its partition.

// Open the private members of the reference classes to this harness only: the hot-path pieces
// (updateBC, computeResidual, RungeKutta, _residuals) are private in EulerSolver.hpp.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>
#include <mpi.h>
#define private public
#define protected public
#include "mesh/Mesh.hpp"
#include "mesh/Metrics.hpp"
#include "parallelization/MPIHandler.hpp"
#include "parser/SimConfig.hpp"
#include "solver/ConvectiveFluxesSchemes.h"
#include "solver/EulerSolver.hpp"
#include "solver/FlowField.hpp"
#include "solver/TimeIntegration.h"
#undef private
#undef protected

using namespace E3D;

namespace {

struct SplitMix64 {
	uint64_t s;
	explicit SplitMix64(uint64_t seed) : s(seed) {}
	uint64_t next() {
		uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
		return z ^ (z >> 31);
	}
	double uniform() { return (double) (next() >> 11) * 0x1.0p-53; }
};

struct Dump {
	FILE *f;
	explicit Dump(const char *path) : f(std::fopen(path, "w")) {
		if (!f) {
			std::perror(path);
			std::exit(2);
		}
	}
	~Dump() { std::fclose(f); }
	void ints(const char *name, const std::vector<int> &v) {
		std::fprintf(f, "I %s %zu\n", name, v.size());
		for (int x : v) std::fprintf(f, "%d\n", x);
	}
	void dbls(const char *name, const std::vector<double> &v) {
		std::fprintf(f, "D %s %zu\n", name, v.size());
		for (double x : v) std::fprintf(f, "%a\n", x);
	}
	void vec3(const char *name, const std::vector<Vector3<double>> &v) {
		std::fprintf(f, "D %s %zu\n", name, v.size() * 3);
		for (const auto &x : v) std::fprintf(f, "%a\n%a\n%a\n", x.x, x.y, x.z);
	}
	void scalar(const char *name, double x) {
		std::fprintf(f, "D %s 1\n%a\n", name, x);
	}
};

struct Setup {
	Parallel::MPIHandler &mpi;
	Parser::SimConfig config;
	Mesh<Parser::MeshPartition> mesh;
	std::unique_ptr<Metrics> metrics;
	std::unique_ptr<Solver::FlowField> field;
	std::unique_ptr<Solver::EulerSolver> solver;

	Setup(Parallel::MPIHandler &m, const std::string &cfg)
	    : mpi(m), config(cfg, m.getRankID(), m.getPoolSize()),
	      mesh(config.getPartitionedMeshFiles()[m.getRankID()], m.getRankID()) {
		mesh.solveConnectivity();
		mpi.updateRequesterID(mesh.getMPIelements());
		mpi.sortInterface();
		metrics = std::make_unique<Metrics>(mesh, mpi.getRankID(), config);
		field = std::make_unique<Solver::FlowField>(config, mesh);
		solver = std::make_unique<Solver::EulerSolver>(*field, mpi, mesh, config, *metrics);
	}

	int nElem() const { return mesh.GetnElemTot(); }
	int nTot() const { return (int) field->Getrho().size(); }
};

void dumpState(Dump &d, Solver::FlowField &ff, const char *prefix) {
	std::string p(prefix);
	d.dbls((p + "rho").c_str(), ff.Getrho());
	d.dbls((p + "u").c_str(), ff.GetU_Velocity());
	d.dbls((p + "v").c_str(), ff.GetV_Velocity());
	d.dbls((p + "w").c_str(), ff.GetW_Velocity());
	d.dbls((p + "p").c_str(), ff.GetP());
	d.dbls((p + "E").c_str(), ff.GetE());
	d.dbls((p + "H").c_str(), ff.GetH());
	d.dbls((p + "M").c_str(), ff.GetMach());
}

void dumpResiduals(Dump &d, const std::vector<Solver::ResidualVar> &r, const char *name) {
	std::vector<double> out;
	out.reserve(r.size() * 5);
	for (const auto &x : r) {
		out.push_back(x.m_rhoV_residual);
		out.push_back(x.m_rho_uV_residual);
		out.push_back(x.m_rho_vV_residual);
		out.push_back(x.m_rho_wV_residual);
		out.push_back(x.m_rho_HV_residual);
	}
	d.dbls(name, out);
}

void modeMesh(Setup &s, Dump &d) {
	auto &m = s.mesh;
	d.ints("counts", {m.GetnElemTot(), m.GetnFace(), m.GetnFaceInt(), (int) s.field->Getrho().size()});
	d.ints("face2element", *m.face2element);
	d.ints("face2nodeStart", *m.face2nodeStart);
	d.ints("face2node", *m.face2node);
	d.ints("element2faceStart", *m.element2faceStart);
	d.ints("element2face", *m.element2face);
	d.ints("element2element", *m.element2element);
	d.ints("wallGhost", m.GetWallGhostCellsIDs());
	d.ints("wallCell", m.GetWallAdjacentGhostCellIDs());
	d.ints("wallFace", m.GetWallAdjacentFaceIDs());
	d.ints("farGhost", m.GetFarfieldGhostCellsIDs());
	d.ints("farCell", m.GetFarfieldAdjacentToGhostCellsIDs());
	d.ints("farFace", m.GetFarfieldAdjacentFaceIDs());
	d.ints("symGhost", m.GetSymmetryGhostCellsIDs());
	d.ints("symCell", m.GetSymmetryAdjacentGhostCellIDs());
	d.ints("symFace", m.GetSymmetryAdjacentFaceIDs());
	d.ints("mpiGhost", m.GetMPIGhostCellsIDs());
	d.ints("mpiCell", m.getMPIadjacentToGhostCellIDs());
	auto &mt = *s.metrics;
	d.vec3("faceNormals", mt.getFaceNormals());
	d.vec3("faceNormalsUnit", mt.getFaceNormalsUnit());
	d.dbls("faceSurfaces", mt.getFaceSurfaces());
	d.vec3("faceCenters", mt.getFaceCenters());
	d.dbls("cellVolumes", mt.getCellVolumes());
	d.vec3("cellCentroids", mt.getCellCentroids());
	d.scalar("u_inf", s.field->getu_inf());
	d.scalar("v_inf", s.field->getv_inf());
	d.scalar("E_inf", s.field->getE_inf());
	d.scalar("H_inf", s.field->getH_inf());
}

// Seeded perturbation of every cell (interior and ghost) about the freestream. Ghost Mach and
// velocity are spread so that the farfield classifier visits all four branches.
void seedState(Setup &s, uint64_t seed) {
	SplitMix64 rng(seed);
	auto &ff = *s.field;
	double g = ff.getgamma_ref();
	int n = s.nTot();
	int nE = s.nElem();
	for (int i = 0; i < n; i++) {
		double rho = 1.0 + 0.4 * (rng.uniform() - 0.5);
		double u = ff.getu_inf() * (1.0 + 0.6 * (rng.uniform() - 0.5));
		double v = ff.getv_inf() + 0.3 * (rng.uniform() - 0.5);
		double w = 0.3 * (rng.uniform() - 0.5);
		double p = 1.0 + 0.4 * (rng.uniform() - 0.5);
		double M = 0.4 + 1.2 * rng.uniform();
		if (i >= nE) {
			// flip the sign of the ghost velocity half of the time (in/outflow branches)
			if (rng.uniform() < 0.5) {
				u = -u;
				v = -v;
				w = -w;
			}
		}
		double E = p / ((g - 1) * rho) + ((u * u + v * v + w * w) / 2.0);
		double H = E + p / rho;
		ff.setRho(i, rho);
		ff.setU(i, u);
		ff.setV(i, v);
		ff.setW(i, w);
		ff.setP(i, p);
		ff.setE(i, E);
		ff.setH(i, H);
		ff.setM(i, M);
	}
}

void modeFlux(Setup &s, Dump &d, uint64_t seed) {
	auto &sv = *s.solver;
	auto &ff = *s.field;
	auto &m = s.mesh;
	auto &mt = *s.metrics;
	seedState(s, seed);
	if (s.mpi.getPoolSize() > 1) {
		// partitioned run: make every interior state one that an update could leave, p and H from
		// (rho, u, v, w, E) as FlowField::Update forms them (FlowField.cpp:121-122), because the
		// exchange recomputes them that way on the receiving side (MPIHandler.cpp:184-196)
		double g = ff.getgamma_ref();
		for (int i = 0; i < s.nElem(); i++) {
			double rho = ff.Getrho()[i], u = ff.GetU_Velocity()[i], v = ff.GetV_Velocity()[i], w = ff.GetW_Velocity()[i];
			double E = ff.GetE()[i];
			double p = (g - 1) * rho * (E - (std::pow(u, 2) + std::pow(v, 2) + std::pow(w, 2)) / 2.0);
			ff.setP(i, p);
			ff.setH(i, E + (p / rho));
		}
	}
	// interface ghosts hold their own cell's copy, as every update leaves them
	// (FlowField.cpp:129-138); the exchange in updateBC then brings the other partition's cells
	const auto &mg = m.GetMPIGhostCellsIDs();
	const auto &mc = m.getMPIadjacentToGhostCellIDs();
	for (size_t i = 0; i < mg.size(); i++) {
		ff.setRho(mg[i], ff.Getrho()[mc[i]]);
		ff.setU(mg[i], ff.GetU_Velocity()[mc[i]]);
		ff.setV(mg[i], ff.GetV_Velocity()[mc[i]]);
		ff.setW(mg[i], ff.GetW_Velocity()[mc[i]]);
		ff.setP(mg[i], ff.GetP()[mc[i]]);
		ff.setE(mg[i], ff.GetE()[mc[i]]);
		ff.setH(mg[i], ff.GetH()[mc[i]]);
		ff.setM(mg[i], ff.GetMach()[mc[i]]);
	}
	dumpState(d, ff, "in_");

	// Same preamble as EulerSolver::Run (EulerSolver.cpp:65-78)
	sv.sortGhostCells();
	sv._residuals.resize(m.GetnElemTot(), Solver::ResidualVar(0.0, 0.0, 0.0, 0.0, 0.0));
	sv._deltaT.resize(m.GetnElemTot(), 0);
	sv._deltaW.resize(m.GetnElemTot(), Solver::ConservativeVar(0.0, 0.0, 0.0, 0.0, 0.0));

	sv.resetResiduals();
	sv.updateBC();
	dumpState(d, ff, "bc_");

	// Per-face flux exactly as computeResidual dispatches it (EulerSolver.cpp:306-358), kept
	// unscaled (F) and scaled (F*S) so both are pinned.
	int nF = m.GetnFace();
	std::vector<double> F(5 * (size_t) nF), FS(5 * (size_t) nF);
	std::vector<int> kind(nF);
	for (int f = 0; f < nF; f++) {
		int *ptr = m.GetFace2ElementID(f);
		int e2 = ptr[1];
		double S = mt.getFaceSurfaces()[f];
		Solver::ResidualVar r;
		if (f < m.GetnFaceInt()) {
			r = Solver::Roe(ff, m, mt, f);
			kind[f] = 0;
		} else if (std::binary_search(sv._sortedWallGhostCellIDs.begin(), sv._sortedWallGhostCellIDs.end(), e2)) {
			auto n = mt.getFaceNormalsUnit()[f];
			r = Solver::ResidualVar(0, n.x * ff.GetP()[e2], n.y * ff.GetP()[e2], n.z * ff.GetP()[e2], 0);
			kind[f] = 1;
		} else if (std::binary_search(sv._sortedMPIGhostCellIDs.begin(), sv._sortedMPIGhostCellIDs.end(), e2)) {
			r = Solver::Roe(ff, m, mt, f);
			kind[f] = 4;
		} else if (std::binary_search(sv._sortedSymmetryGhostCellIDs.begin(), sv._sortedSymmetryGhostCellIDs.end(), e2)) {
			r = Solver::Roe(ff, m, mt, f);
			kind[f] = 2;
		} else {
			auto n = mt.getFaceNormalsUnit()[f];
			double V = ff.GetU_Velocity()[e2] * n.x + ff.GetV_Velocity()[e2] * n.y + ff.GetW_Velocity()[e2] * n.z;
			r = Solver::Fc(ff, mt, e2, f, V);
			kind[f] = 3;
		}
		Solver::ResidualVar rs = r * S;
		double a[5] = {r.m_rhoV_residual, r.m_rho_uV_residual, r.m_rho_vV_residual, r.m_rho_wV_residual, r.m_rho_HV_residual};
		double b[5] = {rs.m_rhoV_residual, rs.m_rho_uV_residual, rs.m_rho_vV_residual, rs.m_rho_wV_residual, rs.m_rho_HV_residual};
		for (int k = 0; k < 5; k++) {
			F[5 * (size_t) f + k] = a[k];
			FS[5 * (size_t) f + k] = b[k];
		}
	}
	d.ints("face_kind", kind);
	d.dbls("face_flux", F);
	d.dbls("face_flux_scaled", FS);

	sv.computeResidual();
	dumpResiduals(d, sv._residuals, "residual");
	sv.updateDeltaTime();
	d.dbls("dt", sv._deltaT);

	// One full RK5 step from here, as Run calls it (EulerSolver.cpp:89-96)
	sv.RungeKutta();
	dumpState(d, ff, "rk_");
	dumpResiduals(d, sv._residuals, "rk_residual");
	d.scalar("rk_rms_sum", sv.computeRMS());
}

void modeRun(Setup &s, Dump &d) {
	auto &sv = *s.solver;
	auto &ff = *s.field;
	sv.Run();
	dumpState(d, ff, "run_");
	dumpResiduals(d, sv._residuals, "run_residual");
	d.scalar("run_rms_sum", sv.computeRMS());
	d.ints("run_iterations", {sv._nbInteration});
	d.scalar("CL", sv._CL);
	d.scalar("CD", sv._CD);
	d.scalar("CM", sv._CM);
	d.dbls("force_coeff", {sv._forceCoefficients.x, sv._forceCoefficients.y, sv._forceCoefficients.z});
	d.dbls("moment_coeff", {sv._momentCoefficients.x, sv._momentCoefficients.y, sv._momentCoefficients.z});
	sv.WriteSummary(0, 0);
}

}// namespace

int main(int argc, char **argv) {
	if (argc < 4) {
		std::fprintf(stderr, "usage: %s mesh|flux|run <cfg.e3d> <out.txt> [seed]\n", argv[0]);
		return 2;
	}
	int fake_argc = 1;
	Parallel::MPIHandler mpi(fake_argc, argv);
	std::string mode = argv[1];
	{
		Setup s(mpi, argv[2]);
		// several ranks: one dump per rank, <out>.r<rank>
		std::string out = argv[3];
		if (mpi.getPoolSize() > 1) out += ".r" + std::to_string(mpi.getRankID());
		Dump d(out.c_str());
		if (mode == "mesh") {
			modeMesh(s, d);
		} else if (mode == "flux") {
			modeFlux(s, d, argc > 4 ? std::stoull(argv[4]) : 0);
		} else if (mode == "run") {
			modeRun(s, d);
		} else {
			std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
			return 2;
		}
	}
	mpi.finalize();
	return 0;
}
