// E3D_Solver-compatible executable: `E3D_Solver <config.e3d>`.
//
// Same command line, configuration file and output files as the reference's solver main
// (src/Solver.cpp:19-73), built on the C ABI of libe3d_hip.so. Every partition the .e3d file lists
// is resident on one GPU and solved as the reference's `mpiexec -np <PARTITION_FILES>` run solves
// it; no MPI launcher is involved. Outputs, per partition r where the reference writes per rank:
//   <outdir>/residuals.dat     include/solver/ResidualsFile.hpp:17-48
//   <outdir>/cp_part_<r>.dat   src/solver/EulerSolver.cpp:282-301
//   <outdir>/summary.txt       src/solver/EulerSolver.cpp:169-198 (TOML; doubles in the shortest
//                              round-trip form, as toml++ prints them)
//   <par_r>.sol, <par_r>.bry   src/solver/WriteSolution.cpp:5-73
#include "../../include/e3d.h"

#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

namespace {

[[noreturn]] void die(const std::string &what) {
	std::cerr << what << ": " << e3d_last_error() << std::endl;
	std::exit(EXIT_FAILURE);
}

std::string tomlDouble(double v) {
	if (std::isnan(v)) return "nan";
	if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
	char buf[64];
	auto res = std::to_chars(buf, buf + sizeof(buf), v);
	std::string s(buf, res.ptr);
	if (s.find_first_of(".eE") == std::string::npos) s += ".0";
	return s;
}

std::string tomlArray(const double v[3]) {
	return "[ " + tomlDouble(v[0]) + ", " + tomlDouble(v[1]) + ", " + tomlDouble(v[2]) + " ]";
}

}// namespace

int main(int argc, char **argv) {
	auto t0 = std::chrono::steady_clock::now();
	if (argc != 2) {
		std::cerr << "Usage : E3D_Solver <configFile.e3d> " << std::endl;
		return EXIT_FAILURE;
	}
	e3d_params prm;
	char meshPath[4096], outDir[4096];
	if (e3d_config_read(argv[1], &prm, meshPath, sizeof(meshPath), outDir, sizeof(outDir)) != E3D_OK) die("configuration");
	std::filesystem::path out(outDir);

	int nPart = 1;
	if (e3d_config_partition(argv[1], 0, nullptr, 0, &nPart) != E3D_OK) die("configuration");
	std::vector<std::string> parPaths(nPart);
	for (int r = 0; r < nPart; r++) {
		char buf[4096];
		if (e3d_config_partition(argv[1], r, buf, sizeof(buf), nullptr) != E3D_OK) die("configuration");
		parPaths[r] = buf;
	}
	e3d_mesh *mesh = nullptr;
	if (nPart == 1) {
		if (e3d_mesh_read(meshPath, 1, &mesh) != E3D_OK) die("mesh");
	} else {
		std::vector<const char *> pp;
		for (auto &x : parPaths) pp.push_back(x.c_str());
		if (e3d_mesh_read_partitions(nPart, pp.data(), &mesh) != E3D_OK) die("mesh");
	}
	if (e3d_mesh_preprocess(mesh, prm.cref) != E3D_OK) die("connectivity/metrics");
	e3d_mesh_info info;
	e3d_mesh_get_info(mesh, &info);
	std::printf("Mesh: %d partition(s), %d cells, %d faces (%d interior), %d ghosts (wall %d, farfield %d, sym %d)\n", nPart,
	            info.n_elem, info.n_face, info.n_face_int, info.n_ghost, info.n_wall, info.n_far, info.n_sym);
	std::vector<e3d_partition_info> parts(nPart);
	for (int r = 0; r < nPart; r++) e3d_mesh_partition_info(mesh, r, &parts[r]);

	e3d_ctx *ctx = nullptr;
	if (e3d_create(mesh, &prm, 0, &ctx) != E3D_OK) die("device context");

	std::vector<double> rows(8 * (size_t) prm.max_iter + 16);
	int nRows = 0, iters = 0;
	double coeffs[3], cf[3], cm[3];
	auto tRun0 = std::chrono::steady_clock::now();
	if (e3d_run(ctx, rows.data(), (int) rows.size() / 8, &nRows, &iters, coeffs, cf, cm) != E3D_OK) die("run");
	auto tRun1 = std::chrono::steady_clock::now();

	{
		std::ofstream f(out / "residuals.dat");
		const int w = 30;
		f << std::setw(w) << "iter" << std::setw(w) << "Max residual" << std::setw(w) << "Cx" << std::setw(w) << "Cy"
		  << std::setw(w) << "Cz" << std::setw(w) << "CMx" << std::setw(w) << "CMy" << std::setw(w) << "CMz" << std::endl;
		for (int r = 0; r < std::min(nRows, (int) rows.size() / 8); r++) {
			const double *x = &rows[8 * (size_t) r];
			f << std::setw(w) << (int) x[0];
			for (int k = 1; k < 8; k++) f << std::scientific << std::setprecision(10) << std::setw(w) << x[k];
			f << std::endl;
			std::printf("iter %6d  rms %.10e  CL %.10e  CD %.10e\n", (int) x[0], x[1],
			            x[2 + prm.orient_cl[0]] * prm.orient_cl[1], x[2 + prm.orient_cd[0]] * prm.orient_cd[1]);
		}
	}

	// primitive state for .sol and the wall cp for cp_part_0.dat
	const int nTot = info.n_elem + info.n_ghost;
	std::vector<double> fld[8];
	double *ptr[8];
	for (int k = 0; k < 8; k++) {
		fld[k].resize(nTot);
		ptr[k] = fld[k].data();
	}
	if (e3d_get_state(ctx, ptr) != E3D_OK) die("state read-back");

	std::vector<int> wallCell(info.n_wall), wallFace(info.n_wall), f2nStart(info.n_face + 1), f2n;
	long long n = 0;
	e3d_mesh_get_i32(mesh, "wallCell", wallCell.data(), info.n_wall, &n);
	e3d_mesh_get_i32(mesh, "wallFace", wallFace.data(), info.n_wall, &n);
	e3d_mesh_get_i32(mesh, "face2nodeStart", f2nStart.data(), info.n_face + 1, &n);
	e3d_mesh_get_i32(mesh, "face2node", nullptr, 0, &n);
	f2n.resize(n);
	e3d_mesh_get_i32(mesh, "face2node", f2n.data(), n, &n);
	std::vector<double> fc(3 * (size_t) info.n_face), fn(3 * (size_t) info.n_face);
	e3d_mesh_get_f64(mesh, "faceCenters", fc.data(), (long long) fc.size(), &n);
	e3d_mesh_get_f64(mesh, "faceNormals", fn.data(), (long long) fn.size(), &n);

	// cp of the last coefficient update (AeroCoefficients::Update at the last sampled iteration, or
	// zeros when none happened), as PrintCp writes it (EulerSolver.cpp:282-300)
	std::vector<double> wallCp(info.n_wall);
	if (e3d_wall_cp(ctx, wallCp.data(), info.n_wall, &n) != E3D_OK) die("wall cp read-back");
	double wsum[3] = {0, 0, 0}; // Wall_Face_Vector_Sum: per partition, then summed in rank order
	for (int r = 0; r < nPart; r++) {
		const e3d_partition_info &P = parts[r];
		{
			std::ofstream f(out / ("cp_part_" + std::to_string(r) + ".dat"));
			f << std::setw(20) << "Cp" << std::setw(20) << "x" << std::setw(20) << "y" << std::setw(20) << "z" << std::endl;
			for (int i = P.wall0; i < P.wall0 + P.n_wall; i++) {
				double cp = wallCp[i];
				const double *c = &fc[3 * (size_t) wallFace[i]];
				f << std::scientific << std::setprecision(10) << std::setw(20) << cp << std::setw(20) << c[0] << std::setw(20)
				  << c[1] << std::setw(20) << c[2] << std::endl;
			}
		}
		{
			std::ofstream f(parPaths[r] + ".sol");
			f << std::setiosflags(std::ios::dec) << std::setiosflags(std::ios::scientific);
			f << std::setw(20) << "density" << std::setw(20) << "velocity_x" << std::setw(20) << "velocity_y" << std::setw(20)
			  << "velocity_z" << std::setw(20) << "pressure" << std::setw(20) << "energy" << std::endl;
			for (int i = P.elem0; i < P.elem0 + P.n_elem; i++)
				f << std::setw(20) << fld[0][i] << std::setw(20) << fld[1][i] << std::setw(20) << fld[2][i] << std::setw(20)
				  << fld[3][i] << std::setw(20) << fld[4][i] << std::setw(20) << fld[5][i] << std::endl;
		}
		{
			// the partition's own numbering: local cell ids and local node ids
			std::ofstream f(parPaths[r] + ".bry");
			f << "NOMBRE DE CELLULES= " << P.n_wall << std::endl;
			for (int i = P.wall0; i < P.wall0 + P.n_wall; i++) {
				f << std::setw(9) << wallCell[i] - P.elem0;
				int a = f2nStart[wallFace[i]], b = f2nStart[wallFace[i] + 1];
				f << std::setw(9) << (b - a);
				for (int j = a; j < b; j++) f << std::setw(9) << f2n[j] - P.node0;
				f << std::endl;
			}
		}
		double ps[3] = {0, 0, 0};
		for (int i = P.wall0; i < P.wall0 + P.n_wall; i++)
			for (int k = 0; k < 3; k++) ps[k] += fn[3 * (size_t) wallFace[i] + k];
		for (int k = 0; k < 3; k++) wsum[k] = r ? wsum[k] + ps[k] : ps[k];
	}
	auto t1 = std::chrono::steady_clock::now();
	long loopMs = (long) std::chrono::duration_cast<std::chrono::milliseconds>(tRun1 - tRun0).count();
	long totalMs = (long) std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
	{
		double perCell = iters > 0 ? (double) (loopMs * nPart) / (double) ((long long) iters * info.n_elem) : 0.0;
		std::ofstream f(out / "summary.txt");
		f << "Iterations = " << iters << "\n";
		f << "Ncores = " << nPart << "\n";
		f << "SolverLoopTime_ms = " << loopMs << "\n";
		f << "TimeCore_p_IterNElem_ms = " << tomlDouble(perCell) << "\n";
		f << "TotalSolverTime_ms = " << totalMs << "\n\n";
		f << "[Coefficients]\n";
		f << "CD = " << tomlDouble(coeffs[1]) << "\n";
		f << "CL = " << tomlDouble(coeffs[0]) << "\n";
		f << "CM = " << tomlDouble(coeffs[2]) << "\n";
		f << "Force_Coeff = " << tomlArray(cf) << "\n";
		f << "Moment_Coeff = " << tomlArray(cm) << "\n\n";
		f << "[Mesh]\n";
		f << "NElem = " << info.n_elem << "\n";
		f << "Wall_Face_Vector_Sum = " << tomlArray(wsum) << "\n";
	}
	std::cout << (iters >= prm.max_iter ? "Solution not converged :  exceeded maximum number of iterations !"
	                                    : "Solution converged OR a NAN encountred !")
	          << "\nSummary printed at: " << (out / "summary.txt") << std::endl;
	e3d_destroy(ctx);
	e3d_mesh_free(mesh);
	return 0;
}
