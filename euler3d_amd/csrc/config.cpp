// .e3d solver configuration reader.
//
// Follows src/parser/SimConfig.cpp:49-236: the keys are recognised by substring match in a fixed
// order between the START and END lines, '#' lines are comments, and each value is read at a
// fixed offset (the key length plus one) from the start of the line, exactly as the reference's
// seekg() does. PARTITION_FILES takes the pattern on the following line and replaces '#' with the
// partition index (SimConfig.cpp:66-86). Mach/velocity conversion follows SimConfig.cpp:176-182,
// the CL/CD/CM axis table SimConfig.cpp:184-231, the output directory SimConfig.cpp:233-235.
//
// Where the reference leaves a member uninitialised when its key is absent (SREF, CREF, SAMPLING,
// RESIDUAL_SMOOTHING, MIN_AEROCOEFF_ERROR: include/parser/SimConfig.hpp:149-179) this reader uses
// the explicit defaults of e3d::Params. The checks of SimConfig::configFileVerification
// (SimConfig.cpp:275-438) raise std::invalid_argument instead of calling exit().
#include "e3d_host.hpp"

#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace e3d {

namespace {

bool has(const std::string &line, const char *key) { return line.find(key) != std::string::npos; }

// value at a fixed offset, like std::stringstream::seekg(off) >> value
template<typename T>
void readAt(const std::string &line, size_t off, T &out) {
	if (line.size() < off) return;
	std::istringstream ss(line.substr(off));
	T v;
	if (ss >> v) out = v;
}

void axisOf(int code, int out[2], const char *what) {
	switch (code) {
		case 0: out[0] = 0; out[1] = 1; break;
		case 1: out[0] = 0; out[1] = -1; break;
		case 2: out[0] = 1; out[1] = 1; break;
		case 3: out[0] = 1; out[1] = -1; break;
		case 4: out[0] = 2; out[1] = 1; break;
		case 5: out[0] = 2; out[1] = -1; break;
		default: throw std::invalid_argument(std::string("invalid ") + what + " argument");
	}
}

}// namespace

Params readConfig(const std::string &path) {
	std::ifstream in(path);
	if (!in) throw std::invalid_argument("Error while opening configuration file ! " + path);
	Params p;
	int orientCL = -1, orientCD = -1;
	int speedChoice = 0, schemeChoice = 0;
	double speedValue = 0.0;
	bool haveOutput = false;
	std::string line;
	while (std::getline(in, line)) {
		if (!has(line, "START")) continue;
		while (!has(line, "END")) {
			if (!std::getline(in, line)) break;
			if (!line.empty() && line[0] == '#') continue;
			if (has(line, "INITIAL_MESH=")) readAt(line, 14, p.initialMesh);
			else if (has(line, "PARTITION_FILES=")) {
				readAt(line, 16, p.nPartition);
				std::string pat;
				if (std::getline(in, line)) {
					std::istringstream ss(line);
					ss >> pat;
				}
				p.partitionFiles.clear();
				for (int i = 0; i < p.nPartition; i++) {
					std::string s = pat;
					size_t pos;
					while ((pos = s.find('#')) != std::string::npos) s.replace(pos, 1, std::to_string(i));
					p.partitionFiles.push_back(s);
				}
			} else if (has(line, "MESH_ORIENTATION_CL=")) readAt(line, 21, orientCL);
			else if (has(line, "MESH_ORIENTATION_CD=")) readAt(line, 21, orientCD);
			else if (has(line, "MESH_REF_POINT_X=")) readAt(line, 18, p.refPoint[0]);
			else if (has(line, "MESH_REF_POINT_Y=")) readAt(line, 18, p.refPoint[1]);
			else if (has(line, "MESH_REF_POINT_Z=")) readAt(line, 18, p.refPoint[2]);
			else if (has(line, "SREF=")) readAt(line, 6, p.sref);
			else if (has(line, "CREF=")) readAt(line, 6, p.cref);
			else if (has(line, "SAMPLING=")) readAt(line, 10, p.sampling);
			else if (has(line, "PRE_LOG=")) readAt(line, 9, p.preLog);
			else if (has(line, "SPEED_OPTION")) readAt(line, 13, speedChoice);
			else if (has(line, "SPEED_VALUE")) readAt(line, 12, speedValue);
			else if (has(line, "AOA")) readAt(line, 4, p.aoa);
			else if (has(line, "AIRFLOW_PRESSURE")) readAt(line, 17, p.pressure);
			else if (has(line, "AIRFLOW_TEMPERATURE")) readAt(line, 20, p.temperature);
			else if (has(line, "AIRFLOW_VISCOSITY")) readAt(line, 18, p.viscosity);
			else if (has(line, "AIRFLOW_DENSITY")) readAt(line, 16, p.density);
			else if (has(line, "GAMMA")) readAt(line, 6, p.gamma);
			else if (has(line, "GAS_CONSTANT")) readAt(line, 13, p.gasConstant);
			else if (has(line, "SPECIFIC_HEAT")) readAt(line, 14, p.specificHeat);
			else if (has(line, "SCHEME")) readAt(line, 8, schemeChoice);
			else if (has(line, "TIME_INTEGRATION")) {
				int t = -1;
				readAt(line, 17, t);
				if (t < 0 || t > 2) throw std::invalid_argument("invalid TIME_INTEGRATION argument");
				p.timeIntegration = t;
			} else if (has(line, "RESIDUAL_SMOOTHING")) {
				int r = 0;
				readAt(line, 19, r);
				p.residualSmoothing = r != 0;
			} else if (has(line, "CFL")) readAt(line, 4, p.cfl);
			else if (has(line, "MIN_RESIDUAL")) readAt(line, 13, p.minResidual);
			else if (has(line, "MIN_AEROCOEFF_ERROR")) readAt(line, 20, p.minAeroCoeffError);
			else if (has(line, "MAX_ITER")) readAt(line, 9, p.maxIter);
			else if (has(line, "SOLVER_LOG=")) readAt(line, 11, p.solverLog);
			else if (has(line, "OUTPUT_FILE")) {
				readAt(line, 13, p.outputFile);
				haveOutput = !p.outputFile.empty();
			} else if (has(line, "POST_LOG=")) readAt(line, 9, p.postLog);
		}
		break;
	}
	p.speedOption = speedChoice;
	p.scheme = schemeChoice;
	if (speedChoice == 0) {
		p.mach = speedValue;
		p.velocity = std::sqrt(p.gamma * p.gasConstant * p.temperature) * p.mach;
	} else {
		p.velocity = speedValue;
		p.mach = p.velocity / (std::sqrt(p.gamma * p.gasConstant * p.temperature));
	}
	axisOf(orientCL, p.orientCL, "MESH_ORIENTATION_CL");
	axisOf(orientCD, p.orientCD, "MESH_ORIENTATION_CD");
	p.orientCM[0] = 3 - p.orientCL[0] - p.orientCD[0];
	p.orientCM[1] = -p.orientCL[1] * p.orientCD[1];
	p.outputDir = std::filesystem::path(p.outputFile).parent_path().string();

	// SimConfig::configFileVerification (SimConfig.cpp:275-438)
	if (p.sampling <= 0) throw std::invalid_argument("The Sampling Period must be a positive integer!");
	if (speedChoice != 0 && speedChoice != 1) throw std::invalid_argument("The Speed Option is incorrect!");
	if (p.mach < 0. || p.mach > 2.5) throw std::invalid_argument("The Mach Number is either too low (below 0) or too high (above 2.5)!");
	if (p.pressure < 0 || p.temperature < 0 || p.viscosity < 0 || p.density < 0 || p.gamma < 0 || p.gasConstant < 0 ||
	    p.specificHeat < 0)
		throw std::invalid_argument("negative airflow property in configuration");
	if (schemeChoice != 0 && schemeChoice != 1) throw std::invalid_argument("The Scheme Option is incorrect! It must be either 0 (Roe) or 1 (Ausm).");
	if (p.cfl <= 0) throw std::invalid_argument("The CFL number must be greater than 0!");
	if (p.minResidual < 0.) throw std::invalid_argument("The Minimum Residual is too low (<0)!");
	if (p.maxIter <= 0) throw std::invalid_argument("The Maximum Number of Iterations must be greater than 0!");
	if (!haveOutput) throw std::invalid_argument("No output file!");
	return p;
}

}// namespace e3d
