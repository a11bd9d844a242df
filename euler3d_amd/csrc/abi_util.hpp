// Shared pieces of the C ABI implementation: the error channel, HIP error checks and the mesh
// handle type.
#pragma once

#include "../../include/e3d.h"
#include "device.hpp"
#include "e3d_host.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

struct e3d_mesh {
	e3d::MeshFile file;
	e3d::SolverMesh sm;
	bool preprocessed = false;
	// several partitions read together (e3d_mesh_read_partitions): their files and, once
	// preprocessed, where each sits in the composite (file is then their merged copy)
	std::vector<e3d::MeshFile> parts;
	std::vector<e3d::PartRange> ranges;
};

namespace e3d_abi {

// message of the last failure on this thread (e3d_last_error)
std::string &lastError();

inline int fail(int code, const std::string &msg) {
	lastError() = msg;
	return code;
}

struct HipError : std::runtime_error {
	using std::runtime_error::runtime_error;
};

inline void hipCheck(hipError_t e, const char *what) {
	if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template<typename F>
int guarded(F &&f) {
	try {
		return f();
	} catch (const HipError &e) {
		return fail(E3D_ERR_DEVICE, e.what());
	} catch (const std::invalid_argument &e) {
		return fail(E3D_ERR_ARG, e.what());
	} catch (const std::domain_error &e) {
		return fail(E3D_ERR_NUMERIC, e.what());
	} catch (const std::exception &e) {
		return fail(E3D_ERR_MESH, e.what());
	}
}

// nondimensional freestream, FlowField::Initialize (FlowField.cpp:72-81; rho_inf = p_inf = 1)
inline void fillFree(const e3d_params &p, e3d::DevFree &q) {
	const double pi = 3.141592653589793238;
	double g = p.gamma;
	q.gamma = g;
	q.cfl = p.cfl;
	q.rho_inf = 1.0;
	q.p_inf = 1.0;
	q.M_inf = p.mach;
	q.u_inf = q.M_inf * std::sqrt(g) * std::cos(p.aoa_deg * (pi / 180));
	q.v_inf = q.M_inf * std::sqrt(g) * std::sin(p.aoa_deg * (pi / 180));
	q.w_inf = 0.0;
	q.E_inf = q.p_inf / ((g - 1) * q.rho_inf) + ((q.u_inf * q.u_inf + q.v_inf * q.v_inf + q.w_inf * q.w_inf) / 2.0);
	q.H_inf = q.E_inf + (q.p_inf / q.rho_inf);
}

}// namespace e3d_abi
