// E3D_PRE-compatible executable: `E3D_PRE <config.e3d>`.
//
// Same command line, configuration file and partition files as the reference's pre-processor
// (src/preprocessor.cpp:16-45, Partition::Write, src/partition/Partition.cpp:357-460): reads
// INITIAL_MESH, solves its connectivity and writes the PARTITION_FILES with the SU2Writer layout at
// its default 10 digits. The reference assigns cells to partitions with METIS, which is not part of
// its sources; this one uses recursive coordinate bisection (a single partition needs neither).
#include "../../include/e3d.h"

#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
	if (argc != 2) {
		std::cerr << "Usage : E3D_PRE <configFile.e3d> " << std::endl;
		return EXIT_FAILURE;
	}
	auto die = [](const std::string &what) {
		std::cerr << what << ": " << e3d_last_error() << std::endl;
		std::exit(EXIT_FAILURE);
	};
	char initial[4096];
	int nPart = 0;
	if (e3d_config_initial_mesh(argv[1], initial, sizeof(initial)) != E3D_OK) die("configuration");
	if (e3d_config_partition(argv[1], 0, nullptr, 0, &nPart) != E3D_OK) die("configuration");
	std::vector<std::string> paths(nPart);
	std::vector<const char *> pp(nPart);
	for (int r = 0; r < nPart; r++) {
		char buf[4096];
		if (e3d_config_partition(argv[1], r, buf, sizeof(buf), nullptr) != E3D_OK) die("configuration");
		paths[r] = buf;
		pp[r] = paths[r].c_str();
	}
	std::cout << "Euler 3D Pre-processor.\n" << std::endl;
	e3d_mesh *mesh = nullptr;
	if (e3d_mesh_read(initial, 0, &mesh) != E3D_OK) die("mesh");
	if (e3d_mesh_preprocess(mesh, 1.0) != E3D_OK) die("connectivity");
	std::cout << std::string(24, '#') << "  Partitionning  " << std::string(24, '#') << "\n" << std::endl;
	std::cout << "      Number of Partitions : " << nPart << "\n";
	if (e3d_mesh_partition_write(mesh, nPart, nullptr, pp.data(), 10) != E3D_OK) die("partitioning");
	e3d_mesh_free(mesh);
	return 0;
}
