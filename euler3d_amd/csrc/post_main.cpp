// E3D_POST-compatible executable: `E3D_POST <config.e3d>`.
//
// Same command line, configuration file and Tecplot output as the reference's post-processor
// (src/postprocessor.cpp:17-40, src/post/post.cpp:38-47), built on e3d_post_tecplot of
// libe3d_hip.so. Reads the partition meshes the .e3d file lists and the .sol file of each from the
// directory of OUTPUT_FILE; no GPU is used.
#include "../../include/e3d.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

int main(int argc, char **argv) {
	if (argc != 2) {
		std::cerr << "Usage : E3D_POST <configFile.e3d> " << std::endl;
		return EXIT_FAILURE;
	}
	std::cout << "Euler 3D Post-processor.\n" << std::endl;
	std::cout << std::string(24, '#') << "  TECPLOT  " << std::string(24, '#') << "\n" << std::endl;
	if (e3d_post_tecplot(argv[1]) != E3D_OK) {
		std::cerr << e3d_last_error() << std::endl;
		return EXIT_FAILURE;
	}
	return 0;
}
