# Builds the reference Euler3D solver (E3D_Solver), its Tecplot post-processor (E3D_POST) and the
# golden-vector harness from the reference sources where they lie, into oracle/_ref/. TEST INFRASTRUCTURE ONLY: this runs only
# where the reference tree is present; the product never links or loads what it builds.
#
# Flags: -O2 -ffp-contract=off (IEEE-strict). The reference's own Release build uses -Ofast
# (CMakeLists.txt:21), which is not bit-reproducible; parity is defined against IEEE arithmetic.
# MPI comes from the MPICH that ships in /opt/conda. Its libstdc++ is older than the system one,
# so the link goes through a directory that holds only the MPI runtime libraries.
#
#   make -f oracle/ref.mk            (from the repository root)

REF      ?= /root/reference
CONDA    ?= /opt/conda
OUT      := oracle/_ref
MPILIB   := $(OUT)/mpilib
CXX      := g++
CXXFLAGS := -std=c++17 -O2 -ffp-contract=off -w -I$(REF)/include -I$(REF)/extern -I$(CONDA)/include
LDFLAGS  := -L$(MPILIB) -Wl,-rpath,$(abspath $(MPILIB)) -lmpicxx -lmpi

SOLVER_SRCS := $(wildcard $(REF)/src/solver/*.cpp) $(wildcard $(REF)/src/mesh/*.cpp) \
               $(REF)/src/parser/SU2MeshParser.cpp $(REF)/src/parser/MeshPartition.cpp \
               $(REF)/src/parser/SimConfig.cpp $(REF)/src/connectivity/Connectivity.cpp \
               $(REF)/src/parallelization/MPIHandler.cpp
OBJS := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(SOLVER_SRCS))

POST_SRCS := $(REF)/src/post/post.cpp $(REF)/src/parser/SolutionPost.cpp $(REF)/src/utils/Logger.cpp
POST_OBJS := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(POST_SRCS)) $(OUT)/obj/parser/SU2MeshParser.o \
             $(OUT)/obj/parser/SimConfig.o

all: $(OUT)/E3D_Solver $(OUT)/ref_harness $(OUT)/E3D_POST

$(MPILIB)/.stamp:
	mkdir -p $(MPILIB)
	for f in libmpi.so libmpi.so.12 libmpicxx.so libmpicxx.so.12 libgfortran.so.4 libquadmath.so.0; do \
	  ln -sf $(CONDA)/lib/$$f $(MPILIB)/$$f; done
	touch $@

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/Solver_main.o: $(REF)/src/Solver.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/Post_main.o: $(REF)/src/postprocessor.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/E3D_POST: $(POST_OBJS) $(OUT)/obj/Post_main.o $(MPILIB)/.stamp
	$(CXX) -o $@ $(POST_OBJS) $(OUT)/obj/Post_main.o $(LDFLAGS) -lpthread

$(OUT)/obj/ref_harness.o: oracle/ref_harness.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/E3D_Solver: $(OBJS) $(OUT)/obj/Solver_main.o $(MPILIB)/.stamp
	$(CXX) -o $@ $(OBJS) $(OUT)/obj/Solver_main.o $(LDFLAGS)

$(OUT)/ref_harness: $(OBJS) $(OUT)/obj/ref_harness.o $(MPILIB)/.stamp
	$(CXX) -o $@ $(OBJS) $(OUT)/obj/ref_harness.o $(LDFLAGS)

.PHONY: all
