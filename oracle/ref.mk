# Builds the reference solver from its own sources, where they lie, plus the
# checker harnesses (oracle/ref_harness.cpp: the octree solver;
# oracle/ast_harness.cpp: the uniform-layout AST solver) into
# oracle/_ref/libastgref.so.
# TEST INFRASTRUCTURE ONLY: nothing in the product links or loads this library.
#
#   make -f oracle/ref.mk -j8            (from the repository root)
#
# Flags follow the reference's CMakeLists.txt:20 (-O3 -ffast-math, C++17,
# GLOBALBENCHMARK), except -march: x86-64-v3 instead of native so the library
# also runs on a different host CPU. EIGEN_USE_BLAS is not defined (no system
# BLAS in this image) and no GLEW/GL code is compiled (only the examples need it).

REF      ?= /root/reference
OUT      := oracle/_ref
OBJ      := $(OUT)/obj
CXX      ?= g++
CXXFLAGS := -std=c++17 -O3 -march=x86-64-v3 -ffast-math -fPIC -DGLOBALBENCHMARK \
            -Wno-unused-variable -Wno-deprecated-declarations -w \
            -I$(REF)/include -I$(REF)/3rdparty

REF_SRCS := $(wildcard $(REF)/src/*.cpp) $(wildcard $(REF)/src/viewer/*.cpp) \
            $(wildcard $(REF)/src/jet/*.cpp) $(wildcard $(REF)/src/geometry/*.cpp)
# 2D-only geometry/level-set units are left out: nothing on the 3D octree path
# uses them and one of them references a symbol defined nowhere in the tree.
REF_SRCS := $(filter-out %2.cpp %2set.cpp,$(REF_SRCS))
REF_OBJS := $(patsubst $(REF)/src/%.cpp,$(OBJ)/%.o,$(REF_SRCS))
HARNESS  := $(OBJ)/ref_harness.o $(OBJ)/ast_harness.o

all: $(OUT)/libastgref.so

$(OBJ)/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/ref_harness.o: oracle/ref_harness.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/ast_harness.o: oracle/ast_harness.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libastgref.so: $(REF_OBJS) $(HARNESS)
	$(CXX) -shared -o $@ $^ -lpthread -lstdc++fs

.PHONY: all
