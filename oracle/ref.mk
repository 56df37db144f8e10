# TEST INFRASTRUCTURE: compiles the reference's own CPU translation units, read
# in place from the reference tree, together with oracle/ref_harness.cpp into
# oracle/_ref/libeosref.so.  Nothing from the reference is copied into this
# repository; the output directory is git-ignored.
#
# Only the renderer's units are compiled (RayTracer, Scene, BVH, Triangle, Box,
# Material, RayIntersection, LightSource).  triangleMesh.cpp is not: it needs
# libassimp, which the reference does not ship; its one caller,
# Scene::addTriangleMesh, is unused by the harness and is dropped by
# --gc-sections (reference units are compiled with hidden visibility so the
# linker may discard what the harness does not reach).
#
# The any-hit path reads an uninitialised float (src/BVH.cpp:697-700, the
# shadow ray's `distance`), so its result is undefined.  The units are built by
# clang with -ftrivial-auto-var-init=pattern, which fills that float with a NaN
# pattern: `t > NaN` is false, every t >= 0 hit occludes, and the build
# realises the "any hit at t >= 0" reading deterministically.  (A plain g++ -O2
# build of the same source drops most occluders: 18 of 254 on the icosphere
# scene.)
#
#   make -f ref.mk REF=/root/reference            parity build (no fast-math)
#   make -f ref.mk REF=/root/reference timing     reference flags (-O3 -ffast-math)
REF     ?= /root/reference
SRC     := $(REF)/src
OUT     := _ref
CXX     := /opt/rocm/llvm/bin/clang++
INC     := -I$(SRC) -I$(SRC)/glm/glm -I$(SRC)/include
COMMON  := -std=c++11 -fPIC -fopenmp -w -ffunction-sections -fdata-sections -fvisibility=hidden \
           -ftrivial-auto-var-init=pattern
PARITY  := $(COMMON) -O2 -ffp-contract=off -fno-fast-math
TIMING  := $(COMMON) -O3 -ffast-math
UNITS   := RayTracer.cpp Scene.cpp BVH.cpp Triangle.cpp Material.cpp RayIntersection.cpp LightSource.cpp
LDFLAGS := -shared -fopenmp -Wl,--gc-sections -Wl,-z,defs -Wl,-rpath,/opt/rocm/llvm/lib

all: $(OUT)/libeosref.so
timing: $(OUT)/libeosref_fast.so

$(OUT)/libeosref.so: ref_harness.cpp
	mkdir -p $(OUT)
	$(CXX) $(PARITY) $(INC) $(addprefix $(SRC)/,$(UNITS)) -x c++ $(SRC)/Box.cu -x none ref_harness.cpp $(LDFLAGS) -o $@

$(OUT)/libeosref_fast.so: ref_harness.cpp
	mkdir -p $(OUT)
	$(CXX) $(TIMING) $(INC) $(addprefix $(SRC)/,$(UNITS)) -x c++ $(SRC)/Box.cu -x none ref_harness.cpp $(LDFLAGS) -o $@

.PHONY: all timing
