# Builds oracle/_ref/libb3ref.so from the reference's own sources where they
# lie under /root/reference (nothing is copied): Bullet3Common, the hull
# computer, b3ConvexUtility, b3QuantizedBvh and the dual-compile shared headers, driven by
# oracle/ref_harness.cpp.  Test infrastructure only.
REF ?= /root/reference/src
CXX ?= g++
CXXFLAGS = -O2 -fPIC -w -I$(REF)
SRCS = ref_harness.cpp \
	$(REF)/Bullet3Common/b3AlignedAllocator.cpp \
	$(REF)/Bullet3Common/b3Logging.cpp \
	$(REF)/Bullet3Common/b3Vector3.cpp \
	$(REF)/Bullet3Geometry/b3ConvexHullComputer.cpp \
	$(REF)/Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.cpp \
	$(REF)/Bullet3OpenCL/NarrowphaseCollision/b3QuantizedBvh.cpp

all: _ref/libb3ref.so _ref/libb3refsolver.so

_ref/libb3ref.so: $(SRCS)
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -shared -o $@ $(SRCS)

# the solver's host emulation (b3Solver.cpp) with its OpenCL-side dependencies,
# all from the reference's own tree; OpenCL is reached only through the
# vendored clew loader, which nothing here initialises
SOLVER_SRCS = ref_solver_harness.cpp \
	$(REF)/Bullet3Common/b3AlignedAllocator.cpp \
	$(REF)/Bullet3Common/b3Logging.cpp \
	$(REF)/Bullet3Common/b3Vector3.cpp \
	$(REF)/Bullet3OpenCL/ParallelPrimitives/b3LauncherCL.cpp \
	$(REF)/Bullet3OpenCL/ParallelPrimitives/b3PrefixScanCL.cpp \
	$(REF)/Bullet3OpenCL/ParallelPrimitives/b3BoundSearchCL.cpp \
	$(REF)/Bullet3OpenCL/ParallelPrimitives/b3RadixSort32CL.cpp \
	$(REF)/Bullet3OpenCL/ParallelPrimitives/b3FillCL.cpp \
	$(REF)/Bullet3OpenCL/Initialize/b3OpenCLUtils.cpp

_ref/clew.o: $(REF)/clew/clew.c
	mkdir -p _ref
	gcc -O2 -fPIC -w -c -o $@ $<

_ref/libb3refsolver.so: $(SOLVER_SRCS) _ref/clew.o
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -DB3_USE_CLEW -I$(REF)/clew -shared -o $@ $(SOLVER_SRCS) _ref/clew.o -ldl
