# Builds oracle/_ref/libb3ref.so from the reference's own sources where they
# lie under /root/reference (nothing is copied): Bullet3Common, the hull
# computer, b3ConvexUtility and the dual-compile shared headers, driven by
# oracle/ref_harness.cpp.  Test infrastructure only.
REF ?= /root/reference/src
CXX ?= g++
CXXFLAGS = -O2 -fPIC -w -I$(REF)
SRCS = ref_harness.cpp \
	$(REF)/Bullet3Common/b3AlignedAllocator.cpp \
	$(REF)/Bullet3Common/b3Logging.cpp \
	$(REF)/Bullet3Common/b3Vector3.cpp \
	$(REF)/Bullet3Geometry/b3ConvexHullComputer.cpp \
	$(REF)/Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.cpp

_ref/libb3ref.so: $(SRCS)
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -shared -o $@ $(SRCS)
