/* hull.h -- polyhedral features of a convex point cloud (host). */
#ifndef B3H_HULL_H
#define B3H_HULL_H
#include <vector>

struct b3hPolyFace
{
	float plane[4];
	std::vector<int> indices;
};
struct b3hPolyhedron
{
	std::vector<float> vertices;    /* xyz triples */
	std::vector<b3hPolyFace> faces; /* merged faces */
	std::vector<float> uniqueEdges; /* xyz triples */
};

/* returns 0 on success, -1 for fewer than 4 distinct or coplanar points */
int b3h_build_polyhedron(const float* pts, int n, int strideFloats, b3hPolyhedron* out);

#endif
