/*
 * b3h_api.cpp -- host side of the C ABI: the narrowphase shape/body registry,
 * the two broadphases and the pipeline that chains the device stages.
 *
 * Mirrors the reference classes' host logic:
 *   b3GpuNarrowPhase       src/Bullet3OpenCL/RigidBody/b3GpuNarrowPhase.cpp
 *   b3GpuGridBroadphase    src/Bullet3OpenCL/BroadphaseCollision/b3GpuGridBroadphase.cpp
 *   b3GpuSapBroadphase     src/Bullet3OpenCL/BroadphaseCollision/b3GpuSapBroadphase.cpp
 *   b3GpuRigidBodyPipeline src/Bullet3OpenCL/RigidBody/b3GpuRigidBodyPipeline.cpp
 * Device buffers are allocated once at the configured capacities (as the
 * reference's b3OpenCLArrays are), so a step allocates nothing and can be
 * captured into a HIP graph.  A step never waits on the host: counts produced
 * by one stage (pairs, contacts) stay on the device and size the next stage.
 */
#include "b3h_internal.h"
#include "hull.h"
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cmath>

static thread_local char g_err[512] = "";

void b3h_set_error(const char* fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
}

extern "C" const char* b3h_last_error(void) { return g_err; }
extern "C" const char* b3h_version(void) { return "b3hip 0.1 (gfx950)"; }

#define CHECK_ARG(c)                                   \
	do                                                 \
	{                                                  \
		if (!(c))                                      \
		{                                              \
			b3h_set_error("invalid argument: %s", #c); \
			return B3H_ERR_INVALID;                    \
		}                                              \
	} while (0)

template <typename T>
static int dalloc(T** p, size_t n)
{
	*p = nullptr;
	if (n == 0) n = 1;
	hipError_t e = hipMalloc((void**)p, n * sizeof(T));
	if (e != hipSuccess)
	{
		b3h_set_error("hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
		*p = nullptr;
		return B3H_ERR_HIP;
	}
	return B3H_OK;
}
template <typename T>
static void dfree(T*& p)
{
	if (p) hipFree((void*)p);
	p = nullptr;
}

static int ensure_device()
{
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n <= 0)
	{
		b3h_set_error("no HIP device available (hipGetDeviceCount: %s, count %d)", hipGetErrorString(e), n);
		return B3H_ERR_NO_DEVICE;
	}
	return B3H_OK;
}

static int make_stream(void* user, hipStream_t* s, bool* own)
{
	if (user)
	{
		*s = (hipStream_t)user;
		*own = false;
		return B3H_OK;
	}
	B3H_CHECK(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
	*own = true;
	return B3H_OK;
}

extern "C" int b3h_config_default(b3hConfig* c)
{
	CHECK_ARG(c);
	c->m_maxConvexBodies = 128 * 1024;
	c->m_maxVerticesPerFace = 64;
	c->m_maxFacesPerShape = 12;
	c->m_maxConvexVertices = 8192;
	c->m_maxConvexIndices = 81920;
	c->m_maxConvexUniqueEdges = 8192;
	c->m_maxCompoundChildShapes = 8192;
	c->m_maxTriConvexPairCapacity = 256 * 1024;
	c->m_maxConvexShapes = c->m_maxConvexBodies;
	c->m_maxBroadphasePairs = 16 * c->m_maxConvexBodies;
	c->m_maxContactCapacity = c->m_maxBroadphasePairs;
	c->m_compoundPairCapacity = 1024 * 1024;
	return B3H_OK;
}

extern "C" int b3h_solver_params_default(b3hSolverParams* p)
{
	CHECK_ARG(p);
	p->m_solverDt = 1.f / 60.f;
	p->m_numIterations = 4;
	p->m_frictionCoeff = 0.7f;
	p->m_positionDrift = 0.005f;
	p->m_positionConstraintCoeff = 0.2f;
	p->m_batchCellSize = 6.f;
	p->m_angularDamping = 0.99f;
	p->m_warmStartMode = 0;
	return B3H_OK;
}

/* unit-sphere directions live in constant memory of the narrowphase module */
int b3h_np_module_init();

/* ======================================================================= */
struct b3hNarrowPhase
{
	b3hConfig cfg;
	hipStream_t stream;
	bool ownStream;
	std::vector<b3hCollidable> cols;
	std::vector<b3hConvexPolyhedronData> convexes;
	std::vector<b3hFloat4> verts, edges;
	std::vector<b3hGpuFace> faces;
	std::vector<int> indices;
	std::vector<b3hAabb> localAabbs;
	std::vector<b3hGpuChildShape> children;
	std::vector<b3hBvhInfo> bvhInfos;
	std::vector<b3hQuantizedBvhNode> bvhNodes;
	std::vector<b3hBvhSubtreeInfo> bvhSubtrees;
	std::vector<b3hRigidBodyData> bodies;
	std::vector<b3hInertiaData> inertias;
	int static0Index;
	/* device */
	b3hRigidBodyData* dBodies;
	b3hInertiaData* dInertias;
	b3hCollidable* dCols;
	b3hConvexPolyhedronData* dConvex;
	b3hFloat4 *dVerts, *dEdges;
	b3hGpuFace* dFaces;
	int* dIndices;
	b3hAabb* dLocalAabbs;
	b3hGpuChildShape* dChildren;
	b3hBvhInfo* dBvhInfos;
	b3hQuantizedBvhNode* dBvhNodes;
	b3hBvhSubtreeInfo* dBvhSubtrees;
	size_t capCols, capConvex, capVerts, capEdges, capFaces, capIndices, capLocal;
	size_t capChildren, capBvhInfos, capBvhNodes, capBvhSubtrees;
	b3hContact4Data* dStaging;
	int *dFlags, *dFlagScan, *dContactPair, *dNumContacts, *dNumPairs;
	unsigned* dScanTmp;
	b3hContact4Data* dContacts;
	int* dSepState;
	b3hFloat4* dSepNormal;
	b3hFloat4* dMpr; /* per pair: MPR direction, point, (depth, hit) */
	/* compound child pairs (allocated once a compound is registered) */
	b3hCompoundWork cw;
	/* per-pair buffers between the SAT passes */
	float* dSatDmin;
	int *dSatFlag, *dSatScan, *dSatList, *dSatCount;
	/* per-body face records for the SAT (b3hSatCache): record ranges follow
	 * the host mirror's collidable indices and are rebuilt when those change */
	b3hSatFace* dSatRecs;
	int *dSatBase, *dSatOwner;
	size_t satCapRecs, satCapBase;
	int satBodies, satTotal;
	float2* dSatDirs;
	int *dSatDslot, *dSatDowner;
	b3hSatBody* dSatDesc;
	size_t satCapDirBodies;
	int satDirBodies;
	unsigned satGen;
	bool satValid;
	std::vector<int> satCols; /* collidable index per body when the ranges were built */
	size_t satNumConvex;
	bool contactsValid;
	int lanesPerPair;
	/* bumped by every call that can change what a captured step refers to
	 * (tables, their device copies, bodies, launch choices) */
	unsigned gen;
};

static int np_alloc_work(b3hNarrowPhase* np)
{
	int mp = np->cfg.m_maxBroadphasePairs, mc = np->cfg.m_maxContactCapacity;
	int r;
	if ((r = dalloc(&np->dBodies, np->cfg.m_maxConvexBodies))) return r;
	if ((r = dalloc(&np->dInertias, np->cfg.m_maxConvexBodies))) return r;
	if ((r = dalloc(&np->dStaging, mp))) return r;
	if ((r = dalloc(&np->dFlags, mp + 1))) return r;
	if ((r = dalloc(&np->dFlagScan, mp + 2))) return r;
	if ((r = dalloc(&np->dScanTmp, b3h_scan_tmp_size(mp)))) return r;
	if ((r = dalloc(&np->dContactPair, mc))) return r;
	if ((r = dalloc(&np->dNumContacts, 2))) return r;
	if ((r = dalloc(&np->dNumPairs, 1))) return r;
	if ((r = dalloc(&np->dContacts, mc))) return r;
	if ((r = dalloc(&np->dSepState, mp))) return r;
	if ((r = dalloc(&np->dSepNormal, mp))) return r;
	if ((r = dalloc(&np->dMpr, 3 * (size_t)mp))) return r;
	if ((r = dalloc(&np->dSatDmin, mp)) || (r = dalloc(&np->dSatFlag, mp + 1)) || (r = dalloc(&np->dSatScan, mp + 2)) ||
		(r = dalloc(&np->dSatList, mp)) || (r = dalloc(&np->dSatCount, 1)))
		return r;
	B3H_CHECK(hipMemset(np->dNumContacts, 0, 2 * sizeof(int)));
	return B3H_OK;
}

extern "C" int b3h_np_create(const b3hConfig* cfg, void* stream, b3hNarrowPhase** out)
{
	CHECK_ARG(cfg && out);
	*out = nullptr;
	int r = ensure_device();
	if (r) return r;
	CHECK_ARG(cfg->m_maxConvexBodies > 0 && cfg->m_maxBroadphasePairs > 0 && cfg->m_maxContactCapacity > 0);
	b3hNarrowPhase* np = new b3hNarrowPhase();
	memset((void*)&np->cfg, 0, sizeof(np->cfg));
	np->cfg = *cfg;
	np->static0Index = -1;
	np->dBodies = nullptr;
	np->dInertias = nullptr;
	np->dCols = nullptr;
	np->dConvex = nullptr;
	np->dVerts = np->dEdges = nullptr;
	np->dFaces = nullptr;
	np->dIndices = nullptr;
	np->dLocalAabbs = nullptr;
	np->dChildren = nullptr;
	np->dBvhInfos = nullptr;
	np->dBvhNodes = nullptr;
	np->dBvhSubtrees = nullptr;
	np->capCols = np->capConvex = np->capVerts = np->capEdges = np->capFaces = np->capIndices = np->capLocal = 0;
	np->capChildren = np->capBvhInfos = np->capBvhNodes = np->capBvhSubtrees = 0;
	np->dStaging = nullptr;
	np->dFlags = np->dFlagScan = np->dContactPair = np->dNumContacts = np->dNumPairs = nullptr;
	np->dScanTmp = nullptr;
	np->dContacts = nullptr;
	np->dSepState = nullptr;
	np->dSepNormal = nullptr;
	np->dMpr = nullptr;
	memset(&np->cw, 0, sizeof(np->cw));
	np->dSatDmin = nullptr;
	np->dSatFlag = np->dSatScan = np->dSatList = np->dSatCount = nullptr;
	np->dSatRecs = nullptr;
	np->dSatBase = np->dSatOwner = nullptr;
	np->satCapRecs = np->satCapBase = 0;
	np->satBodies = np->satTotal = 0;
	np->dSatDirs = nullptr;
	np->dSatDslot = np->dSatDowner = nullptr;
	np->dSatDesc = nullptr;
	np->satCapDirBodies = 0;
	np->satDirBodies = 0;
	np->satGen = 0;
	np->satValid = false;
	np->satNumConvex = 0;
	np->contactsValid = true;
	np->lanesPerPair = 32;
	if ((r = make_stream(stream, &np->stream, &np->ownStream)) || (r = np_alloc_work(np)) || (r = b3h_np_module_init()))
	{
		b3h_np_destroy(np);
		return r;
	}
	*out = np;
	return B3H_OK;
}

extern "C" int b3h_np_destroy(b3hNarrowPhase* np)
{
	if (!np) return B3H_OK;
	hipStreamSynchronize(np->stream);
	dfree(np->dBodies);
	dfree(np->dInertias);
	dfree(np->dCols);
	dfree(np->dConvex);
	dfree(np->dVerts);
	dfree(np->dEdges);
	dfree(np->dFaces);
	dfree(np->dIndices);
	dfree(np->dLocalAabbs);
	dfree(np->dChildren);
	dfree(np->dBvhInfos);
	dfree(np->dBvhNodes);
	dfree(np->dBvhSubtrees);
	dfree(np->dStaging);
	dfree(np->dFlags);
	dfree(np->dFlagScan);
	dfree(np->dScanTmp);
	dfree(np->dContactPair);
	dfree(np->dNumContacts);
	dfree(np->dNumPairs);
	dfree(np->dContacts);
	dfree(np->dSepState);
	dfree(np->dSepNormal);
	dfree(np->dMpr);
	dfree(np->cw.counts);
	dfree(np->cw.offsets);
	dfree(np->cw.cpairs);
	dfree(np->cw.num);
	dfree(np->cw.flags);
	dfree(np->cw.scan);
	dfree(np->cw.scanTmp);
	dfree(np->dSatDmin);
	dfree(np->dSatFlag);
	dfree(np->dSatScan);
	dfree(np->dSatList);
	dfree(np->dSatCount);
	dfree(np->dSatRecs);
	dfree(np->dSatBase);
	dfree(np->dSatOwner);
	dfree(np->dSatDirs);
	dfree(np->dSatDslot);
	dfree(np->dSatDowner);
	dfree(np->dSatDesc);
	if (np->ownStream) hipStreamDestroy(np->stream);
	delete np;
	return B3H_OK;
}

/* One polyhedron into the shape tables (b3GpuNarrowPhase::registerConvexHullShape(b3ConvexUtility*)
 * and registerConvexHullShapeInternal, b3GpuNarrowPhase.cpp:233-296, 321-368): the record copies the
 * caller's mC / mE / m_extents / m_radius, the local centre is the vertex mean (written back through
 * centerOut, as the reference writes it into the utility), and the local AABB bounds the vertices. */
static int np_add_polyhedron(b3hNarrowPhase* np, int nv, const float* v, int vstride, int nf, const float* planes4,
							 const int* faceCounts, const int* indices, int ne, const float* e, int estride,
							 const float* mC4, const float* mE4, const float* extents4, float radius,
							 float* centerOut4, int* outCollidable)
{
	np->gen++;
	*outCollidable = -1;
	if ((int)np->cols.size() >= np->cfg.m_maxConvexShapes)
	{
		b3h_set_error("registerConvexHullShape: exceeding the number of shapes (%d)", np->cfg.m_maxConvexShapes);
		return B3H_ERR_CAPACITY;
	}
	if ((int)np->verts.size() + nv > np->cfg.m_maxConvexVertices)
	{
		b3h_set_error("registerConvexHullShape: exceeding the number of hull vertices (%d)", np->cfg.m_maxConvexVertices);
		return B3H_ERR_CAPACITY;
	}
	size_t ni = 0;
	for (int f = 0; f < nf; f++)
	{
		if (faceCounts[f] < 0)
		{
			b3h_set_error("registerConvexHullShape: negative index count on face %d", f);
			return B3H_ERR_INVALID;
		}
		ni += (size_t)faceCounts[f];
	}
	for (size_t i = 0; i < ni; i++)
		if (indices[i] < 0 || indices[i] >= nv)
		{
			b3h_set_error("registerConvexHullShape: face index %d out of range (%d vertices)", indices[i], nv);
			return B3H_ERR_INVALID;
		}
	b3hConvexPolyhedronData cv;
	memset(&cv, 0, sizeof(cv));
	if (mC4) cv.mC = f4(mC4[0], mC4[1], mC4[2], mC4[3]);
	if (mE4) cv.mE = f4(mE4[0], mE4[1], mE4[2], mE4[3]);
	if (extents4) cv.m_extents = f4(extents4[0], extents4[1], extents4[2], extents4[3]);
	cv.m_radius = radius;
	/* local centre = vertex mean, accumulated in order (b3GpuNarrowPhase.cpp:330-336) */
	float c[3] = {0.f, 0.f, 0.f};
	for (int i = 0; i < nv; i++)
		for (int k = 0; k < 3; k++) c[k] += v[(size_t)i * vstride + k];
	float inv = 1.f / (float)nv;
	cv.m_localCenter = f4(c[0] * inv, c[1] * inv, c[2] * inv, 0.f);
	if (centerOut4)
	{
		centerOut4[0] = cv.m_localCenter.x;
		centerOut4[1] = cv.m_localCenter.y;
		centerOut4[2] = cv.m_localCenter.z;
		centerOut4[3] = 0.f;
	}
	cv.m_uniqueEdgesOffset = (int)np->edges.size();
	cv.m_numUniqueEdges = ne;
	for (int i = 0; i < ne; i++)
		np->edges.push_back(f4(e[(size_t)i * estride], e[(size_t)i * estride + 1], e[(size_t)i * estride + 2], 0.f));
	cv.m_faceOffset = (int)np->faces.size();
	cv.m_numFaces = nf;
	size_t w = 0;
	for (int f = 0; f < nf; f++)
	{
		b3hGpuFace g;
		memset(&g, 0, sizeof(g));
		g.m_plane = f4(planes4[4 * f], planes4[4 * f + 1], planes4[4 * f + 2], planes4[4 * f + 3]);
		g.m_indexOffset = (int)np->indices.size();
		g.m_numIndices = faceCounts[f];
		for (int k = 0; k < faceCounts[f]; k++) np->indices.push_back(indices[w++]);
		np->faces.push_back(g);
	}
	cv.m_vertexOffset = (int)np->verts.size();
	cv.m_numVertices = nv;
	for (int i = 0; i < nv; i++)
		np->verts.push_back(f4(v[(size_t)i * vstride], v[(size_t)i * vstride + 1], v[(size_t)i * vstride + 2], 0.f));
	b3hCollidable col;
	col.m_numChildShapes = 0;
	col.m_radius = 0.f;
	col.m_shapeType = B3H_SHAPE_CONVEX_HULL;
	col.m_shapeIndex = (int)np->convexes.size();
	np->convexes.push_back(cv);
	/* local AABB of the vertices, setMin / setMax (b3GpuNarrowPhase.cpp:341-365) */
	float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
	for (int i = 0; i < nv; i++)
		for (int k = 0; k < 3; k++)
		{
			float x = v[(size_t)i * vstride + k];
			if (x < mn[k]) mn[k] = x;
			if (mx[k] < x) mx[k] = x;
		}
	b3hAabb la;
	la.m_min = f4(mn[0], mn[1], mn[2], 0.f);
	la.m_max = f4(mx[0], mx[1], mx[2], 0.f);
	np->localAabbs.push_back(la);
	*outCollidable = (int)np->cols.size();
	np->cols.push_back(col);
	return B3H_OK;
}

extern "C" int b3h_np_register_convex_hull_shape(b3hNarrowPhase* np, const float* vertices, int strideInBytes,
												  int numVertices, const float* scaling3, int* outCollidable)
{
	CHECK_ARG(np && vertices && numVertices > 0 && strideInBytes >= 12 && (strideInBytes % 4) == 0 && outCollidable);
	*outCollidable = -1;
	float sc[3] = {1.f, 1.f, 1.f};
	if (scaling3)
		for (int k = 0; k < 3; k++) sc[k] = scaling3[k];
	std::vector<float> pts(3 * (size_t)numVertices);
	int stride = strideInBytes / 4;
	for (int i = 0; i < numVertices; i++)
		for (int k = 0; k < 3; k++) pts[3 * i + k] = vertices[(size_t)i * stride + k] * sc[k];
	b3hPolyhedron poly;
	if (b3h_build_polyhedron(pts.data(), numVertices, 3, &poly) != 0)
	{
		b3h_set_error("registerConvexHullShape: degenerate point set (fewer than 3 distinct, non-collinear points)");
		return B3H_ERR_INVALID;
	}
	int nv = (int)poly.vertices.size() / 3, nf = (int)poly.faces.size(), ne = (int)poly.uniqueEdges.size() / 3;
	std::vector<float> planes(4 * (size_t)nf);
	std::vector<int> counts(nf), idx;
	for (int f = 0; f < nf; f++)
	{
		for (int k = 0; k < 4; k++) planes[4 * f + k] = poly.faces[f].plane[k];
		counts[f] = (int)poly.faces[f].indices.size();
		idx.insert(idx.end(), poly.faces[f].indices.begin(), poly.faces[f].indices.end());
	}
	/* b3ConvexUtility leaves mC, mE, m_extents and m_radius unset unless built with
	 * TEST_INTERNAL_OBJECTS (b3ConvexUtility.cpp:399-470); no kernel reads them: zero */
	return np_add_polyhedron(np, nv, poly.vertices.data(), 3, nf, planes.data(), counts.data(), idx.data(), ne,
							 poly.uniqueEdges.data(), 3, nullptr, nullptr, nullptr, 0.f, nullptr, outCollidable);
}

/* b3GpuNarrowPhase::registerSphereShape(radius), b3GpuNarrowPhase.cpp:159-192 */
extern "C" int b3h_np_register_sphere_shape(b3hNarrowPhase* np, float radius, int* outCollidable)
{
	CHECK_ARG(np && outCollidable);
	*outCollidable = -1;
	if ((int)np->cols.size() >= np->cfg.m_maxConvexShapes)
	{
		b3h_set_error("registerSphereShape: exceeding the number of shapes (%d)", np->cfg.m_maxConvexShapes);
		return B3H_ERR_CAPACITY;
	}
	np->gen++;
	b3hCollidable col;
	col.m_numChildShapes = 0;
	col.m_radius = radius;
	col.m_shapeType = B3H_SHAPE_SPHERE;
	col.m_shapeIndex = 0;
	b3hAabb la;
	la.m_min = f4(-radius, -radius, -radius, 0.f);
	la.m_max = f4(radius, radius, radius, 0.f);
	np->localAabbs.push_back(la);
	*outCollidable = (int)np->cols.size();
	np->cols.push_back(col);
	return B3H_OK;
}

/* b3GpuNarrowPhase::registerPlaneShape(planeNormal, planeConstant) and
 * registerFace, b3GpuNarrowPhase.cpp:194-232: the plane n.x = constant is
 * stored as a face (plane (n, constant)); the collidable's radius holds the
 * constant; its local AABB is +-1e30 */
extern "C" int b3h_np_register_plane_shape(b3hNarrowPhase* np, const float* normal3, float constant, int* outCollidable)
{
	CHECK_ARG(np && normal3 && outCollidable);
	*outCollidable = -1;
	if ((int)np->cols.size() >= np->cfg.m_maxConvexShapes)
	{
		b3h_set_error("registerPlaneShape: exceeding the number of shapes (%d)", np->cfg.m_maxConvexShapes);
		return B3H_ERR_CAPACITY;
	}
	np->gen++;
	b3hGpuFace face;
	memset(&face, 0, sizeof(face));
	face.m_plane = f4(normal3[0], normal3[1], normal3[2], constant);
	const int faceIndex = (int)np->faces.size();
	np->faces.push_back(face);
	b3hCollidable col;
	col.m_numChildShapes = 0;
	col.m_radius = constant;
	col.m_shapeType = B3H_SHAPE_PLANE;
	col.m_shapeIndex = faceIndex;
	b3hAabb la;
	la.m_min = f4(-1e30f, -1e30f, -1e30f, 0.f);
	la.m_max = f4(1e30f, 1e30f, 1e30f, 0.f);
	np->localAabbs.push_back(la);
	*outCollidable = (int)np->cols.size();
	np->cols.push_back(col);
	return B3H_OK;
}

/* b3GpuNarrowPhase::registerCompoundShape, b3GpuNarrowPhase.cpp:370-517: a
 * compound of registered convex hulls placed by child transforms; its local
 * AABB bounds the children's transformed local AABBs, and a quantized AABB
 * tree over the children serves the compound-compound pair search */
extern "C" int b3h_np_register_compound_shape(b3hNarrowPhase* np, const b3hGpuChildShape* children, int numChildren,
											  int* outCollidable)
{
	CHECK_ARG(np && children && numChildren > 0 && outCollidable);
	*outCollidable = -1;
	if ((int)np->cols.size() >= np->cfg.m_maxConvexShapes)
	{
		b3h_set_error("registerCompoundShape: exceeding the number of shapes (%d)", np->cfg.m_maxConvexShapes);
		return B3H_ERR_CAPACITY;
	}
	if ((long long)np->children.size() + numChildren >= (long long)np->cfg.m_maxCompoundChildShapes)
	{
		b3h_set_error("registerCompoundShape: exceeding the number of child shapes (%d)", np->cfg.m_maxCompoundChildShapes);
		return B3H_ERR_CAPACITY;
	}
	std::vector<b3hAabb> childLocal((size_t)numChildren);
	for (int i = 0; i < numChildren; i++)
	{
		int c = children[i].m_shapeIndex;
		if (c < 0 || c >= (int)np->cols.size() || np->cols[(size_t)c].m_shapeType != B3H_SHAPE_CONVEX_HULL)
		{
			b3h_set_error("registerCompoundShape: child %d refers to collidable %d, which is not a convex hull", i, c);
			return B3H_ERR_INVALID;
		}
		childLocal[(size_t)i] = np->localAabbs[(size_t)c];
	}
	b3hAabb aabb;
	b3hBvhInfo info;
	std::vector<b3hQuantizedBvhNode> nodes;
	std::vector<b3hBvhSubtreeInfo> subtrees;
	int r = b3h_build_compound(numChildren, children, childLocal.data(), &aabb, &info, nodes, subtrees);
	if (r) return r;
	np->gen++;
	info.m_nodeOffset = (int)np->bvhNodes.size();
	info.m_subTreeOffset = (int)np->bvhSubtrees.size();
	b3hCollidable col;
	col.m_numChildShapes = numChildren;
	col.m_compoundBvhIndex = (int)np->bvhInfos.size();
	col.m_shapeType = B3H_SHAPE_COMPOUND_OF_CONVEX_HULLS;
	col.m_shapeIndex = (int)np->children.size();
	np->children.insert(np->children.end(), children, children + numChildren);
	np->bvhInfos.push_back(info);
	np->bvhNodes.insert(np->bvhNodes.end(), nodes.begin(), nodes.end());
	np->bvhSubtrees.insert(np->bvhSubtrees.end(), subtrees.begin(), subtrees.end());
	np->localAabbs.push_back(aabb);
	*outCollidable = (int)np->cols.size();
	np->cols.push_back(col);
	return B3H_OK;
}

/* the host tables of a compound without registering it (tests and tools) */
extern "C" int b3h_compute_compound_tables(const b3hGpuChildShape* children, int numChildren, const b3hAabb* childLocal,
										   b3hAabb* aabbOut, b3hBvhInfo* infoOut, b3hQuantizedBvhNode* nodesOut,
										   int nodesCap, int* numNodesOut, b3hBvhSubtreeInfo* subtreesOut,
										   int subtreesCap, int* numSubtreesOut)
{
	CHECK_ARG(children && numChildren > 0 && childLocal && aabbOut && infoOut && numNodesOut && numSubtreesOut);
	std::vector<b3hQuantizedBvhNode> nodes;
	std::vector<b3hBvhSubtreeInfo> subtrees;
	int r = b3h_build_compound(numChildren, children, childLocal, aabbOut, infoOut, nodes, subtrees);
	if (r) return r;
	*numNodesOut = (int)nodes.size();
	*numSubtreesOut = (int)subtrees.size();
	CHECK_ARG((int)nodes.size() <= nodesCap && (int)subtrees.size() <= subtreesCap && nodesOut && subtreesOut);
	memcpy(nodesOut, nodes.data(), nodes.size() * sizeof(b3hQuantizedBvhNode));
	memcpy(subtreesOut, subtrees.data(), subtrees.size() * sizeof(b3hBvhSubtreeInfo));
	return B3H_OK;
}

extern "C" int b3h_np_register_convex_polyhedron(b3hNarrowPhase* np, const b3hConvexUtility* u, int* outCollidable)
{
	CHECK_ARG(np && u && outCollidable && u->numVertices > 0 && u->vertices && u->numFaces >= 0 &&
			  u->numUniqueEdges >= 0);
	if (u->numFaces > 0) CHECK_ARG(u->facePlanes && u->faceIndexCounts && u->indices);
	if (u->numUniqueEdges > 0) CHECK_ARG(u->uniqueEdges);
	return np_add_polyhedron(np, u->numVertices, &u->vertices[0].x, 4, u->numFaces, &u->facePlanes[0].x,
							 u->faceIndexCounts, u->indices, u->numUniqueEdges,
							 u->numUniqueEdges ? &u->uniqueEdges[0].x : nullptr, 4, &u->mC.x, &u->mE.x,
							 &u->m_extents.x, u->m_radius, u->localCenterOut, outCollidable);
}

extern "C" int b3h_np_load_shape_tables(b3hNarrowPhase* np, const int* counts, const b3hCollidable* cols,
										const b3hConvexPolyhedronData* cv, const b3hFloat4* verts,
										const b3hFloat4* edges, const b3hGpuFace* faces, const int* indices,
										const b3hAabb* la)
{
	CHECK_ARG(np && counts);
	np->gen++;
	for (int k = 0; k < 7; k++) CHECK_ARG(counts[k] >= 0);
	CHECK_ARG((counts[0] == 0 || cols) && (counts[1] == 0 || cv) && (counts[2] == 0 || verts) &&
			  (counts[3] == 0 || edges) && (counts[4] == 0 || faces) && (counts[5] == 0 || indices) &&
			  (counts[6] == 0 || la) && counts[6] == counts[0]);
	if (counts[0] > np->cfg.m_maxConvexShapes || counts[2] > np->cfg.m_maxConvexVertices)
	{
		b3h_set_error("load_shape_tables: %d shapes / %d vertices exceed the configured capacity", counts[0], counts[2]);
		return B3H_ERR_CAPACITY;
	}
	/* every offset a kernel follows must stay inside the tables it indexes */
	for (int i = 0; i < counts[0]; i++)
		if (cols[i].m_shapeType == B3H_SHAPE_CONVEX_HULL && (cols[i].m_shapeIndex < 0 || cols[i].m_shapeIndex >= counts[1]))
		{
			b3h_set_error("load_shape_tables: collidable %d names hull %d of %d", i, cols[i].m_shapeIndex, counts[1]);
			return B3H_ERR_INVALID;
		}
	for (int i = 0; i < counts[1]; i++)
	{
		const b3hConvexPolyhedronData& c = cv[i];
		bool ok = c.m_numVertices > 0 && c.m_vertexOffset >= 0 && c.m_vertexOffset + c.m_numVertices <= counts[2] &&
				  c.m_numFaces >= 0 && c.m_faceOffset >= 0 && c.m_faceOffset + c.m_numFaces <= counts[4] &&
				  c.m_numUniqueEdges >= 0 && c.m_uniqueEdgesOffset >= 0 &&
				  c.m_uniqueEdgesOffset + c.m_numUniqueEdges <= counts[3];
		for (int f = 0; ok && f < c.m_numFaces; f++)
		{
			const b3hGpuFace& g = faces[c.m_faceOffset + f];
			ok = g.m_numIndices >= 0 && g.m_indexOffset >= 0 && g.m_indexOffset + g.m_numIndices <= counts[5];
			for (int k = 0; ok && k < g.m_numIndices; k++)
				ok = indices[g.m_indexOffset + k] >= 0 && indices[g.m_indexOffset + k] < c.m_numVertices;
		}
		if (!ok)
		{
			b3h_set_error("load_shape_tables: hull %d has offsets or indices outside the tables", i);
			return B3H_ERR_INVALID;
		}
	}
	np->cols.assign(cols, cols + counts[0]);
	np->convexes.assign(cv, cv + counts[1]);
	np->verts.assign(verts, verts + counts[2]);
	np->edges.assign(edges, edges + counts[3]);
	np->faces.assign(faces, faces + counts[4]);
	np->indices.assign(indices, indices + counts[5]);
	np->localAabbs.assign(la, la + counts[6]);
	return B3H_OK;
}

/* b3Matrix3x3::setRotation (scalar path, b3Matrix3x3.h:201-268) */
void b3h_quat_to_basis(const float q[4], float m[3][3])
{
	float d = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
	float s = 2.0f / d;
	float xs = q[0] * s, ys = q[1] * s, zs = q[2] * s;
	float wx = q[3] * xs, wy = q[3] * ys, wz = q[3] * zs;
	float xx = q[0] * xs, xy = q[0] * ys, xz = q[0] * zs;
	float yy = q[1] * ys, yz = q[1] * zs, zz = q[2] * zs;
	m[0][0] = 1.0f - (yy + zz);
	m[0][1] = xy - wz;
	m[0][2] = xz + wy;
	m[1][0] = xy + wz;
	m[1][1] = 1.0f - (xx + zz);
	m[1][2] = yz - wx;
	m[2][0] = xz - wy;
	m[2][1] = yz + wx;
	m[2][2] = 1.0f - (xx + yy);
}

/* inertia record of a registered body: a box of edges 2*(aabbMax-aabbMin)
 * (the AABB already carries the 0.01 margin), world inverse inertia
 * R diag(1/I) R^T (b3GpuNarrowPhase.cpp:860-901) */
static void body_inertia(float mass, const float* aabbMin, const float* aabbMax, const float* orn, b3hInertiaData* in)
{
	memset(in, 0, sizeof(*in));
	if (mass == 0.f) return;
	float he[3] = {aabbMax[0] - aabbMin[0], aabbMax[1] - aabbMin[1], aabbMax[2] - aabbMin[2]};
	float lx = 2.f * he[0], ly = 2.f * he[1], lz = 2.f * he[2];
	float li[3] = {(mass / 12.0f) * (ly * ly + lz * lz), (mass / 12.0f) * (lx * lx + lz * lz),
				   (mass / 12.0f) * (lx * lx + ly * ly)};
	float il[3] = {1.f / li[0], 1.f / li[1], 1.f / li[2]};
	in->m_initInvInertia.m_row[0] = f4(il[0], 0, 0, 0);
	in->m_initInvInertia.m_row[1] = f4(0, il[1], 0, 0);
	in->m_initInvInertia.m_row[2] = f4(0, 0, il[2], 0);
	float m[3][3];
	b3h_quat_to_basis(orn, m);
	float ms[3][3];
	for (int r = 0; r < 3; r++)
		for (int k = 0; k < 3; k++) ms[r][k] = m[r][k] * il[k];
	/* (m.scaled(il)) * m^T: row r, column k = ms[r] . m[k] (b3Matrix3x3 operator*, via tdotx/y/z of m^T) */
	for (int r = 0; r < 3; r++)
	{
		float w[3];
		for (int k = 0; k < 3; k++) w[k] = m[k][0] * ms[r][0] + m[k][1] * ms[r][1] + m[k][2] * ms[r][2];
		in->m_invInertiaWorld.m_row[r] = f4(w[0], w[1], w[2], 0.f);
	}
}

extern "C" int b3h_np_register_rigid_body(b3hNarrowPhase* np, int collidableIndex, float mass, const float* pos,
										  const float* orn, const float* aabbMin, const float* aabbMax, int writeToGpu,
										  int* outBody)
{
	CHECK_ARG(np && pos && orn && aabbMin && aabbMax && outBody);
	np->gen++;
	*outBody = -1;
	int idx = (int)np->bodies.size();
	if (idx >= np->cfg.m_maxConvexBodies)
	{
		b3h_set_error("registerRigidBody: exceeding the number of rigid bodies, %d > %d", idx, np->cfg.m_maxConvexBodies);
		return B3H_ERR_CAPACITY;
	}
	b3hRigidBodyData b;
	memset(&b, 0, sizeof(b));
	b.m_frictionCoeff = 1.f;
	b.m_restituitionCoeff = 0.f;
	b.m_angVel = f4(0, 0, 0, 0);
	b.m_linVel = f4(0, 0, 0, 0);
	b.m_pos = f4(pos[0], pos[1], pos[2], 0.f);
	b.m_quat = f4(orn[0], orn[1], orn[2], orn[3]);
	b.m_collidableIdx = collidableIndex;
	b.m_invMass = mass ? 1.f / mass : 0.f;
	b3hInertiaData in;
	if (mass == 0.f && idx == 0) np->static0Index = 0;
	body_inertia(mass, aabbMin, aabbMax, orn, &in);
	np->bodies.push_back(b);
	np->inertias.push_back(in);
	if (writeToGpu)
	{
		B3H_CHECK(hipMemcpyAsync(np->dBodies + idx, &np->bodies[idx], sizeof(b), hipMemcpyHostToDevice, np->stream));
		B3H_CHECK(hipMemcpyAsync(np->dInertias + idx, &np->inertias[idx], sizeof(in), hipMemcpyHostToDevice, np->stream));
		B3H_CHECK(hipStreamSynchronize(np->stream));
	}
	*outBody = idx;
	return B3H_OK;
}

template <typename T>
static int upload_table(T** d, size_t* cap, const std::vector<T>& h, hipStream_t s)
{
	if (h.size() > *cap || !*d)
	{
		dfree(*d);
		int r = dalloc(d, h.size() + 16);
		if (r) return r;
		*cap = h.size() + 16;
	}
	if (!h.empty()) B3H_CHECK(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
	return B3H_OK;
}

extern "C" int b3h_np_write_all_bodies_to_gpu(b3hNarrowPhase* np)
{
	CHECK_ARG(np);
	np->gen++;
	int r;
	if ((r = upload_table(&np->dCols, &np->capCols, np->cols, np->stream))) return r;
	if ((r = upload_table(&np->dLocalAabbs, &np->capLocal, np->localAabbs, np->stream))) return r;
	if ((r = upload_table(&np->dConvex, &np->capConvex, np->convexes, np->stream))) return r;
	if ((r = upload_table(&np->dVerts, &np->capVerts, np->verts, np->stream))) return r;
	if ((r = upload_table(&np->dEdges, &np->capEdges, np->edges, np->stream))) return r;
	if ((r = upload_table(&np->dFaces, &np->capFaces, np->faces, np->stream))) return r;
	if ((r = upload_table(&np->dIndices, &np->capIndices, np->indices, np->stream))) return r;
	if ((r = upload_table(&np->dChildren, &np->capChildren, np->children, np->stream))) return r;
	if ((r = upload_table(&np->dBvhInfos, &np->capBvhInfos, np->bvhInfos, np->stream))) return r;
	if ((r = upload_table(&np->dBvhNodes, &np->capBvhNodes, np->bvhNodes, np->stream))) return r;
	if ((r = upload_table(&np->dBvhSubtrees, &np->capBvhSubtrees, np->bvhSubtrees, np->stream))) return r;
	if (!np->bodies.empty())
	{
		B3H_CHECK(hipMemcpyAsync(np->dBodies, np->bodies.data(), np->bodies.size() * sizeof(b3hRigidBodyData),
								 hipMemcpyHostToDevice, np->stream));
		B3H_CHECK(hipMemcpyAsync(np->dInertias, np->inertias.data(), np->inertias.size() * sizeof(b3hInertiaData),
								 hipMemcpyHostToDevice, np->stream));
	}
	B3H_CHECK(hipStreamSynchronize(np->stream));
	return B3H_OK;
}

extern "C" int b3h_np_reset(b3hNarrowPhase* np)
{
	CHECK_ARG(np);
	np->gen++;
	np->cols.clear();
	np->convexes.clear();
	np->verts.clear();
	np->edges.clear();
	np->faces.clear();
	np->indices.clear();
	np->localAabbs.clear();
	np->children.clear();
	np->bvhInfos.clear();
	np->bvhNodes.clear();
	np->bvhSubtrees.clear();
	np->bodies.clear();
	np->inertias.clear();
	np->static0Index = -1;
	return B3H_OK;
}

extern "C" int b3h_np_readback_all_bodies_to_cpu(b3hNarrowPhase* np, b3hRigidBodyData* dst)
{
	CHECK_ARG(np);
	if (np->bodies.empty()) return B3H_OK;
	B3H_CHECK(hipMemcpyAsync(np->bodies.data(), np->dBodies, np->bodies.size() * sizeof(b3hRigidBodyData),
							 hipMemcpyDeviceToHost, np->stream));
	B3H_CHECK(hipStreamSynchronize(np->stream));
	if (dst) memcpy(dst, np->bodies.data(), np->bodies.size() * sizeof(b3hRigidBodyData));
	return B3H_OK;
}

extern "C" int b3h_np_set_object_transform_cpu(b3hNarrowPhase* np, const float* p, const float* o, int i)
{
	CHECK_ARG(np && p && o && i >= 0 && i < (int)np->bodies.size());
	np->gen++;
	np->bodies[i].m_pos = f4(p[0], p[1], p[2], 0.f);
	np->bodies[i].m_quat = f4(o[0], o[1], o[2], o[3]);
	return B3H_OK;
}
extern "C" int b3h_np_set_object_velocity_cpu(b3hNarrowPhase* np, const float* lv, const float* av, int i)
{
	CHECK_ARG(np && lv && av && i >= 0 && i < (int)np->bodies.size());
	np->gen++;
	np->bodies[i].m_linVel = f4(lv[0], lv[1], lv[2], 0.f);
	np->bodies[i].m_angVel = f4(av[0], av[1], av[2], 0.f);
	return B3H_OK;
}
extern "C" int b3h_np_get_object_transform_from_cpu(b3hNarrowPhase* np, float* p, float* o, int i)
{
	CHECK_ARG(np && p && o && i >= 0 && i < (int)np->bodies.size());
	const b3hRigidBodyData& b = np->bodies[i];
	p[0] = b.m_pos.x;
	p[1] = b.m_pos.y;
	p[2] = b.m_pos.z;
	p[3] = 1.f;
	o[0] = b.m_quat.x;
	o[1] = b.m_quat.y;
	o[2] = b.m_quat.z;
	o[3] = b.m_quat.w;
	return B3H_OK;
}

static b3hShapeTables np_tables(b3hNarrowPhase* np)
{
	b3hShapeTables t;
	t.collidables = np->dCols;
	t.convexes = np->dConvex;
	t.vertices = np->dVerts;
	t.uniqueEdges = np->dEdges;
	t.faces = np->dFaces;
	t.indices = np->dIndices;
	t.children = np->dChildren;
	t.bvhInfos = np->dBvhInfos;
	t.bvhNodes = np->dBvhNodes;
	t.bvhSubtrees = np->dBvhSubtrees;
	return t;
}

/* Record ranges of the per-body face records (b3hSatCache).  Called outside
 * graph capture before every narrowphase entry; the ranges are rebuilt only
 * when a body's collidable or the shape tables changed. */
/* the compound buffers, once a compound shape exists: child-pair lists and
 * a staging area for their contacts after the pairs' records */
static int np_ensure_compound(b3hNarrowPhase* np)
{
	if (np->cw.cap > 0) return B3H_OK;
	bool any = false;
	for (const b3hCollidable& c : np->cols) any = any || c.m_shapeType == B3H_SHAPE_COMPOUND_OF_CONVEX_HULLS;
	if (!any) return B3H_OK;
	int mp = np->cfg.m_maxBroadphasePairs;
	int cap = np->cfg.m_compoundPairCapacity > 0 ? np->cfg.m_compoundPairCapacity : 1;
	B3H_CHECK(hipStreamSynchronize(np->stream));
	b3hCompoundWork& w = np->cw;
	int r;
	size_t tmp = b3h_scan_tmp_size(mp > cap ? mp : cap);
	if ((r = dalloc(&w.counts, (size_t)mp + 1)) || (r = dalloc(&w.offsets, (size_t)mp + 2)) ||
		(r = dalloc(&w.cpairs, (size_t)cap)) || (r = dalloc(&w.num, 2)) || (r = dalloc(&w.flags, (size_t)cap + 1)) ||
		(r = dalloc(&w.scan, (size_t)cap + 2)) || (r = dalloc(&w.scanTmp, tmp)))
		return r;
	B3H_CHECK(hipMemset(w.num, 0, 2 * sizeof(int)));
	dfree(np->dStaging);
	if ((r = dalloc(&np->dStaging, (size_t)mp + (size_t)cap))) return r;
	w.stagingBase = mp;
	w.cap = cap;
	np->gen++;
	return B3H_OK;
}

static int np_ensure_sat(b3hNarrowPhase* np)
{
	if (int rc = np_ensure_compound(np)) return rc;
	if (np->satValid && np->satGen == np->gen) return B3H_OK;
	int nb = (int)np->bodies.size();
	bool same = np->satValid && (int)np->satCols.size() == nb && np->satNumConvex == np->convexes.size();
	for (int b = 0; same && b < nb; b++) same = np->satCols[(size_t)b] == np->bodies[(size_t)b].m_collidableIdx;
	if (same)
	{
		np->satGen = np->gen;
		return B3H_OK;
	}
	np->satValid = false;
	auto hullOf = [&](int b) -> const b3hConvexPolyhedronData* {
		int ci = np->bodies[(size_t)b].m_collidableIdx;
		if (ci < 0 || ci >= (int)np->cols.size() || np->cols[(size_t)ci].m_shapeType != B3H_SHAPE_CONVEX_HULL)
			return nullptr;
		int si = np->cols[(size_t)ci].m_shapeIndex;
		return (si >= 0 && si < (int)np->convexes.size()) ? &np->convexes[(size_t)si] : nullptr;
	};
	/* face-record ranges */
	std::vector<int> base((size_t)nb + 1);
	long long total = 0;
	for (int b = 0; b < nb; b++)
	{
		base[(size_t)b] = (int)std::min<long long>(total, 1ll << 30);
		if (const b3hConvexPolyhedronData* h = hullOf(b)) total += h->m_numFaces;
	}
	if (total > (1ll << 30)) total = 0; /* too large: the SAT projects the faces itself */
	base[(size_t)nb] = (int)total;
	std::vector<int> owner((size_t)total);
	for (int b = 0; total > 0 && b < nb; b++)
		for (int r = base[(size_t)b]; r < base[(size_t)b + 1]; r++) owner[(size_t)r] = b;
	/* fixed-direction intervals for every body that can meet a pair whose
	 * edge-pair count exceeds the 162 directions */
	long long maxE = 0;
	for (int b = 0; b < nb; b++)
		if (const b3hConvexPolyhedronData* h = hullOf(b)) maxE = std::max<long long>(maxE, h->m_numUniqueEdges);
	std::vector<int> dslot((size_t)nb, -1), downer;
	for (int b = 0; b < nb; b++)
	{
		const b3hConvexPolyhedronData* h = hullOf(b);
		if (h && (long long)h->m_numUniqueEdges * maxE > B3H_SAT_DIRS)
		{
			dslot[(size_t)b] = (int)downer.size();
			downer.push_back(b);
		}
	}
	B3H_CHECK(hipStreamSynchronize(np->stream)); /* earlier steps may still read the old ranges */
	int r;
	size_t nbAlloc = (size_t)(nb > 0 ? nb : 1);
	if (np->satCapBase < nbAlloc + 1)
	{
		dfree(np->dSatBase);
		dfree(np->dSatDslot);
		dfree(np->dSatDesc);
		np->satCapBase = 0;
		if ((r = dalloc(&np->dSatBase, nbAlloc + 1)) || (r = dalloc(&np->dSatDslot, nbAlloc + 1)) ||
			(r = dalloc(&np->dSatDesc, nbAlloc)))
			return r;
		np->satCapBase = nbAlloc + 1;
	}
	if (np->satCapDirBodies < downer.size())
	{
		dfree(np->dSatDirs);
		dfree(np->dSatDowner);
		np->satCapDirBodies = 0;
		if ((r = dalloc(&np->dSatDirs, downer.size() * B3H_SAT_DIRS)) || (r = dalloc(&np->dSatDowner, downer.size())))
			return r;
		np->satCapDirBodies = downer.size();
	}
	if (np->satCapRecs < (size_t)total)
	{
		dfree(np->dSatRecs);
		dfree(np->dSatOwner);
		np->satCapRecs = 0;
		if ((r = dalloc(&np->dSatRecs, (size_t)total)) || (r = dalloc(&np->dSatOwner, (size_t)total))) return r;
		np->satCapRecs = (size_t)total;
	}
	if (nb > 0)
	{
		B3H_CHECK(hipMemcpy(np->dSatDslot, dslot.data(), sizeof(int) * (size_t)nb, hipMemcpyHostToDevice));
		B3H_CHECK(hipMemcpy(np->dSatBase, base.data(), sizeof(int) * ((size_t)nb + 1), hipMemcpyHostToDevice));
	}
	if (!downer.empty())
		B3H_CHECK(hipMemcpy(np->dSatDowner, downer.data(), sizeof(int) * downer.size(), hipMemcpyHostToDevice));
	if (total > 0)
		B3H_CHECK(hipMemcpy(np->dSatOwner, owner.data(), sizeof(int) * (size_t)total, hipMemcpyHostToDevice));
	np->satDirBodies = (int)downer.size();
	np->satCols.resize((size_t)nb);
	for (int b = 0; b < nb; b++) np->satCols[(size_t)b] = np->bodies[(size_t)b].m_collidableIdx;
	np->satNumConvex = np->convexes.size();
	np->satBodies = nb;
	np->satTotal = (int)total;
	np->satGen = np->gen;
	np->satValid = true;
	return B3H_OK;
}

/* device-count form used inside a step; the per-body SAT data must be set up
 * (np_ensure_sat) before, outside any graph capture */
static int np_compute_contacts_dev(b3hNarrowPhase* np, b3hInt4* pairs, const int* dNumPairs, int maxPairs,
								   hipEvent_t evAfterKernel = nullptr, hipEvent_t evAfterMpr = nullptr)
{
	if (!np->satValid || np->satGen != np->gen || np->satBodies != (int)np->bodies.size())
	{
		b3h_set_error("narrowphase: per-body SAT data not set up for the current bodies");
		return B3H_ERR_STATE;
	}
	bool prim = false;
	for (const b3hCollidable& c : np->cols) prim = prim || c.m_shapeType != B3H_SHAPE_CONVEX_HULL;
	b3hSatCache sc = {np->satTotal > 0 ? np->dSatRecs : nullptr,
					  np->dSatBase,
					  np->dSatOwner,
					  np->satBodies,
					  np->satTotal,
					  np->satDirBodies > 0 ? np->dSatDirs : nullptr,
					  np->dSatDslot,
					  np->dSatDowner,
					  np->satDirBodies,
					  np->dSatDesc};
	b3h_narrowphase(pairs, dNumPairs, maxPairs, np->dBodies, np_tables(np), np->dStaging, np->dFlags, np->dFlagScan,
					np->dScanTmp, np->dContactPair, np->cfg.m_maxContactCapacity, np->dNumContacts, np->dSepState,
					np->dSepNormal, np->lanesPerPair, np->stream, evAfterKernel, np->dMpr, evAfterMpr, prim, &sc,
					b3hSatWork{np->dSatDmin, np->dSatFlag, np->dSatScan, np->dSatList, np->dSatCount}, np->cw);
	np->contactsValid = false;
	return B3H_OK;
}

extern "C" int b3h_np_compute_contacts(b3hNarrowPhase* np, b3hInt4* pairs, int numPairs, const b3hAabb* aabbs, int numObjects)
{
	CHECK_ARG(np && numPairs >= 0 && numPairs <= np->cfg.m_maxBroadphasePairs);
	(void)aabbs;
	(void)numObjects;
	if (numPairs > 0) CHECK_ARG(pairs);
	if (int re = np_ensure_sat(np)) return re;
	B3H_CHECK(hipMemcpyAsync(np->dNumPairs, &numPairs, sizeof(int), hipMemcpyHostToDevice, np->stream));
	if (int rc = np_compute_contacts_dev(np, pairs, np->dNumPairs, numPairs > 0 ? numPairs : 1)) return rc;
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}

extern "C" int b3h_np_get_bodies_gpu(b3hNarrowPhase* np, b3hRigidBodyData** dev, int* n)
{
	CHECK_ARG(np && dev);
	*dev = np->dBodies;
	if (n) *n = (int)np->bodies.size();
	return B3H_OK;
}
extern "C" int b3h_np_get_body_inertias_gpu(b3hNarrowPhase* np, b3hInertiaData** dev, int* n)
{
	CHECK_ARG(np && dev);
	*dev = np->dInertias;
	if (n) *n = (int)np->inertias.size();
	return B3H_OK;
}
extern "C" int b3h_np_get_collidables_gpu(b3hNarrowPhase* np, b3hCollidable** dev, int* n)
{
	CHECK_ARG(np && dev);
	*dev = np->dCols;
	if (n) *n = (int)np->cols.size();
	return B3H_OK;
}
extern "C" int b3h_np_get_contacts_gpu(b3hNarrowPhase* np, b3hContact4Data** dev, int* n)
{
	CHECK_ARG(np && dev);
	if (!np->contactsValid)
	{
		b3h_gather_contacts(np->dContactPair, np->dNumContacts, np->cfg.m_maxContactCapacity, np->dStaging,
							np->dContacts, np->stream);
		np->contactsValid = true;
	}
	*dev = np->dContacts;
	if (n)
	{
		B3H_CHECK(hipMemcpyAsync(n, np->dNumContacts, sizeof(int), hipMemcpyDeviceToHost, np->stream));
	}
	B3H_CHECK(hipStreamSynchronize(np->stream));
	return B3H_OK;
}
extern "C" int b3h_np_get_static0_index(b3hNarrowPhase* np, int* index)
{
	CHECK_ARG(np && index);
	*index = np->static0Index;
	return B3H_OK;
}
extern "C" int b3h_np_get_num_rigid_bodies(b3hNarrowPhase* np, int* n)
{
	CHECK_ARG(np && n);
	*n = (int)np->bodies.size();
	return B3H_OK;
}
extern "C" int b3h_np_get_local_space_aabb(b3hNarrowPhase* np, int ci, b3hAabb* out)
{
	CHECK_ARG(np && out && ci >= 0 && ci < (int)np->localAabbs.size());
	*out = np->localAabbs[ci];
	return B3H_OK;
}
extern "C" int b3h_np_get_shape_tables(b3hNarrowPhase* np, int* counts, b3hCollidable* cols, b3hConvexPolyhedronData* cv,
									   b3hFloat4* verts, b3hFloat4* edges, b3hGpuFace* faces, int* indices, b3hAabb* la)
{
	CHECK_ARG(np && counts);
	counts[0] = (int)np->cols.size();
	counts[1] = (int)np->convexes.size();
	counts[2] = (int)np->verts.size();
	counts[3] = (int)np->edges.size();
	counts[4] = (int)np->faces.size();
	counts[5] = (int)np->indices.size();
	counts[6] = (int)np->localAabbs.size();
	if (cols && !np->cols.empty()) memcpy(cols, np->cols.data(), np->cols.size() * sizeof(*cols));
	if (cv && !np->convexes.empty()) memcpy(cv, np->convexes.data(), np->convexes.size() * sizeof(*cv));
	if (verts && !np->verts.empty()) memcpy(verts, np->verts.data(), np->verts.size() * sizeof(*verts));
	if (edges && !np->edges.empty()) memcpy(edges, np->edges.data(), np->edges.size() * sizeof(*edges));
	if (faces && !np->faces.empty()) memcpy(faces, np->faces.data(), np->faces.size() * sizeof(*faces));
	if (indices && !np->indices.empty()) memcpy(indices, np->indices.data(), np->indices.size() * sizeof(*indices));
	if (la && !np->localAabbs.empty()) memcpy(la, np->localAabbs.data(), np->localAabbs.size() * sizeof(*la));
	return B3H_OK;
}
/* the compound tables (counts: children, tree infos, tree nodes, subtrees) */
extern "C" int b3h_np_get_compound_tables(b3hNarrowPhase* np, int* counts, b3hGpuChildShape* children,
										  b3hBvhInfo* infos, b3hQuantizedBvhNode* nodes, b3hBvhSubtreeInfo* subtrees)
{
	CHECK_ARG(np && counts);
	counts[0] = (int)np->children.size();
	counts[1] = (int)np->bvhInfos.size();
	counts[2] = (int)np->bvhNodes.size();
	counts[3] = (int)np->bvhSubtrees.size();
	if (children && counts[0]) memcpy(children, np->children.data(), np->children.size() * sizeof(*children));
	if (infos && counts[1]) memcpy(infos, np->bvhInfos.data(), np->bvhInfos.size() * sizeof(*infos));
	if (nodes && counts[2]) memcpy(nodes, np->bvhNodes.data(), np->bvhNodes.size() * sizeof(*nodes));
	if (subtrees && counts[3]) memcpy(subtrees, np->bvhSubtrees.data(), np->bvhSubtrees.size() * sizeof(*subtrees));
	return B3H_OK;
}
extern "C" int b3h_np_get_inertias_cpu(b3hNarrowPhase* np, b3hInertiaData* dst, int n)
{
	CHECK_ARG(np && dst && n <= (int)np->inertias.size());
	if (n > 0) memcpy(dst, np->inertias.data(), (size_t)n * sizeof(*dst));
	return B3H_OK;
}
extern "C" int b3h_np_set_pairs_per_group(b3hNarrowPhase* np, int lanes)
{
	CHECK_ARG(np && lanes >= 1 && lanes <= 64 && (lanes & (lanes - 1)) == 0);
	np->gen++;
	np->lanesPerPair = lanes;
	return B3H_OK;
}
extern "C" int b3h_np_get_mpr_debug(b3hNarrowPhase* np, b3hFloat4** mpr)
{
	CHECK_ARG(np && mpr);
	*mpr = np->dMpr;
	return B3H_OK;
}
extern "C" int b3h_np_get_sat_debug(b3hNarrowPhase* np, int** st, b3hFloat4** sn)
{
	CHECK_ARG(np && st && sn);
	*st = np->dSepState;
	*sn = np->dSepNormal;
	return B3H_OK;
}

/* ======================================================================= */
struct b3hBroadphase
{
	int kind;
	b3hConfig cfg;
	hipStream_t stream;     /* stream every call is issued on */
	hipStream_t homeStream; /* the stream given at create (or owned) */
	bool ownStream;
	std::vector<b3hAabb> all;
	std::vector<int> smallMap, largeMap;
	b3hGridParams gp;
	float cellSize[3];
	int numCells;
	b3hAabb* dAll;
	int *dSmall, *dLarge;
	b3hInt4* dPairs;
	int* dPairCount; /* [count, total] */
	b3hBroadphaseScratch sc;
	int scCapSmall;
	int scCapLarge;
	int scCells;
	unsigned gen; /* as b3hNarrowPhase::gen: proxies, grid parameters, scratch reallocation */
};

static void bp_free_scratch(b3hBroadphase* bp)
{
	dfree(bp->sc.keys);
	dfree(bp->sc.vals);
	dfree(bp->sc.tmpKeys);
	dfree(bp->sc.tmpVals);
	dfree(bp->sc.hist);
	dfree(bp->sc.cellStart);
	dfree(bp->sc.counts);
	dfree(bp->sc.offsets);
	dfree(bp->sc.scanTmp);
	dfree(bp->sc.sortedAabbs);
	dfree(bp->sc.sapPartials);
	dfree(bp->sc.axis);
	dfree(bp->sc.stash);
	dfree(bp->sc.lkeys);
	dfree(bp->sc.lvals);
	dfree(bp->sc.ltmpKeys);
	dfree(bp->sc.ltmpVals);
	dfree(bp->sc.lhist);
	dfree(bp->sc.lmaxExt);
	dfree(bp->sc.lminx);
	dfree(bp->sc.bvhMerge);
	dfree(bp->sc.bvhChild);
	dfree(bp->sc.bvhBox);
	dfree(bp->sc.bvhFlag);
	dfree(bp->sc.bvhMeta);
}

extern "C" int b3h_bp_create(int kind, const b3hConfig* cfg, void* stream, b3hBroadphase** out)
{
	CHECK_ARG(cfg && out && (kind == B3H_BROADPHASE_GRID || kind == B3H_BROADPHASE_SAP || kind == B3H_BROADPHASE_PLBVH));
	*out = nullptr;
	int r = ensure_device();
	if (r) return r;
	b3hBroadphase* bp = new b3hBroadphase();
	bp->kind = kind;
	bp->cfg = *cfg;
	memset(&bp->sc, 0, sizeof(bp->sc));
	bp->scCapSmall = 0;
	bp->scCapLarge = 0;
	bp->scCells = 0;
	bp->dAll = nullptr;
	bp->dSmall = bp->dLarge = nullptr;
	bp->dPairs = nullptr;
	bp->dPairCount = nullptr;
	float cs[3] = {3.f, 3.f, 3.f};
	int dims[3] = {128, 128, 128};
	for (int k = 0; k < 3; k++)
	{
		bp->cellSize[k] = cs[k];
		bp->gp.invCell[k] = 1.f / cs[k];
		bp->gp.dims[k] = dims[k];
	}
	bp->gp.maxBodiesPerCell = 256;
	bp->numCells = dims[0] * dims[1] * dims[2];
	bp->ownStream = false;
	if ((r = make_stream(stream, &bp->homeStream, &bp->ownStream)))
	{
		delete bp;
		return r;
	}
	bp->stream = bp->homeStream;
	if ((r = dalloc(&bp->dAll, cfg->m_maxConvexBodies)) ||
		(r = dalloc(&bp->dPairs, cfg->m_maxBroadphasePairs)) || (r = dalloc(&bp->dPairCount, 2)) ||
		(r = dalloc(&bp->dSmall, cfg->m_maxConvexBodies)) || (r = dalloc(&bp->dLarge, cfg->m_maxConvexBodies)))
	{
		b3h_bp_destroy(bp);
		return r;
	}
	hipMemset(bp->dPairCount, 0, 2 * sizeof(int));
	*out = bp;
	return B3H_OK;
}

extern "C" int b3h_bp_destroy(b3hBroadphase* bp)
{
	if (!bp) return B3H_OK;
	hipStreamSynchronize(bp->stream);
	bp_free_scratch(bp);
	dfree(bp->dAll);
	dfree(bp->dSmall);
	dfree(bp->dLarge);
	dfree(bp->dPairs);
	dfree(bp->dPairCount);
	if (bp->ownStream) hipStreamDestroy(bp->homeStream);
	delete bp;
	return B3H_OK;
}

extern "C" int b3h_bp_set_grid_params(b3hBroadphase* bp, const float* cs, const int* dims, int maxPerCell)
{
	CHECK_ARG(bp && cs && dims && maxPerCell > 0);
	bp->gen++;
	for (int k = 0; k < 3; k++)
	{
		CHECK_ARG(cs[k] > 0.f && dims[k] > 0 && (dims[k] & (dims[k] - 1)) == 0);
	}
	CHECK_ARG((long long)dims[0] * dims[1] * dims[2] <= (1ll << 26));
	for (int k = 0; k < 3; k++)
	{
		bp->cellSize[k] = cs[k];
		bp->gp.invCell[k] = 1.f / cs[k];
		bp->gp.dims[k] = dims[k];
	}
	bp->gp.maxBodiesPerCell = maxPerCell;
	bp->numCells = dims[0] * dims[1] * dims[2];
	return B3H_OK;
}

static int bp_proxy(b3hBroadphase* bp, const float* mn, const float* mx, int user, bool large)
{
	CHECK_ARG(bp && mn && mx);
	bp->gen++;
	if ((int)bp->all.size() >= bp->cfg.m_maxConvexBodies)
	{
		b3h_set_error("createProxy: exceeding capacity %d", bp->cfg.m_maxConvexBodies);
		return B3H_ERR_CAPACITY;
	}
	b3hAabb a;
	a.m_min = f4(mn[0], mn[1], mn[2], i2f(user));
	a.m_max = f4(mx[0], mx[1], mx[2], i2f((int)bp->all.size()));
	(large ? bp->largeMap : bp->smallMap).push_back((int)bp->all.size());
	bp->all.push_back(a);
	return B3H_OK;
}
extern "C" int b3h_bp_create_proxy(b3hBroadphase* bp, const float* mn, const float* mx, int user, int g, int m)
{
	(void)g;
	(void)m;
	return bp_proxy(bp, mn, mx, user, false);
}
extern "C" int b3h_bp_create_proxies(b3hBroadphase* bp, int n, const float* mn, const float* mx, const int* user)
{
	CHECK_ARG(bp && n >= 0 && (n == 0 || (mn && mx)));
	bp->gen++;
	if ((long long)bp->all.size() + n > bp->cfg.m_maxConvexBodies)
	{
		b3h_set_error("createProxies: %d more proxies exceed capacity %d", n, bp->cfg.m_maxConvexBodies);
		return B3H_ERR_CAPACITY;
	}
	bp->all.reserve(bp->all.size() + n);
	bp->smallMap.reserve(bp->smallMap.size() + n);
	for (int i = 0; i < n; i++)
	{
		int r = bp_proxy(bp, mn + 3 * (size_t)i, mx + 3 * (size_t)i, user ? user[i] : i, false);
		if (r) return r;
	}
	return B3H_OK;
}
extern "C" int b3h_bp_create_large_proxy(b3hBroadphase* bp, const float* mn, const float* mx, int user, int g, int m)
{
	(void)g;
	(void)m;
	return bp_proxy(bp, mn, mx, user, true);
}

static int bp_ensure_scratch(b3hBroadphase* bp)
{
	int ns = (int)bp->smallMap.size();
	int nl = (int)bp->largeMap.size();
	if (ns <= bp->scCapSmall && nl <= bp->scCapLarge && bp->scCells == bp->numCells && bp->sc.keys) return B3H_OK;
	bp->gen++;
	bp_free_scratch(bp);
	int cap = ns + 64;
	int capL = nl + 64;
	int r;
	if ((r = dalloc(&bp->sc.lkeys, capL)) || (r = dalloc(&bp->sc.lvals, capL)) || (r = dalloc(&bp->sc.ltmpKeys, capL)) ||
		(r = dalloc(&bp->sc.ltmpVals, capL)) || (r = dalloc(&bp->sc.lhist, b3h_radix_hist_size(capL))) ||
		(r = dalloc(&bp->sc.lmaxExt, 1)) || (r = dalloc(&bp->sc.lminx, capL)))
		return r;
	bp->scCapLarge = nl;
	if ((r = dalloc(&bp->sc.keys, cap)) || (r = dalloc(&bp->sc.vals, cap)) || (r = dalloc(&bp->sc.tmpKeys, cap)) ||
		(r = dalloc(&bp->sc.tmpVals, cap)) || (r = dalloc(&bp->sc.hist, b3h_radix_hist_size(cap))) ||
		(r = dalloc(&bp->sc.cellStart, bp->numCells)) || (r = dalloc(&bp->sc.counts, 2 * cap + 1)) ||
		(r = dalloc(&bp->sc.offsets, 2 * cap + 2)) || (r = dalloc(&bp->sc.scanTmp, b3h_scan_tmp_size(2 * cap))) ||
		(r = dalloc(&bp->sc.sortedAabbs, cap)) || (r = dalloc(&bp->sc.sapPartials, 6 * (size_t)(b3h_sap_num_partials(cap) + 1))) ||
		(r = dalloc(&bp->sc.axis, 1)) || (r = dalloc(&bp->sc.stash, (size_t)cap * B3H_GRID_STASH)))
		return r;
	if (bp->kind == B3H_BROADPHASE_PLBVH &&
		((r = dalloc(&bp->sc.bvhMerge, b3h_plbvh_reduce_blocks(cap) + 1)) || (r = dalloc(&bp->sc.bvhChild, cap)) ||
		 (r = dalloc(&bp->sc.bvhBox, 2 * (size_t)cap)) || (r = dalloc(&bp->sc.bvhFlag, cap)) ||
		 (r = dalloc(&bp->sc.bvhMeta, 2))))
		return r;
	bp->scCapSmall = ns;
	bp->scCells = bp->numCells;
	B3H_CHECK(hipMemset(bp->sc.axis, 0, sizeof(int)));
	if (bp->sc.bvhMeta) B3H_CHECK(hipMemset(bp->sc.bvhMeta, 0, 2 * sizeof(int)));
	return B3H_OK;
}

extern "C" int b3h_bp_write_aabbs_to_gpu(b3hBroadphase* bp)
{
	CHECK_ARG(bp);
	bp->gen++;
	int r = bp_ensure_scratch(bp);
	if (r) return r;
	if (!bp->all.empty())
		B3H_CHECK(hipMemcpyAsync(bp->dAll, bp->all.data(), bp->all.size() * sizeof(b3hAabb), hipMemcpyHostToDevice, bp->stream));
	if (!bp->smallMap.empty())
		B3H_CHECK(hipMemcpyAsync(bp->dSmall, bp->smallMap.data(), bp->smallMap.size() * sizeof(int), hipMemcpyHostToDevice,
								 bp->stream));
	if (!bp->largeMap.empty())
		B3H_CHECK(hipMemcpyAsync(bp->dLarge, bp->largeMap.data(), bp->largeMap.size() * sizeof(int), hipMemcpyHostToDevice,
								 bp->stream));
	B3H_CHECK(hipStreamSynchronize(bp->stream));
	return B3H_OK;
}

static int bp_calc_pairs(b3hBroadphase* bp, int maxPairs, hipStream_t s)
{
	if (maxPairs > bp->cfg.m_maxBroadphasePairs) maxPairs = bp->cfg.m_maxBroadphasePairs;
	int r = bp_ensure_scratch(bp);
	if (r) return r;
	int ns = (int)bp->smallMap.size(), nl = (int)bp->largeMap.size();
	if (bp->kind == B3H_BROADPHASE_GRID)
		b3h_grid_pairs(bp->dAll, bp->dSmall, ns, bp->dLarge, nl, bp->gp, bp->numCells, bp->dPairs, maxPairs, bp->dPairCount,
					   bp->dPairCount + 1, bp->sc, s);
	else if (bp->kind == B3H_BROADPHASE_SAP)
		b3h_sap_pairs(bp->dAll, bp->dSmall, ns, bp->dLarge, nl, bp->dPairs, maxPairs, bp->dPairCount, bp->dPairCount + 1,
					  bp->sc, s);
	else
		b3h_plbvh_pairs(bp->dAll, bp->dSmall, ns, bp->dLarge, nl, bp->dPairs, maxPairs, bp->dPairCount,
						bp->dPairCount + 1, bp->sc, s);
	return B3H_OK;
}

extern "C" int b3h_bp_calculate_overlapping_pairs(b3hBroadphase* bp, int maxPairs)
{
	CHECK_ARG(bp && maxPairs > 0);
	int r = bp_calc_pairs(bp, maxPairs, bp->stream);
	if (r) return r;
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}
/* inclusive overlap test on x, y, z (b3TestAabbAgainstAabb2, b3AabbUtil.h; TestAabbAgainstAabb2 in
 * b3GpuSapBroadphase.cpp) */
static bool aabb_overlap(const b3hAabb& a, const b3hAabb& b)
{
	return !(a.m_min.x > b.m_max.x || a.m_max.x < b.m_min.x) && !(a.m_min.z > b.m_max.z || a.m_max.z < b.m_min.z) &&
		   !(a.m_min.y > b.m_max.y || a.m_max.y < b.m_min.y);
}

extern "C" int b3h_bp_calculate_overlapping_pairs_host(b3hBroadphase* bp, int maxPairs)
{
	CHECK_ARG(bp && maxPairs >= 0);
	if (maxPairs > bp->cfg.m_maxBroadphasePairs) maxPairs = bp->cfg.m_maxBroadphasePairs;
	int n = (int)bp->all.size();
	std::vector<b3hAabb> a(n);
	std::vector<int> scene;
	if (n) B3H_CHECK(hipMemcpyAsync(a.data(), bp->dAll, n * sizeof(b3hAabb), hipMemcpyDeviceToHost, bp->stream));
	if (bp->sc.sceneOf && n)
	{
		scene.resize(n);
		B3H_CHECK(hipMemcpyAsync(scene.data(), bp->sc.sceneOf, n * sizeof(int), hipMemcpyDeviceToHost, bp->stream));
	}
	B3H_CHECK(hipStreamSynchronize(bp->stream));
	auto user = [&](int h) {
		int u;
		memcpy(&u, &a[h].m_min.w, 4);
		return u;
	};
	std::vector<b3hInt4> out;
	long long total = 0;
	auto add = [&](int h0, int h1) {
		if (!scene.empty() && scene[h0] != scene[h1]) return;
		int x = user(h0), y = user(h1);
		total++;
		if ((int)out.size() < maxPairs) out.push_back(x <= y ? b3hInt4{x, y, -1, -1} : b3hInt4{y, x, -1, -1});
	};
	if (bp->kind == B3H_BROADPHASE_GRID)
	{
		/* every proxy against every later one (b3GpuGridBroadphase.cpp:260-292) */
		for (int i = 0; i < n; i++)
			for (int j = i + 1; j < n; j++)
				if (aabb_overlap(a[i], a[j])) add(j, i);
	}
	else
	{
		/* small x small, then small x large (b3GpuSapBroadphase.cpp:862-981;
		 * b3GpuParallelLinearBvhBroadphase::calculateOverlappingPairsHost,
		 * b3GpuParallelLinearBvhBroadphase.cpp:59-62, asserts instead: the
		 * BVH broadphase takes the SAP's host loop) */
		int ns = (int)bp->smallMap.size(), nl = (int)bp->largeMap.size();
		for (int i = 0; i < ns; i++)
			for (int j = i + 1; j < ns; j++)
				if (aabb_overlap(a[bp->smallMap[i]], a[bp->smallMap[j]])) add(bp->smallMap[i], bp->smallMap[j]);
		for (int i = 0; i < ns; i++)
			for (int j = 0; j < nl; j++)
				if (aabb_overlap(a[bp->smallMap[i]], a[bp->largeMap[j]])) add(bp->largeMap[j], bp->smallMap[i]);
	}
	int cnt[2] = {(int)out.size(), (int)(total > 0x7fffffff ? 0x7fffffff : total)};
	if (!out.empty())
		B3H_CHECK(hipMemcpyAsync(bp->dPairs, out.data(), out.size() * sizeof(b3hInt4), hipMemcpyHostToDevice, bp->stream));
	B3H_CHECK(hipMemcpyAsync(bp->dPairCount, cnt, sizeof(cnt), hipMemcpyHostToDevice, bp->stream));
	B3H_CHECK(hipStreamSynchronize(bp->stream));
	return B3H_OK;
}

extern "C" int b3h_bp_get_all_aabbs_gpu(b3hBroadphase* bp, b3hAabb** dev, int* n)
{
	CHECK_ARG(bp && dev);
	*dev = bp->dAll;
	if (n) *n = (int)bp->all.size();
	return B3H_OK;
}
extern "C" int b3h_bp_get_all_aabbs_cpu(b3hBroadphase* bp, b3hAabb* host, int capacity, int* n)
{
	CHECK_ARG(bp && n);
	*n = (int)bp->all.size();
	if (!host) return B3H_OK;
	if (capacity < *n)
	{
		b3h_set_error("getAllAabbsCPU: capacity %d < %d proxies", capacity, *n);
		return B3H_ERR_CAPACITY;
	}
	if (*n) memcpy(host, bp->all.data(), (size_t)*n * sizeof(b3hAabb));
	return B3H_OK;
}
extern "C" int b3h_bp_get_small_aabb_indices_gpu(b3hBroadphase* bp, int** dev, int* n)
{
	CHECK_ARG(bp && dev);
	*dev = bp->dSmall;
	if (n) *n = (int)bp->smallMap.size();
	return B3H_OK;
}
extern "C" int b3h_bp_get_large_aabb_indices_gpu(b3hBroadphase* bp, int** dev, int* n)
{
	CHECK_ARG(bp && dev);
	*dev = bp->dLarge;
	if (n) *n = (int)bp->largeMap.size();
	return B3H_OK;
}
extern "C" int b3h_bp_get_aabb_buffer_ws(b3hBroadphase* bp, b3hAabb** dev, int* n)
{
	CHECK_ARG(bp && dev);
	*dev = bp->dAll;
	if (n) *n = (int)bp->all.size();
	return B3H_OK;
}
static int read_int(const int* d, int* h, hipStream_t s)
{
	B3H_CHECK(hipMemcpyAsync(h, d, sizeof(int), hipMemcpyDeviceToHost, s));
	B3H_CHECK(hipStreamSynchronize(s));
	return B3H_OK;
}
extern "C" int b3h_bp_get_num_overlap(b3hBroadphase* bp, int* n)
{
	CHECK_ARG(bp && n);
	return read_int(bp->dPairCount, n, bp->stream);
}
extern "C" int b3h_bp_get_total_found(b3hBroadphase* bp, int* n)
{
	CHECK_ARG(bp && n);
	return read_int(bp->dPairCount + 1, n, bp->stream);
}
extern "C" int b3h_bp_get_overlapping_pair_buffer(b3hBroadphase* bp, b3hInt4** dev)
{
	CHECK_ARG(bp && dev);
	*dev = bp->dPairs;
	return B3H_OK;
}
extern "C" int b3h_bp_get_sap_axis(b3hBroadphase* bp, int* axis)
{
	CHECK_ARG(bp && axis);
	if (!bp->sc.axis)
	{
		*axis = 0;
		return B3H_OK;
	}
	return read_int(bp->sc.axis, axis, bp->stream);
}
extern "C" int b3h_bp_get_bvh_info(b3hBroadphase* bp, int* root, int* overflow)
{
	CHECK_ARG(bp && root && overflow);
	if (bp->kind != B3H_BROADPHASE_PLBVH || !bp->sc.bvhMeta)
	{
		b3h_set_error("getBvhInfo: not a BVH broadphase, or no pairs calculated yet");
		return B3H_ERR_STATE;
	}
	int m[2];
	B3H_CHECK(hipMemcpyAsync(m, bp->sc.bvhMeta, sizeof(m), hipMemcpyDeviceToHost, bp->stream));
	B3H_CHECK(hipStreamSynchronize(bp->stream));
	*root = m[0];
	*overflow = m[1];
	return B3H_OK;
}
extern "C" int b3h_bp_get_grid_sort(b3hBroadphase* bp, const unsigned** k, const unsigned** v, int* n)
{
	CHECK_ARG(bp && k && v && n);
	*k = bp->sc.keys;
	*v = bp->sc.vals;
	*n = (int)bp->smallMap.size();
	return B3H_OK;
}

extern "C" int b3h_bp_get_proxy_maps(b3hBroadphase* bp, int* counts2, int* sm, int* lm)
{
	CHECK_ARG(bp && counts2);
	counts2[0] = (int)bp->smallMap.size();
	counts2[1] = (int)bp->largeMap.size();
	if (sm && !bp->smallMap.empty()) memcpy(sm, bp->smallMap.data(), bp->smallMap.size() * sizeof(int));
	if (lm && !bp->largeMap.empty()) memcpy(lm, bp->largeMap.data(), bp->largeMap.size() * sizeof(int));
	return B3H_OK;
}

/* ======================================================================= */
struct b3hPipeline
{
	b3hNarrowPhase* np;
	b3hBroadphase* bp;
	b3hConfig cfg;
	hipStream_t stream;
	bool ownStream;
	F4 gravity;
	b3hSolverParams prm;
	b3hSolverBuffers sb;
	bool trackResidual;
	bool timing;
	hipEvent_t ev[B3H_NUM_STAGES + 1];
	bool evValid;
	hipGraph_t graph;
	hipGraphExec_t graphExec;
	float graphDt;
	int graphBodies;
	unsigned graphNpGen, graphBpGen; /* component generations the graph was captured at */
	float* dWarm;
	int warmCount;
	/* warm start and contact logging */
	b3hWarmBuffers wb;
	bool warmAlloc, warmFeatAlloc;
	/* batched scenes: scene of every body, per-scene contact counts */
	int numScenes;
	int sceneBodies;
	int* dSceneOf;
	int* dSceneCounts;
	b3hMlp* warmNet;
	bool logRows;
	float matchTol;
	/* ray casts: BVH scratch (sized for raySmallCap small proxies) and the
	 * device copies of the host ray / hit arrays (rayCap rays) */
	b3hBroadphaseScratch raySc;
	int raySmallCap;
	b3hRayInfo* dRays;
	b3hRayHit* dHits;
	int rayCap;
	/* joints: the host list (b3GpuRigidBodyPipeline's m_cpuConstraints), its
	 * device image and the solver's buffers; jointsDirty: the list changed
	 * since the last upload (rows, slots and batches are rebuilt then) */
	std::vector<b3hGpuGenericConstraint> joints;
	int jointUid;
	bool jointsDirty;
	b3hJointBuffers jb;
	b3hJointParams jp;
};

static void joint_free(b3hPipeline* p)
{
	b3hJointBuffers& jb = p->jb;
	dfree(jb.constraints);
	dfree(jb.rowOffset);
	dfree(jb.rowCount);
	dfree(jb.slotOf);
	dfree(jb.rows);
	dfree(jb.batchStart);
	dfree(jb.order);
	dfree(jb.delta);
	dfree(jb.bodySlot);
	jb.numConstraints = jb.numBatches = jb.numSlots = jb.numBodies = 0;
}

static void ray_free(b3hPipeline* p)
{
	b3hBroadphaseScratch& sc = p->raySc;
	dfree(sc.keys);
	dfree(sc.vals);
	dfree(sc.tmpKeys);
	dfree(sc.tmpVals);
	dfree(sc.hist);
	dfree(sc.sortedAabbs);
	dfree(sc.bvhMerge);
	dfree(sc.bvhChild);
	dfree(sc.bvhBox);
	dfree(sc.bvhFlag);
	dfree(sc.bvhMeta);
	p->raySmallCap = -1;
	dfree(p->dRays);
	dfree(p->dHits);
	p->rayCap = 0;
}

static void pl_free(b3hPipeline* p)
{
	dfree(p->sb.cellOf);
	dfree(p->sb.cellCounts);
	dfree(p->sb.cellOffsets);
	dfree(p->sb.cellCursor);
	dfree(p->sb.cellSlots);
	dfree(p->sb.perm);
	dfree(p->sb.batchIdx);
	dfree(p->sb.batchSizes);
	dfree(p->sb.sortKeys);
	dfree(p->sb.constraints);
	dfree(p->sb.sortedContacts);
	dfree(p->sb.resid);
	dfree(p->sb.cellBodies);
	dfree(p->sb.cellBodyCount);
	dfree(p->sb.cslot);
	dfree(p->sb.recN);
	dfree(p->sb.recF);
	dfree(p->sb.lamN);
	dfree(p->sb.lamF);
	dfree(p->sb.posOf);
	dfree(p->sb.levelOrder);
	dfree(p->sb.levelStart);
	dfree(p->sb.partMeta);
	dfree(p->sb.partBodies);
	dfree(p->dSceneOf);
	dfree(p->dSceneCounts);
	dfree(p->sb.cellLevels);
	dfree(p->sb.lastClass);
	dfree(p->sb.residSnap);
	dfree(p->sb.residL2);
	b3hWarmBuffers& w = p->wb;
	dfree(w.htKeys);
	dfree(w.htVals);
	dfree(w.hist);
	dfree(w.histKeys);
	dfree(w.histValid);
	dfree(w.prev);
	dfree(w.warm);
	dfree(w.X);
	dfree(w.Y);
	dfree(w.T);
	dfree(w.rowCount);
	dfree(w.rowScan);
	dfree(w.scanTmp);
	dfree(w.logCount);
	ray_free(p);
	joint_free(p);
}

/* warm-start buffers, allocated on first use (features: the per-point rows of
 * the learned mode and of contact logging) */
static int warm_alloc(b3hPipeline* p, bool features)
{
	b3hWarmBuffers& w = p->wb;
	size_t mc = (size_t)p->cfg.m_maxContactCapacity;
	int r = 0;
	if (!p->warmAlloc)
	{
		int cap = 1;
		while ((size_t)cap < 2 * mc) cap *= 2;
		w.capC = (int)mc;
		w.htCap = cap;
		if ((r = dalloc(&w.htKeys, (size_t)cap)) || (r = dalloc(&w.htVals, (size_t)cap)) || (r = dalloc(&w.hist, mc)) ||
			(r = dalloc(&w.histKeys, mc)) || (r = dalloc(&w.histValid, 1)) || (r = dalloc(&w.prev, 8 * mc)) || (r = dalloc(&w.warm, 6 * mc)) ||
			(r = dalloc(&w.logCount, 1)))
			return r;
		B3H_CHECK(hipMemset(w.histValid, 0, sizeof(int)));
		B3H_CHECK(hipMemset(w.logCount, 0, sizeof(int)));
		B3H_CHECK(hipMemset(w.htKeys, 0xff, sizeof(unsigned long long) * cap));
		p->warmAlloc = true;
	}
	if (features && !p->warmFeatAlloc)
	{
		if ((r = dalloc(&w.X, 4 * mc * B3H_WARM_FEATURES)) || (r = dalloc(&w.Y, 4 * mc * B3H_WARM_OUTPUTS)) ||
			(r = dalloc(&w.T, 4 * mc * B3H_WARM_OUTPUTS)) || (r = dalloc(&w.rowCount, mc + 1)) ||
			(r = dalloc(&w.rowScan, mc + 2)) || (r = dalloc(&w.scanTmp, b3h_scan_tmp_size((int)mc))))
			return r;
		p->warmFeatAlloc = true;
	}
	return B3H_OK;
}
static void drop_graph(b3hPipeline* p)
{
	if (p->graphExec)
	{
		hipGraphExecDestroy(p->graphExec);
		p->graphExec = nullptr;
	}
}

extern "C" int b3h_pipeline_create(b3hNarrowPhase* np, b3hBroadphase* bp, const b3hConfig* cfg, void* stream,
								   b3hPipeline** out)
{
	CHECK_ARG(np && bp && cfg && out);
	*out = nullptr;
	b3hPipeline* p = new b3hPipeline();
	memset((void*)&p->sb, 0, sizeof(p->sb));
	p->np = np;
	p->bp = bp;
	p->cfg = *cfg;
	p->gravity = f4(0.f, -9.8f, 0.f, 0.f);
	b3h_solver_params_default(&p->prm);
	p->trackResidual = false;
	p->timing = false;
	p->evValid = false;
	p->graph = nullptr;
	p->graphExec = nullptr;
	p->graphDt = 0.f;
	p->graphBodies = -1;
	p->dWarm = nullptr;
	p->warmCount = 0;
	memset((void*)&p->wb, 0, sizeof(p->wb));
	p->warmAlloc = p->warmFeatAlloc = false;
	p->warmNet = nullptr;
	p->logRows = false;
	p->matchTol = 0.1f;
	p->jointUid = 0;
	p->jointsDirty = false;
	memset((void*)&p->jb, 0, sizeof(p->jb));
	/* b3GpuPgsConstraintSolver::solveJoints (b3GpuPgsConstraintSolver.cpp:927-937)
	 * with b3ContactSolverInfo's defaults (b3ContactSolverInfo.h:60-80) */
	p->jp.timeStep = 1.f / 60.f;
	p->jp.erp = 0.2f;
	p->jp.globalCfm = 0.f;
	p->jp.damping = 1.f;
	p->jp.iterations = 4;
	memset((void*)&p->raySc, 0, sizeof(p->raySc));
	p->raySmallCap = -1;
	p->dRays = nullptr;
	p->dHits = nullptr;
	p->rayCap = 0;
	int r;
	/* one stream for all three objects keeps the step in order: while the
	 * pipeline exists the broadphase issues its stand-alone calls on the
	 * narrowphase's stream too, so they are ordered after the pipeline's
	 * AABB update */
	p->stream = np->stream;
	p->ownStream = false;
	(void)stream;
	hipStreamSynchronize(bp->stream);
	bp->stream = np->stream;
	if ((r = b3h_solver_module_init()))
	{
		delete p;
		return r;
	}
	int mc = cfg->m_maxContactCapacity;
	p->sb.maxContacts = mc;
	p->sb.pgsBodyCap = -1;
	p->sb.numContacts = np->dNumContacts;
	if ((r = dalloc(&p->sb.cellOf, mc)) || (r = dalloc(&p->sb.cellCounts, B3H_N_CELLS)) ||
		(r = dalloc(&p->sb.cellOffsets, B3H_N_CELLS + 1)) || (r = dalloc(&p->sb.cellCursor, B3H_N_CELLS)) ||
		(r = dalloc(&p->sb.cellSlots, mc)) || (r = dalloc(&p->sb.perm, mc)) || (r = dalloc(&p->sb.batchIdx, mc)) ||
		(r = dalloc(&p->sb.batchSizes, B3H_N_CELLS)) || (r = dalloc(&p->sb.sortKeys, 4 * (size_t)mc + 64)) ||
		(r = dalloc(&p->sb.constraints, mc)) || (r = dalloc(&p->sb.sortedContacts, mc)) || (r = dalloc(&p->sb.resid, 4 * B3H_MAX_ITERATIONS)) ||
		(r = dalloc(&p->sb.cellBodies, 2 * (size_t)mc + 64)) || (r = dalloc(&p->sb.cellBodyCount, B3H_N_CELLS)) ||
		(r = dalloc(&p->sb.cslot, 2 * (size_t)mc + 2)) || (r = dalloc(&p->sb.recN, 16 * (size_t)mc)) ||
		(r = dalloc(&p->sb.recF, 11 * (size_t)mc)) || (r = dalloc(&p->sb.lamN, mc)) ||
		(r = dalloc(&p->sb.lamF, mc)) || (r = dalloc(&p->sb.posOf, mc)) || (r = dalloc(&p->sb.levelOrder, mc)) ||
		(r = dalloc(&p->sb.levelStart, (size_t)mc + B3H_N_CELLS * B3H_PGS_PARTS + 8)) ||
		(r = dalloc(&p->sb.partMeta, 2 * B3H_N_CELLS * B3H_PGS_PARTS)) ||
		(r = dalloc(&p->sb.partBodies, 2 * (2 * (size_t)mc + 64))) || (r = dalloc(&p->sb.cellLevels, B3H_N_CELLS)) ||
		(r = dalloc(&p->sb.lastClass, (size_t)cfg->m_maxConvexBodies + 16)) ||
		(r = dalloc(&p->sb.residSnap, 6 * (size_t)mc)) || (r = dalloc(&p->sb.residL2, 4 * B3H_MAX_ITERATIONS)))
	{
		pl_free(p);
		delete p;
		return r;
	}
	B3H_CHECK(hipMemset(p->sb.cellCounts, 0, B3H_N_CELLS * sizeof(int)));
	B3H_CHECK(hipMemset(p->sb.cellOffsets, 0, (B3H_N_CELLS + 1) * sizeof(int)));
	B3H_CHECK(hipMemset(p->sb.batchSizes, 0, B3H_N_CELLS * sizeof(int)));
	*out = p;
	return B3H_OK;
}

extern "C" int b3h_pipeline_destroy(b3hPipeline* p)
{
	if (!p) return B3H_OK;
	hipStreamSynchronize(p->stream);
	p->bp->stream = p->bp->homeStream;
	if (p->graphExec) hipGraphExecDestroy(p->graphExec);
	if (p->graph) hipGraphDestroy(p->graph);
	if (p->evValid)
		for (int i = 0; i <= B3H_NUM_STAGES; i++) hipEventDestroy(p->ev[i]);
	pl_free(p);
	delete p;
	return B3H_OK;
}

static int num_bodies(b3hPipeline* p) { return (int)p->np->bodies.size(); }
static int joints_upload(b3hPipeline* p);
static int joints_pull_flags(b3hPipeline* p);

static void pl_aabbs(b3hPipeline* p)
{
	b3h_launch_world_aabbs(num_bodies(p), p->np->dBodies, p->np->dCols, p->np->dLocalAabbs, p->bp->dAll, p->stream);
}

static void pl_solve(b3hPipeline* p, hipEvent_t evSetupDone = nullptr, hipEvent_t evNormalDone = nullptr,
					 const b3hIntegrateParams* integrate = nullptr)
{
	hipStream_t s = p->stream;
	int mc = p->cfg.m_maxContactCapacity;
	b3h_solve_setup(num_bodies(p), p->np->dBodies, p->np->dInertias, p->np->dNumContacts, mc, p->np->dStaging,
					p->np->dContactPair, p->prm, p->np->static0Index, p->sb, true, s);
	int mode = p->prm.m_warmStartMode;
	bool warmOn = (mode >= 1 || p->logRows) && p->warmAlloc;
	b3hWarmParams wp;
	float g = sqrtf(dot3(p->gravity, p->gravity)) * p->prm.m_solverDt;
	wp.velUnit = g > 0.f ? g : 1.f;
	wp.matchTol = p->matchTol;
	wp.mode = mode;
	if (warmOn)
		b3h_warm_prepare(p->np->dNumContacts, mc, p->sb.constraints, p->np->dBodies, p->wb, wp,
						 mode == 2 ? p->warmNet : nullptr, p->logRows, s);
	p->sb.warmImpulses = p->dWarm ? p->dWarm : ((warmOn && mode >= 1) ? p->wb.warm : nullptr);
	if (evSetupDone) hipEventRecord(evSetupDone, s);
	b3h_solve_sweeps(p->np->dBodies, p->np->dInertias, p->prm, p->sb, p->trackResidual, s, evNormalDone, integrate);
	if (warmOn) b3h_warm_finish(p->np->dNumContacts, mc, p->sb.constraints, p->wb, wp, p->logRows, s);
}

static int pl_step(b3hPipeline* p, float dt, bool events)
{
	hipStream_t s = p->stream;
	/* stage events: 0 start | aabb | 1 | broadphase | 2 | narrowphase MPR | 3 |
	 * narrowphase SAT and clipping | 4 | narrowphase compaction | 5 | solver
	 * set-up | 6 | normal sweeps | 7 | friction sweeps | 8 | integrate | 9 */
	if (events) hipEventRecord(p->ev[0], s);
	pl_aabbs(p);
	if (events) hipEventRecord(p->ev[1], s);
	p->bp->sc.sceneOf = p->numScenes > 0 ? p->dSceneOf : nullptr;
	int r = bp_calc_pairs(p->bp, p->cfg.m_maxBroadphasePairs, s);
	if (r) return r;
	if (events) hipEventRecord(p->ev[2], s);
	if ((r = np_compute_contacts_dev(p->np, p->bp->dPairs, p->bp->dPairCount, p->cfg.m_maxBroadphasePairs,
									 events ? p->ev[4] : nullptr, events ? p->ev[3] : nullptr)))
		return r;
	if (events) hipEventRecord(p->ev[5], s);
	/* joints before the contact solve (b3GpuRigidBodyPipeline.cpp:361-372) */
	if (p->jb.numConstraints > 0) b3h_solve_joints(p->jb, p->np->dBodies, p->np->dInertias, p->jp, s);
	/* integration is fused into the last friction sweep for every body a
	 * contact touches; the rest are integrated by the pass below */
	b3hIntegrateParams ip = {dt, p->prm.m_angularDamping, p->gravity};
	bool fused = p->prm.m_numIterations > 0;
	pl_solve(p, events ? p->ev[6] : nullptr, events ? p->ev[7] : nullptr, fused ? &ip : nullptr);
	if (events) hipEventRecord(p->ev[8], s);
	b3h_launch_integrate(num_bodies(p), p->np->dBodies, dt, p->prm.m_angularDamping, p->gravity, s,
						 fused ? p->sb.lastClass : nullptr);
	if (events) hipEventRecord(p->ev[9], s);
	return B3H_OK;
}

static int check_scenes(b3hPipeline* p)
{
	if (p->numScenes > 0 && p->sceneBodies != num_bodies(p))
	{
		b3h_set_error("stepSimulation: scenes cover %d bodies, the pipeline has %d (set the scenes again)",
					  p->sceneBodies, num_bodies(p));
		return B3H_ERR_STATE;
	}
	return B3H_OK;
}

extern "C" int b3h_pipeline_set_scenes(b3hPipeline* p, int numScenes, const int* bodyOffsets)
{
	CHECK_ARG(p && numScenes >= 0 && (numScenes == 0 || bodyOffsets));
	int nb = num_bodies(p);
	B3H_CHECK(hipStreamSynchronize(p->stream));
	if (numScenes == 0)
	{
		p->numScenes = 0;
		p->bp->sc.sceneOf = nullptr;
		drop_graph(p);
		return B3H_OK;
	}
	if (bodyOffsets[0] != 0 || bodyOffsets[numScenes] != nb)
	{
		b3h_set_error("setScenes: offsets must run from 0 to the body count %d (got %d .. %d)", nb, bodyOffsets[0],
					  bodyOffsets[numScenes]);
		return B3H_ERR_INVALID;
	}
	std::vector<int> sceneOf((size_t)nb);
	for (int k = 0; k < numScenes; k++)
	{
		if (bodyOffsets[k + 1] < bodyOffsets[k])
		{
			b3h_set_error("setScenes: offsets decrease at scene %d", k);
			return B3H_ERR_INVALID;
		}
		for (int b = bodyOffsets[k]; b < bodyOffsets[k + 1]; b++) sceneOf[(size_t)b] = k;
	}
	dfree(p->dSceneOf);
	dfree(p->dSceneCounts);
	p->dSceneOf = nullptr;
	p->dSceneCounts = nullptr;
	int r;
	if ((r = dalloc(&p->dSceneOf, (size_t)(nb > 0 ? nb : 1))) || (r = dalloc(&p->dSceneCounts, (size_t)numScenes)))
		return r;
	if (nb > 0) B3H_CHECK(hipMemcpy(p->dSceneOf, sceneOf.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
	p->numScenes = numScenes;
	p->sceneBodies = nb;
	p->bp->sc.sceneOf = p->dSceneOf;
	drop_graph(p);
	return B3H_OK;
}

extern "C" int b3h_pipeline_get_scene_contact_counts(b3hPipeline* p, int* counts, int numScenes)
{
	CHECK_ARG(p && counts && numScenes == p->numScenes && numScenes > 0);
	b3h_scene_contact_counts(p->bp->dPairCount, p->cfg.m_maxBroadphasePairs, p->bp->dPairs, p->dSceneOf,
							 p->dSceneCounts, numScenes, p->stream);
	B3H_CHECK(hipGetLastError());
	B3H_CHECK(hipMemcpyAsync(counts, p->dSceneCounts, sizeof(int) * numScenes, hipMemcpyDeviceToHost, p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	return B3H_OK;
}

extern "C" int b3h_pipeline_step_simulation(b3hPipeline* p, float dt)
{
	CHECK_ARG(p);
	if (num_bodies(p) == 0) return B3H_OK;
	if (int rs = check_scenes(p)) return rs;
	if (p->bp->all.size() != p->np->bodies.size())
	{
		b3h_set_error("stepSimulation: %zu proxies for %zu bodies (register bodies through registerPhysicsInstance)",
					  p->bp->all.size(), p->np->bodies.size());
		return B3H_ERR_STATE;
	}
	if (p->timing && !p->evValid)
	{
		for (int i = 0; i <= B3H_NUM_STAGES; i++) B3H_CHECK(hipEventCreate(&p->ev[i]));
		p->evValid = true;
	}
	int r = joints_upload(p);
	if (r) return r;
	if ((r = np_ensure_sat(p->np))) return r;
	r = pl_step(p, dt, p->timing);
	if (r) return r;
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}

extern "C" int b3h_pipeline_step_graph(b3hPipeline* p, float dt, int count)
{
	CHECK_ARG(p && count >= 0);
	if (num_bodies(p) == 0) return B3H_OK;
	if (int rs = check_scenes(p)) return rs;
	if (int ru = joints_upload(p)) return ru;
	if (!p->graphExec || p->graphDt != dt || p->graphBodies != num_bodies(p) || p->graphNpGen != p->np->gen ||
		p->graphBpGen != p->bp->gen)
	{
		if (p->graphExec) hipGraphExecDestroy(p->graphExec);
		if (p->graph) hipGraphDestroy(p->graph);
		p->graphExec = nullptr;
		p->graph = nullptr;
		/* scratch sizes and one-time module setup must exist before capture */
		int r = bp_ensure_scratch(p->bp);
		if (r) return r;
		if ((r = np_ensure_sat(p->np))) return r;
		B3H_CHECK(hipStreamSynchronize(p->stream));
		B3H_CHECK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
		r = pl_step(p, dt, false);
		hipGraph_t g = nullptr;
		hipError_t e = hipStreamEndCapture(p->stream, &g);
		if (r) return r;
		if (e != hipSuccess)
		{
			b3h_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
			return B3H_ERR_HIP;
		}
		p->graph = g;
		B3H_CHECK(hipGraphInstantiate(&p->graphExec, g, nullptr, nullptr, 0));
		p->graphDt = dt;
		p->graphBodies = num_bodies(p);
		p->graphNpGen = p->np->gen;
		p->graphBpGen = p->bp->gen;
	}
	for (int i = 0; i < count; i++) B3H_CHECK(hipGraphLaunch(p->graphExec, p->stream));
	return B3H_OK;
}

extern "C" int b3h_pipeline_integrate(b3hPipeline* p, float dt)
{
	CHECK_ARG(p);
	b3h_launch_integrate(num_bodies(p), p->np->dBodies, dt, p->prm.m_angularDamping, p->gravity, p->stream);
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}
extern "C" int b3h_pipeline_setup_gpu_aabbs_full(b3hPipeline* p)
{
	CHECK_ARG(p);
	pl_aabbs(p);
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}
extern "C" int b3h_pipeline_solve_contacts(b3hPipeline* p)
{
	CHECK_ARG(p);
	pl_solve(p);
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}

/* b3TransformAabb with the b3Transform of (position, orientation)
 * (b3AabbUtil.h:182-196, b3Transform.h:90-93) */
static void transform_aabb(const b3hAabb& la, float margin, const float* pos, const float* orn, float* mn, float* mx)
{
	float m[3][3];
	b3h_quat_to_basis(orn, m);
	float he[3], lc[3];
	const float lmin[3] = {la.m_min.x, la.m_min.y, la.m_min.z};
	const float lmax[3] = {la.m_max.x, la.m_max.y, la.m_max.z};
	for (int k = 0; k < 3; k++)
	{
		he[k] = 0.5f * (lmax[k] - lmin[k]);
		he[k] += margin;
		lc[k] = 0.5f * (lmax[k] + lmin[k]);
	}
	for (int r = 0; r < 3; r++)
	{
		float c = lc[0] * m[r][0] + lc[1] * m[r][1] + lc[2] * m[r][2];
		c = c + pos[r];
		float e = he[0] * std::fabs(m[r][0]) + he[1] * std::fabs(m[r][1]) + he[2] * std::fabs(m[r][2]);
		mn[r] = c - e;
		mx[r] = c + e;
	}
}

extern "C" int b3h_pipeline_register_physics_instance(b3hPipeline* p, float mass, const float* pos, const float* orn,
													  int collidableIndex, int userIndex, int writeToGpu, int* outBody)
{
	CHECK_ARG(p && pos && orn && outBody);
	(void)userIndex;
	*outBody = -1;
	if (collidableIndex < 0 || collidableIndex >= (int)p->np->localAabbs.size())
	{
		b3h_set_error("registerPhysicsInstance using invalid collidableIndex %d", collidableIndex);
		return B3H_ERR_INVALID;
	}
	float mn[3], mx[3];
	transform_aabb(p->np->localAabbs[collidableIndex], 0.01f, pos, orn, mn, mx);
	int body = -1;
	int r = b3h_np_register_rigid_body(p->np, collidableIndex, mass, pos, orn, mn, mx, 0, &body);
	if (r) return r;
	if (mass)
		r = b3h_bp_create_proxy(p->bp, mn, mx, body, 1, 1);
	else
		r = b3h_bp_create_large_proxy(p->bp, mn, mx, body, 1, 1);
	if (r) return r;
	if (writeToGpu)
	{
		if ((r = b3h_np_write_all_bodies_to_gpu(p->np)) || (r = b3h_bp_write_aabbs_to_gpu(p->bp))) return r;
	}
	*outBody = body;
	return B3H_OK;
}

extern "C" int b3h_compute_body_registration(const float* localAabbMin3, const float* localAabbMax3, float mass,
											 const float* pos3, const float* orn4, float* aabbMin3, float* aabbMax3,
											 b3hInertiaData* inertia)
{
	CHECK_ARG(localAabbMin3 && localAabbMax3 && pos3 && orn4 && aabbMin3 && aabbMax3 && inertia);
	b3hAabb la;
	la.m_min = f4(localAabbMin3[0], localAabbMin3[1], localAabbMin3[2], 0.f);
	la.m_max = f4(localAabbMax3[0], localAabbMax3[1], localAabbMax3[2], 0.f);
	transform_aabb(la, 0.01f, pos3, orn4, aabbMin3, aabbMax3);
	body_inertia(mass, aabbMin3, aabbMax3, orn4, inertia);
	return B3H_OK;
}

extern "C" int b3h_pipeline_register_convex_polyhedron(b3hPipeline* p, const b3hConvexUtility* u, int* outCollidable)
{
	CHECK_ARG(p);
	return b3h_np_register_convex_polyhedron(p->np, u, outCollidable);
}

extern "C" int b3h_pipeline_write_all_instances_to_gpu(b3hPipeline* p)
{
	CHECK_ARG(p);
	int r;
	if ((r = b3h_np_write_all_bodies_to_gpu(p->np)) || (r = b3h_bp_write_aabbs_to_gpu(p->bp))) return r;
	if ((r = joints_pull_flags(p))) return r;
	p->jointsDirty = p->jointsDirty || !p->joints.empty();
	return joints_upload(p);
}
extern "C" int b3h_pipeline_set_gravity(b3hPipeline* p, const float* g)
{
	CHECK_ARG(p && g);
	drop_graph(p);
	p->gravity = f4(g[0], g[1], g[2], 0.f);
	if (p->graphExec)
	{
		hipGraphExecDestroy(p->graphExec);
		p->graphExec = nullptr;
	}
	return B3H_OK;
}
extern "C" int b3h_pipeline_reset(b3hPipeline* p)
{
	CHECK_ARG(p);
	drop_graph(p);
	p->bp->all.clear();
	p->bp->smallMap.clear();
	p->bp->largeMap.clear();
	p->joints.clear();
	p->jointsDirty = true;
	if (p->warmAlloc) B3H_CHECK(hipMemsetAsync(p->wb.histValid, 0, sizeof(int), p->stream));
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_body_buffer(b3hPipeline* p, b3hRigidBodyData** dev)
{
	CHECK_ARG(p && dev);
	*dev = p->np->dBodies;
	return B3H_OK;
}
/* BVH scratch of the ray caster for the broadphase's current small proxies */
static int ray_ensure(b3hPipeline* p)
{
	const int ns = (int)p->bp->smallMap.size();
	if (ns <= p->raySmallCap && p->raySc.bvhMeta) return B3H_OK;
	b3hBroadphaseScratch& sc = p->raySc;
	dfree(sc.keys);
	dfree(sc.vals);
	dfree(sc.tmpKeys);
	dfree(sc.tmpVals);
	dfree(sc.hist);
	dfree(sc.sortedAabbs);
	dfree(sc.bvhMerge);
	dfree(sc.bvhChild);
	dfree(sc.bvhBox);
	dfree(sc.bvhFlag);
	dfree(sc.bvhMeta);
	p->raySmallCap = -1;
	const int cap = ns + 64;
	int r;
	if ((r = dalloc(&sc.keys, cap)) || (r = dalloc(&sc.vals, cap)) || (r = dalloc(&sc.tmpKeys, cap)) ||
		(r = dalloc(&sc.tmpVals, cap)) || (r = dalloc(&sc.hist, b3h_radix_hist_size(cap))) ||
		(r = dalloc(&sc.sortedAabbs, cap)) || (r = dalloc(&sc.bvhMerge, b3h_plbvh_reduce_blocks(cap) + 1)) ||
		(r = dalloc(&sc.bvhChild, cap)) || (r = dalloc(&sc.bvhBox, 2 * (size_t)cap)) ||
		(r = dalloc(&sc.bvhFlag, cap)) || (r = dalloc(&sc.bvhMeta, 2)))
		return r;
	B3H_CHECK(hipMemset(sc.bvhMeta, 0, 2 * sizeof(int)));
	p->raySmallCap = ns;
	return B3H_OK;
}

extern "C" int b3h_pipeline_cast_rays_gpu(b3hPipeline* p, const b3hRayInfo* rays, b3hRayHit* hits, int numRays)
{
	CHECK_ARG(p && numRays >= 0);
	if (numRays == 0) return B3H_OK;
	CHECK_ARG(rays && hits);
	int r = ray_ensure(p);
	if (r) return r;
	b3hBroadphase* bp = p->bp;
	b3h_cast_rays(numRays, rays, hits, bp->dAll, bp->dSmall, (int)bp->smallMap.size(), bp->dLarge,
				  (int)bp->largeMap.size(), p->np->dBodies, np_tables(p->np), p->raySc, p->stream);
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}

extern "C" int b3h_pipeline_cast_rays(b3hPipeline* p, const b3hRayInfo* rays, b3hRayHit* hits, int numRays)
{
	CHECK_ARG(p && numRays >= 0);
	if (numRays == 0) return B3H_OK;
	CHECK_ARG(rays && hits);
	if (numRays > p->rayCap)
	{
		dfree(p->dRays);
		dfree(p->dHits);
		p->rayCap = 0;
		int r;
		if ((r = dalloc(&p->dRays, numRays)) || (r = dalloc(&p->dHits, numRays))) return r;
		p->rayCap = numRays;
	}
	B3H_CHECK(hipMemcpyAsync(p->dRays, rays, sizeof(b3hRayInfo) * numRays, hipMemcpyHostToDevice, p->stream));
	B3H_CHECK(hipMemcpyAsync(p->dHits, hits, sizeof(b3hRayHit) * numRays, hipMemcpyHostToDevice, p->stream));
	int r = b3h_pipeline_cast_rays_gpu(p, p->dRays, p->dHits, numRays);
	if (r) return r;
	B3H_CHECK(hipMemcpyAsync(hits, p->dHits, sizeof(b3hRayHit) * numRays, hipMemcpyDeviceToHost, p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	return B3H_OK;
}

extern "C" int b3h_pipeline_get_ray_bvh_info(b3hPipeline* p, int* root, int* overflow)
{
	CHECK_ARG(p && root && overflow);
	if (!p->raySc.bvhMeta)
	{
		b3h_set_error("no ray cast yet");
		return B3H_ERR_STATE;
	}
	int m[2];
	B3H_CHECK(hipMemcpyAsync(m, p->raySc.bvhMeta, sizeof(m), hipMemcpyDeviceToHost, p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	*root = m[0];
	*overflow = m[1];
	return B3H_OK;
}

/* ---- joints ---- */
/* the device's enabled flags (cleared by breaking) into the host list */
static int joints_pull_flags(b3hPipeline* p)
{
	if (p->jointsDirty || p->jb.numConstraints == 0 || (int)p->joints.size() != p->jb.numConstraints) return B3H_OK;
	std::vector<b3hGpuGenericConstraint> dev((size_t)p->jb.numConstraints);
	B3H_CHECK(hipMemcpyAsync(dev.data(), p->jb.constraints, dev.size() * sizeof(b3hGpuGenericConstraint),
							 hipMemcpyDeviceToHost, p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	for (size_t i = 0; i < dev.size(); i++) p->joints[i].m_flags = dev[i].m_flags;
	return B3H_OK;
}

/* Greedy joint batches on the host, as b3GpuPgsConstraintSolver::
 * sortConstraintByBatch3 (b3GpuPgsConstraintSolver.cpp:862-948) builds them
 * with simdWidth = numConstraints + 1 (never reached, so the used-body map is
 * cleared only between batches) and staticIdx -1: pass after pass over the
 * constraints not yet batched, in their current order, a constraint joins
 * the batch when neither of its non-static bodies is used yet, and is swapped
 * to the front.  Static means a negative signed index (invMass 0,
 * initBatchConstraintsKernel, jointSolver.cl:400-416), so body 0 counts as
 * dynamic even when static (-0 == 0).  order receives the original constraint
 * index per position, sizes the batch sizes. */
static void joints_batch(const std::vector<b3hGpuGenericConstraint>& cs, const std::vector<b3hRigidBodyData>& bodies,
						 std::vector<int>& order, std::vector<int>& sizes)
{
	struct BC
	{
		int a, b, orig;
	};
	const int n = (int)cs.size();
	std::vector<BC> bc((size_t)n);
	for (int i = 0; i < n; i++)
	{
		int ra = cs[i].m_rbA, rb = cs[i].m_rbB;
		bc[i].a = bodies[ra].m_invMass != 0.f ? ra : -ra;
		bc[i].b = bodies[rb].m_invMass != 0.f ? rb : -rb;
		bc[i].orig = i;
	}
	const int staticIdx = -1;
	std::vector<unsigned> used(bodies.size() / 32 + 1, 0u);
	std::vector<int> cur;
	sizes.clear();
	int valid = 0;
	while (valid < n)
	{
		int inBatch = 0;
		for (int b : cur) used[b / 32] = 0;
		cur.clear();
		for (int i = valid; i < n; i++)
		{
			const int as = bc[i].a, bs = bc[i].b;
			const int a = as < 0 ? -as : as, b = bs < 0 ? -bs : bs;
			const bool aStatic = as < 0 || as == staticIdx, bStatic = bs < 0 || bs == staticIdx;
			unsigned aUn = 0, bUn = 0;
			if (!aStatic) aUn = used[a / 32] & (1u << (a & 31));
			if (!aUn && !bStatic) bUn = used[b / 32] & (1u << (b & 31));
			if (aUn == 0 && bUn == 0)
			{
				if (!aStatic)
				{
					used[a / 32] |= 1u << (a & 31);
					cur.push_back(a);
				}
				if (!bStatic)
				{
					used[b / 32] |= 1u << (b & 31);
					cur.push_back(b);
				}
				if (i != valid) std::swap(bc[i], bc[valid]);
				valid++;
				inBatch++;
			}
		}
		sizes.push_back(inBatch);
	}
	order.resize((size_t)n);
	for (int i = 0; i < n; i++) order[i] = bc[i].orig;
}

/* upload the host list and rebuild rows, slots and batches (the reference
 * uploads at writeAllInstancesToGpu, b3GpuRigidBodyPipeline.cpp:597-601, and
 * re-batches after recomputeBatches) */
static int joints_upload(b3hPipeline* p)
{
	if (!p->jointsDirty) return B3H_OK;
	drop_graph(p);
	joint_free(p);
	b3hJointBuffers& jb = p->jb;
	const int n = (int)p->joints.size();
	const int nb = num_bodies(p);
	p->jointsDirty = false;
	if (n == 0) return B3H_OK;
	std::vector<int> rowOffset(n), rowCount(n), slotOf(2 * (size_t)n), bodySlot((size_t)nb, -1), order, sizes;
	int rows = 0, slots = 0;
	for (int i = 0; i < n; i++)
	{
		const b3hGpuGenericConstraint& c = p->joints[i];
		rowCount[i] = c.m_constraintType == B3H_POINT2POINT_CONSTRAINT_TYPE ? 3 : (c.m_constraintType == B3H_FIXED_CONSTRAINT_TYPE ? 6 : 0);
		rowOffset[i] = rows;
		rows += rowCount[i];
		const int ab[2] = {c.m_rbA, c.m_rbB};
		for (int k = 0; k < 2; k++)
		{
			if (bodySlot[ab[k]] < 0) bodySlot[ab[k]] = slots++;
			slotOf[2 * i + k] = bodySlot[ab[k]];
		}
	}
	joints_batch(p->joints, p->np->bodies, order, sizes);
	std::vector<int> start(sizes.size() + 1, 0);
	for (size_t k = 0; k < sizes.size(); k++) start[k + 1] = start[k] + sizes[k];
	int r;
	if ((r = dalloc(&jb.constraints, n)) || (r = dalloc(&jb.rowOffset, n)) || (r = dalloc(&jb.rowCount, n)) ||
		(r = dalloc(&jb.slotOf, 2 * (size_t)n)) || (r = dalloc(&jb.rows, rows > 0 ? rows : 1)) ||
		(r = dalloc(&jb.batchStart, start.size())) || (r = dalloc(&jb.order, n)) || (r = dalloc(&jb.delta, 2 * (size_t)slots)) ||
		(r = dalloc(&jb.bodySlot, nb > 0 ? nb : 1)))
	{
		joint_free(p);
		p->jointsDirty = true;
		return r;
	}
	hipStream_t s = p->stream;
	B3H_CHECK(hipMemcpyAsync(jb.constraints, p->joints.data(), n * sizeof(b3hGpuGenericConstraint), hipMemcpyHostToDevice, s));
	B3H_CHECK(hipMemcpyAsync(jb.rowOffset, rowOffset.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
	B3H_CHECK(hipMemcpyAsync(jb.rowCount, rowCount.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
	B3H_CHECK(hipMemcpyAsync(jb.slotOf, slotOf.data(), 2 * n * sizeof(int), hipMemcpyHostToDevice, s));
	B3H_CHECK(hipMemcpyAsync(jb.batchStart, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, s));
	B3H_CHECK(hipMemcpyAsync(jb.order, order.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
	if (nb > 0) B3H_CHECK(hipMemcpyAsync(jb.bodySlot, bodySlot.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
	B3H_CHECK(hipStreamSynchronize(s));
	jb.numConstraints = n;
	jb.numBatches = (int)sizes.size();
	jb.numSlots = slots;
	jb.numBodies = nb;
	return B3H_OK;
}

static int joint_check_bodies(b3hPipeline* p, int a, int b)
{
	const int nb = num_bodies(p);
	if (a < 0 || a >= nb || b < 0 || b >= nb)
	{
		b3h_set_error("create constraint: body %d or %d out of range (%d bodies)", a, b, nb);
		return B3H_ERR_INVALID;
	}
	return B3H_OK;
}

extern "C" int b3h_pipeline_create_point2point_constraint(b3hPipeline* p, int bodyA, int bodyB, const float* pivotInA3,
														   const float* pivotInB3, float breakingThreshold, int* uidOut)
{
	CHECK_ARG(p && pivotInA3 && pivotInB3 && uidOut);
	int r = joint_check_bodies(p, bodyA, bodyB);
	if (r || (r = joints_pull_flags(p))) return r;
	b3hGpuGenericConstraint c;
	memset(&c, 0, sizeof(c));
	c.m_uid = p->jointUid++;
	c.m_flags = B3H_CONSTRAINT_FLAG_ENABLED;
	c.m_rbA = bodyA;
	c.m_rbB = bodyB;
	c.m_pivotInA = f4(pivotInA3[0], pivotInA3[1], pivotInA3[2], 0.f);
	c.m_pivotInB = f4(pivotInB3[0], pivotInB3[1], pivotInB3[2], 0.f);
	c.m_relTargetAB = f4(0.f, 0.f, 0.f, 1.f);
	c.m_breakingImpulseThreshold = breakingThreshold;
	c.m_constraintType = B3H_POINT2POINT_CONSTRAINT_TYPE;
	p->joints.push_back(c);
	p->jointsDirty = true;
	drop_graph(p);
	*uidOut = c.m_uid;
	return B3H_OK;
}

extern "C" int b3h_pipeline_create_fixed_constraint(b3hPipeline* p, int bodyA, int bodyB, const float* pivotInA3,
													 const float* pivotInB3, const float* relTargetAB4,
													 float breakingThreshold, int* uidOut)
{
	CHECK_ARG(p && pivotInA3 && pivotInB3 && relTargetAB4 && uidOut);
	int r = joint_check_bodies(p, bodyA, bodyB);
	if (r || (r = joints_pull_flags(p))) return r;
	b3hGpuGenericConstraint c;
	memset(&c, 0, sizeof(c));
	c.m_uid = p->jointUid++;
	c.m_flags = B3H_CONSTRAINT_FLAG_ENABLED;
	c.m_rbA = bodyA;
	c.m_rbB = bodyB;
	c.m_pivotInA = f4(pivotInA3[0], pivotInA3[1], pivotInA3[2], 0.f);
	c.m_pivotInB = f4(pivotInB3[0], pivotInB3[1], pivotInB3[2], 0.f);
	c.m_relTargetAB = f4(relTargetAB4[0], relTargetAB4[1], relTargetAB4[2], relTargetAB4[3]);
	c.m_breakingImpulseThreshold = breakingThreshold;
	c.m_constraintType = B3H_FIXED_CONSTRAINT_TYPE;
	p->joints.push_back(c);
	p->jointsDirty = true;
	drop_graph(p);
	*uidOut = c.m_uid;
	return B3H_OK;
}

extern "C" int b3h_pipeline_remove_constraint_by_uid(b3hPipeline* p, int uid)
{
	CHECK_ARG(p);
	int r = joints_pull_flags(p);
	if (r) return r;
	for (size_t i = 0; i < p->joints.size(); i++)
		if (p->joints[i].m_uid == uid)
		{
			std::swap(p->joints[i], p->joints.back());
			p->joints.pop_back();
			p->jointsDirty = true;
			drop_graph(p);
			return B3H_OK;
		}
	b3h_set_error("removeConstraintByUid: no constraint with uid %d", uid);
	return B3H_ERR_INVALID;
}

extern "C" int b3h_pipeline_copy_constraints_to_host(b3hPipeline* p) { CHECK_ARG(p); return joints_pull_flags(p); }

extern "C" int b3h_pipeline_get_constraints(b3hPipeline* p, b3hGpuGenericConstraint* host, int capacity, int* n)
{
	CHECK_ARG(p && n);
	int r = joints_pull_flags(p);
	if (r) return r;
	*n = (int)p->joints.size();
	if (!host) return B3H_OK;
	if (capacity < *n)
	{
		b3h_set_error("getConstraints: capacity %d < %d constraints", capacity, *n);
		return B3H_ERR_CAPACITY;
	}
	if (*n) memcpy(host, p->joints.data(), (size_t)*n * sizeof(b3hGpuGenericConstraint));
	return B3H_OK;
}

extern "C" int b3h_pipeline_get_num_bodies(b3hPipeline* p, int* n)
{
	CHECK_ARG(p && n);
	*n = num_bodies(p);
	return B3H_OK;
}
extern "C" int b3h_pipeline_set_solver_params(b3hPipeline* p, const b3hSolverParams* prm)
{
	CHECK_ARG(p && prm && prm->m_numIterations >= 0 && prm->m_numIterations <= B3H_MAX_ITERATIONS && prm->m_solverDt > 0.f &&
			  prm->m_batchCellSize > 0.f);
	CHECK_ARG(prm->m_warmStartMode >= 0 && prm->m_warmStartMode <= 2);
	if (prm->m_warmStartMode == 2 && !p->warmNet)
	{
		b3h_set_error("setSolverParams: the learned warm start (mode 2) needs a network "
					  "(b3h_pipeline_set_warm_start_model first)");
		return B3H_ERR_STATE;
	}
	if (prm->m_warmStartMode >= 1)
	{
		int r = warm_alloc(p, prm->m_warmStartMode == 2 || p->logRows);
		if (r) return r;
	}
	p->prm = *prm;
	drop_graph(p);
	return B3H_OK;
}
extern "C" int b3h_pipeline_set_warm_start_model(b3hPipeline* p, b3hMlp* net)
{
	CHECK_ARG(p);
	if (net)
	{
		int in = 0, out = 0;
		b3h_mlp_get_dims(net, &in, nullptr, &out);
		if (in != B3H_WARM_FEATURES || out != B3H_WARM_OUTPUTS)
		{
			b3h_set_error("setWarmStartModel: the network maps %d features to %d outputs, the warm start needs %d -> %d",
						  in, out, B3H_WARM_FEATURES, B3H_WARM_OUTPUTS);
			return B3H_ERR_INVALID;
		}
	}
	else if (p->prm.m_warmStartMode == 2)
	{
		b3h_set_error("setWarmStartModel: cannot remove the network while the learned warm start is selected");
		return B3H_ERR_STATE;
	}
	p->warmNet = net;
	drop_graph(p);
	return B3H_OK;
}
extern "C" int b3h_pipeline_set_contact_log(b3hPipeline* p, float* featuresDev, float* targetsDev, int capacityRows)
{
	CHECK_ARG(p && capacityRows >= 0 && ((featuresDev && targetsDev) || capacityRows == 0));
	if (capacityRows > 0)
	{
		int r = warm_alloc(p, true);
		if (r) return r;
	}
	p->logRows = capacityRows > 0;
	p->wb.logX = capacityRows > 0 ? featuresDev : nullptr;
	p->wb.logT = capacityRows > 0 ? targetsDev : nullptr;
	p->wb.logCap = capacityRows;
	if (p->warmAlloc) B3H_CHECK(hipMemsetAsync(p->wb.logCount, 0, sizeof(int), p->stream));
	drop_graph(p);
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_contact_log_count(b3hPipeline* p, int* rows)
{
	CHECK_ARG(p && rows);
	*rows = 0;
	if (!p->warmAlloc) return B3H_OK;
	return read_int(p->wb.logCount, rows, p->stream);
}
extern "C" int b3h_pipeline_get_warm_start_debug(b3hPipeline* p, float** warmDev, float** prevDev,
												 float** featuresDev)
{
	CHECK_ARG(p);
	if (warmDev) *warmDev = p->wb.warm;
	if (prevDev) *prevDev = p->wb.prev;
	if (featuresDev) *featuresDev = p->wb.X;
	return B3H_OK;
}
extern "C" int b3h_pipeline_clear_warm_start_history(b3hPipeline* p)
{
	CHECK_ARG(p);
	if (p->warmAlloc) B3H_CHECK(hipMemsetAsync(p->wb.histValid, 0, sizeof(int), p->stream));
	return B3H_OK;
}

/* Simulation state as one byte image: header, bodies, inertias, then the
 * warm-start history (key and entry per constraint of the last solve).  The
 * shape tables and registrations are set-up, not state: a state loads into a
 * pipeline with the same bodies registered. */
namespace
{
const char kStateMagic[8] = {'B', '3', 'H', 'S', 'T', 'A', 'T', 'E'};
const int kStateVersion = 1;
struct StateHeader
{
	char magic[8];
	int version;
	int headerBytes;
	int numBodies;
	int numHist;
	F4 gravity;
	b3hSolverParams prm;
};
size_t state_bytes(int nb, int nh)
{
	return sizeof(StateHeader) + (size_t)nb * (sizeof(b3hRigidBodyData) + sizeof(b3hInertiaData)) +
		   (size_t)nh * (sizeof(unsigned long long) + sizeof(b3hWarmHist));
}
}  // namespace

extern "C" int b3h_pipeline_save_state(b3hPipeline* p, void* host, size_t capacityBytes, size_t* bytesOut)
{
	CHECK_ARG(p && bytesOut);
	int nb = num_bodies(p), nh = 0;
	B3H_CHECK(hipStreamSynchronize(p->stream));
	if (p->warmAlloc) B3H_CHECK(hipMemcpy(&nh, p->wb.histValid, sizeof(int), hipMemcpyDeviceToHost));
	size_t need = state_bytes(nb, nh);
	*bytesOut = need;
	if (!host) return B3H_OK;
	if (capacityBytes < need)
	{
		b3h_set_error("saveState: %zu bytes needed, %zu given", need, capacityBytes);
		return B3H_ERR_CAPACITY;
	}
	StateHeader h;
	memset(&h, 0, sizeof(h));
	memcpy(h.magic, kStateMagic, 8);
	h.version = kStateVersion;
	h.headerBytes = (int)sizeof(StateHeader);
	h.numBodies = nb;
	h.numHist = nh;
	h.gravity = p->gravity;
	h.prm = p->prm;
	char* o = (char*)host;
	memcpy(o, &h, sizeof(h));
	o += sizeof(h);
	if (nb > 0)
	{
		B3H_CHECK(hipMemcpy(o, p->np->dBodies, nb * sizeof(b3hRigidBodyData), hipMemcpyDeviceToHost));
		o += nb * sizeof(b3hRigidBodyData);
		B3H_CHECK(hipMemcpy(o, p->np->dInertias, nb * sizeof(b3hInertiaData), hipMemcpyDeviceToHost));
		o += nb * sizeof(b3hInertiaData);
	}
	if (nh > 0)
	{
		B3H_CHECK(hipMemcpy(o, p->wb.histKeys, nh * sizeof(unsigned long long), hipMemcpyDeviceToHost));
		o += nh * sizeof(unsigned long long);
		B3H_CHECK(hipMemcpy(o, p->wb.hist, nh * sizeof(b3hWarmHist), hipMemcpyDeviceToHost));
	}
	return B3H_OK;
}

extern "C" int b3h_pipeline_load_state(b3hPipeline* p, const void* host, size_t bytes)
{
	CHECK_ARG(p && host);
	StateHeader h;
	if (bytes < sizeof(h))
	{
		b3h_set_error("loadState: %zu bytes is shorter than the header", bytes);
		return B3H_ERR_INVALID;
	}
	memcpy(&h, host, sizeof(h));
	if (memcmp(h.magic, kStateMagic, 8) != 0 || h.version != kStateVersion || h.headerBytes != (int)sizeof(h))
	{
		b3h_set_error("loadState: not a state image of this version");
		return B3H_ERR_INVALID;
	}
	int nb = num_bodies(p);
	if (h.numBodies != nb || h.numHist < 0 || bytes != state_bytes(h.numBodies, h.numHist))
	{
		b3h_set_error("loadState: image holds %d bodies and %d history entries in %zu bytes; the pipeline has %d bodies",
					  h.numBodies, h.numHist, bytes, nb);
		return B3H_ERR_STATE;
	}
	if (h.numHist > p->cfg.m_maxContactCapacity)
	{
		b3h_set_error("loadState: %d history entries exceed the contact capacity %d", h.numHist,
					  p->cfg.m_maxContactCapacity);
		return B3H_ERR_CAPACITY;
	}
	const char* src = (const char*)host + sizeof(h);
	const b3hRigidBodyData* bodies = (const b3hRigidBodyData*)src;
	for (int i = 0; i < nb; i++)
		if (bodies[i].m_collidableIdx != p->np->bodies[i].m_collidableIdx)
		{
			b3h_set_error("loadState: body %d has collidable %d in the image, %d in the pipeline", i,
						  bodies[i].m_collidableIdx, p->np->bodies[i].m_collidableIdx);
			return B3H_ERR_STATE;
		}
	int r = b3h_pipeline_set_solver_params(p, &h.prm);
	if (r) return r;
	if (h.numHist > 0 && (r = warm_alloc(p, false))) return r;
	B3H_CHECK(hipStreamSynchronize(p->stream));
	p->gravity = h.gravity;
	drop_graph(p);
	if (nb > 0)
	{
		memcpy(p->np->bodies.data(), src, nb * sizeof(b3hRigidBodyData));
		B3H_CHECK(hipMemcpy(p->np->dBodies, src, nb * sizeof(b3hRigidBodyData), hipMemcpyHostToDevice));
		src += nb * sizeof(b3hRigidBodyData);
		memcpy(p->np->inertias.data(), src, nb * sizeof(b3hInertiaData));
		B3H_CHECK(hipMemcpy(p->np->dInertias, src, nb * sizeof(b3hInertiaData), hipMemcpyHostToDevice));
		src += nb * sizeof(b3hInertiaData);
	}
	if (p->warmAlloc)
	{
		if (h.numHist > 0)
		{
			B3H_CHECK(hipMemcpy(p->wb.histKeys, src, h.numHist * sizeof(unsigned long long), hipMemcpyHostToDevice));
			src += h.numHist * sizeof(unsigned long long);
			B3H_CHECK(hipMemcpy(p->wb.hist, src, h.numHist * sizeof(b3hWarmHist), hipMemcpyHostToDevice));
		}
		b3h_warm_rebuild(h.numHist, p->wb, p->stream);
		B3H_CHECK(hipGetLastError());
		B3H_CHECK(hipStreamSynchronize(p->stream));
	}
	return B3H_OK;
}

/* Autotuner: times `steps` eager steps from the current state for every
 * combination of narrowphase lanes per pair (64, 32, 16) and PGS parts per
 * cell (8, 4, 1), restoring the state image before each trial, then keeps the
 * fastest and restores the state once more.  The knobs do not change results,
 * so the simulation continues exactly as if no trial had run. */
extern "C" int b3h_pipeline_autotune(b3hPipeline* p, float dt, int steps, int* lanesOut, int* partsOut,
									 float* msOut)
{
	CHECK_ARG(p && steps >= 1 && steps <= 64);
	if (num_bodies(p) == 0) return B3H_OK;
	size_t n = 0;
	int r = b3h_pipeline_save_state(p, nullptr, 0, &n);
	if (r) return r;
	std::vector<unsigned char> image(n);
	if ((r = b3h_pipeline_save_state(p, image.data(), n, &n))) return r;
	const int lanes[3] = {64, 32, 16}, parts[3] = {B3H_PGS_PARTS, 4, 1};
	hipEvent_t e0, e1;
	B3H_CHECK(hipEventCreate(&e0));
	B3H_CHECK(hipEventCreate(&e1));
	float best = INFINITY;
	int bestL = p->np->lanesPerPair, bestP = p->sb.pgsMaxParts > 0 ? p->sb.pgsMaxParts : B3H_PGS_PARTS;
	/* trial steps append nothing to a contact log */
	bool logSave = p->logRows;
	p->logRows = false;
	for (int li = 0; li < 3 && !r; li++)
		for (int pi = 0; pi < 3 && !r; pi++)
		{
			if ((r = b3h_pipeline_load_state(p, image.data(), n)) || (r = np_ensure_sat(p->np))) break;
			p->np->lanesPerPair = lanes[li];
			p->sb.pgsMaxParts = parts[pi];
			hipEventRecord(e0, p->stream);
			for (int k = 0; k < steps && !r; k++) r = pl_step(p, dt, false);
			hipEventRecord(e1, p->stream);
			if (hipEventSynchronize(e1) != hipSuccess)
			{
				b3h_set_error("autotune: %s", hipGetErrorString(hipGetLastError()));
				r = B3H_ERR_HIP;
				break;
			}
			float ms = 0.f;
			hipEventElapsedTime(&ms, e0, e1);
			ms /= (float)steps;
			if (ms < best)
			{
				best = ms;
				bestL = lanes[li];
				bestP = parts[pi];
			}
		}
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	p->logRows = logSave;
	p->np->lanesPerPair = bestL;
	p->np->gen++;
	p->sb.pgsMaxParts = bestP;
	int r2 = b3h_pipeline_load_state(p, image.data(), n);
	drop_graph(p);
	if (r) return r;
	if (r2) return r2;
	if (lanesOut) *lanesOut = bestL;
	if (partsOut) *partsOut = bestP;
	if (msOut) *msOut = best;
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_solver_params(b3hPipeline* p, b3hSolverParams* prm)
{
	CHECK_ARG(p && prm);
	*prm = p->prm;
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_solver_debug(b3hPipeline* p, b3hContact4Data** sc, b3hContactConstraint4** cs, int** cc,
											 int** co, int** bs, int* n)
{
	CHECK_ARG(p);
	if (sc) *sc = p->sb.sortedContacts;
	if (cs) *cs = p->sb.constraints;
	if (cc) *cc = p->sb.cellCounts;
	if (co) *co = p->sb.cellOffsets;
	if (bs) *bs = p->sb.batchSizes;
	if (n) return read_int(p->np->dNumContacts, n, p->stream);
	return B3H_OK;
}
extern "C" int b3h_pipeline_set_pgs_body_cache(b3hPipeline* p, int maxBodiesPerCell)
{
	CHECK_ARG(p);
	drop_graph(p);
	p->sb.pgsBodyCap = maxBodiesPerCell;
	if (p->graphExec)
	{
		hipGraphExecDestroy(p->graphExec);
		p->graphExec = nullptr;
	}
	if (p->graph)
	{
		hipGraphDestroy(p->graph);
		p->graph = nullptr;
	}
	return B3H_OK;
}

extern "C" int b3h_pipeline_set_pgs_max_parts(b3hPipeline* p, int maxParts)
{
	CHECK_ARG(p && maxParts >= 1 && maxParts <= B3H_PGS_PARTS);
	p->sb.pgsMaxParts = maxParts;
	drop_graph(p);
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_pgs_levels(b3hPipeline* p, int* levels)
{
	CHECK_ARG(p && levels);
	B3H_CHECK(hipMemcpyAsync(levels, p->sb.cellLevels, sizeof(int) * B3H_N_CELLS, hipMemcpyDeviceToHost, p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_pgs_parts(b3hPipeline* p, int* parts, int* largest)
{
	CHECK_ARG(p && parts);
	std::vector<b3hInt4> meta(2 * (size_t)B3H_N_CELLS * B3H_PGS_PARTS);
	B3H_CHECK(hipMemcpyAsync(meta.data(), p->sb.partMeta, meta.size() * sizeof(b3hInt4), hipMemcpyDeviceToHost,
							 p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	for (int c = 0; c < B3H_N_CELLS; c++)
	{
		int k = 0, big = 0;
		for (int q = 0; q < B3H_PGS_PARTS; q++)
		{
			const b3hInt4& m = meta[2 * ((size_t)c * B3H_PGS_PARTS + q)];
			if (m.y > 0) k++;
			if (m.y > big) big = m.y;
		}
		parts[c] = k;
		if (largest) largest[c] = big;
	}
	return B3H_OK;
}

extern "C" int b3h_pipeline_set_track_residual(b3hPipeline* p, int enable)
{
	CHECK_ARG(p);
	drop_graph(p);
	p->trackResidual = enable != 0;
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_residuals(b3hPipeline* p, float* out, int n)
{
	CHECK_ARG(p && out && n >= 0 && n <= 2 * p->prm.m_numIterations);
	std::vector<float> tmp(4 * (size_t)p->prm.m_numIterations + 4);
	B3H_CHECK(hipMemcpyAsync(tmp.data(), p->sb.resid, sizeof(float) * 4 * p->prm.m_numIterations, hipMemcpyDeviceToHost,
							 p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	for (int i = 0; i < n; i++)
	{
		float d = tmp[2 * i], l = tmp[2 * i + 1];
		out[i] = d / (l > 1e-9f ? l : 1e-9f);
	}
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_residuals_l2(b3hPipeline* p, float* out, int n)
{
	CHECK_ARG(p && out && n >= 0 && n <= 2 * p->prm.m_numIterations);
	std::vector<double> tmp(4 * (size_t)p->prm.m_numIterations + 4);
	B3H_CHECK(hipMemcpyAsync(tmp.data(), p->sb.residL2, sizeof(double) * 4 * p->prm.m_numIterations, hipMemcpyDeviceToHost,
							 p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	for (int i = 0; i < n; i++)
	{
		double d = tmp[2 * i], l = tmp[2 * i + 1];
		out[i] = (float)(l > 0.0 ? sqrt(d / l) : 0.0);
	}
	return B3H_OK;
}
extern "C" int b3h_pipeline_set_warm_start_impulses(b3hPipeline* p, const float* dev, int n)
{
	CHECK_ARG(p && n >= 0);
	drop_graph(p);
	p->dWarm = (float*)dev;
	p->warmCount = n;
	if (p->graphExec)
	{
		hipGraphExecDestroy(p->graphExec);
		p->graphExec = nullptr;
	}
	return B3H_OK;
}
extern "C" int b3h_pipeline_enable_timing(b3hPipeline* p, int enable)
{
	CHECK_ARG(p);
	p->timing = enable != 0;
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_stage_times(b3hPipeline* p, float* ms)
{
	CHECK_ARG(p && ms);
	if (!p->evValid)
	{
		b3h_set_error("timing not enabled");
		return B3H_ERR_STATE;
	}
	B3H_CHECK(hipEventSynchronize(p->ev[B3H_NUM_STAGES]));
	for (int i = 0; i < B3H_NUM_STAGES; i++)
	{
		float t = 0.f;
		B3H_CHECK(hipEventElapsedTime(&t, p->ev[i], p->ev[i + 1]));
		ms[i] = t;
	}
	return B3H_OK;
}
extern "C" int b3h_pipeline_get_overflow(b3hPipeline* p, int* po, int* co)
{
	CHECK_ARG(p);
	int pc[2], cc[2];
	B3H_CHECK(hipMemcpyAsync(pc, p->bp->dPairCount, sizeof(pc), hipMemcpyDeviceToHost, p->stream));
	B3H_CHECK(hipMemcpyAsync(cc, p->np->dNumContacts, sizeof(cc), hipMemcpyDeviceToHost, p->stream));
	B3H_CHECK(hipStreamSynchronize(p->stream));
	if (po) *po = pc[1] > pc[0];
	if (co) *co = cc[1] > cc[0];
	return B3H_OK;
}

extern "C" int b3h_radix_sort32(unsigned* keys, unsigned* vals, int n, int bits)
{
	CHECK_ARG(n >= 0 && bits >= 0 && bits <= 32);
	if (n == 0) return B3H_OK;
	CHECK_ARG(keys && vals);
	int r = ensure_device();
	if (r) return r;
	unsigned *tk = nullptr, *tv = nullptr, *h = nullptr;
	if ((r = dalloc(&tk, n)) || (r = dalloc(&tv, n)) || (r = dalloc(&h, b3h_radix_hist_size(n))))
	{
		dfree(tk);
		dfree(tv);
		return r;
	}
	b3h_radix_sort(keys, vals, tk, tv, h, n, bits, 0);
	hipError_t e = hipDeviceSynchronize();
	dfree(tk);
	dfree(tv);
	dfree(h);
	B3H_CHECK(e);
	return B3H_OK;
}

extern "C" int b3h_prefix_scan(const int* src, int* dst, int n, int* sumOut)
{
	CHECK_ARG(n >= 0 && src && dst);
	int r = ensure_device();
	if (r) return r;
	unsigned* tmp = nullptr;
	if ((r = dalloc(&tmp, b3h_scan_tmp_size(n)))) return r;
	b3h_exclusive_scan(src, dst, nullptr, n, tmp, 0);
	hipError_t e = hipDeviceSynchronize();
	if (e == hipSuccess && sumOut) e = hipMemcpy(sumOut, dst + n, sizeof(int), hipMemcpyDeviceToHost);
	dfree(tmp);
	B3H_CHECK(e);
	return B3H_OK;
}

extern "C" int b3h_compute_polyhedral_features(const float* vertices, int strideInBytes, int numVertices, int cap,
											   int* counts, float* v4, float* p4, int* fc, int* idx, float* e4)
{
	CHECK_ARG(vertices && counts && numVertices > 0 && strideInBytes >= 12 && (strideInBytes % 4) == 0);
	b3hPolyhedron poly;
	if (b3h_build_polyhedron(vertices, numVertices, strideInBytes / 4, &poly) != 0)
	{
		b3h_set_error("degenerate point set");
		return B3H_ERR_INVALID;
	}
	int nv = (int)poly.vertices.size() / 3, nf = (int)poly.faces.size(), ne = (int)poly.uniqueEdges.size() / 3;
	int ni = 0;
	for (auto& f : poly.faces) ni += (int)f.indices.size();
	counts[0] = nv;
	counts[1] = nf;
	counts[2] = ni;
	counts[3] = ne;
	if (nv > cap || nf > cap || ni > cap || ne > cap)
	{
		if (v4 || p4 || fc || idx || e4)
		{
			b3h_set_error("capacity %d too small", cap);
			return B3H_ERR_CAPACITY;
		}
		return B3H_OK;
	}
	if (v4)
		for (int i = 0; i < nv; i++)
			for (int k = 0; k < 4; k++) v4[4 * i + k] = k < 3 ? poly.vertices[3 * i + k] : 0.f;
	int w = 0;
	for (int f = 0; f < nf; f++)
	{
		if (p4)
			for (int k = 0; k < 4; k++) p4[4 * f + k] = poly.faces[f].plane[k];
		if (fc) fc[f] = (int)poly.faces[f].indices.size();
		for (int id : poly.faces[f].indices)
		{
			if (idx) idx[w] = id;
			w++;
		}
	}
	if (e4)
		for (int i = 0; i < ne; i++)
			for (int k = 0; k < 4; k++) e4[4 * i + k] = k < 3 ? poly.uniqueEdges[3 * i + k] : 0.f;
	return B3H_OK;
}

/* ---------------- device memory helpers ---------------- */
extern "C" int b3h_device_count(int* n)
{
	CHECK_ARG(n);
	*n = 0;
	hipError_t e = hipGetDeviceCount(n);
	if (e != hipSuccess) *n = 0;
	return B3H_OK;
}
extern "C" int b3h_malloc(void** d, size_t bytes)
{
	CHECK_ARG(d);
	int r = ensure_device();
	if (r) return r;
	B3H_CHECK(hipMalloc(d, bytes ? bytes : 1));
	return B3H_OK;
}
extern "C" int b3h_free(void* d)
{
	if (d) B3H_CHECK(hipFree(d));
	return B3H_OK;
}
extern "C" int b3h_memcpy_htod(void* d, const void* h, size_t b)
{
	B3H_CHECK(hipMemcpy(d, h, b, hipMemcpyHostToDevice));
	return B3H_OK;
}
extern "C" int b3h_memcpy_dtoh(void* h, const void* d, size_t b)
{
	B3H_CHECK(hipMemcpy(h, d, b, hipMemcpyDeviceToHost));
	return B3H_OK;
}
extern "C" int b3h_memcpy_dtod(void* d, const void* s, size_t b)
{
	B3H_CHECK(hipMemcpy(d, s, b, hipMemcpyDeviceToDevice));
	return B3H_OK;
}
extern "C" int b3h_synchronize(void)
{
	B3H_CHECK(hipDeviceSynchronize());
	return B3H_OK;
}
