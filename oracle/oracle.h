/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * Declarations of the oracle's C entry points (bound from Python with ctypes).
 */
#ifndef B3ORACLE_H
#define B3ORACLE_H

#include "oracle_math.h"

typedef struct b3oShapes
{
	const b3hCollidable* collidables;
	const b3hConvexPolyhedronData* convexes;
	const b3hFloat4* vertices;
	const b3hFloat4* uniqueEdges;
	const b3hGpuFace* faces;
	const int* indices;
} b3oShapes;

#ifdef __cplusplus
extern "C" {
#endif

void b3o_compute_world_aabbs(int n, const b3hRigidBodyData* bodies, const b3hCollidable* cols,
							 const b3hAabb* localAabbs, b3hAabb* worldAabbs);
void b3o_integrate(int n, b3hRigidBodyData* bodies, float dt, float angDamp, const float* grav3);
void b3o_sort_keys(int n, b3hSortData* kv);
int b3o_grid_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
				   const float* cellSize3, const int* dims3, int maxBodiesPerCell,
				   b3hInt4* pairsOut, int maxPairs, int* totalFound);
int b3o_sap_axis(const b3hAabb* all, const int* smallMap, int nSmall);
int b3o_sap_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
				  int axisOverride, b3hInt4* pairsOut, int maxPairs, int* totalFound, int* axisOut);
/* parallel linear BVH (pairs in the device's fixed order, see oracle_bp.cpp);
 * rootOut (may be NULL): the root node, sign bit set for an internal node */
int b3o_plbvh_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
					b3hInt4* pairsOut, int maxPairs, int* totalFound, int* rootOut);
/* ray casts against the bodies (nearest hit per ray; hits in/out) */
int b3o_cast_rays(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
				  const b3hRigidBodyData* bodies, const b3hCollidable* cols, const b3hConvexPolyhedronData* convexes,
				  const b3hGpuFace* faces, const b3hRayInfo* rays, b3hRayHit* hits, int numRays);
int b3o_brute_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
					b3hInt4* pairsOut, int maxPairs);

/* narrowphase: returns number of contacts; pairs[i].z receives the contact slot
 * or -1.  sepOut (optional, nPairs) receives the SAT state/normal per pair. */
int b3o_mpr_contacts(b3hInt4* pairs, int nPairs, const b3hRigidBodyData* bodies,
					 const b3hCollidable* cols, const b3hConvexPolyhedronData* convexes,
					 const b3hFloat4* vertices, b3hContact4Data* contacts, int capacity);
int b3o_narrowphase(b3hInt4* pairs, int nPairs, const b3hRigidBodyData* bodies,
					const b3hCollidable* cols, const b3hConvexPolyhedronData* convexes,
					const b3hFloat4* vertices, const b3hFloat4* uniqueEdges, const b3hGpuFace* faces,
					const int* indices, b3hContact4Data* contacts, int capacity,
					int* hasSepOut, b3hFloat4* sepOut);
/* compound tables (registerCompoundShape); capacity 0: no compounds */
typedef struct b3oCompoundTables
{
	const b3hGpuChildShape* children;
	const b3hBvhInfo* infos;
	const b3hQuantizedBvhNode* nodes;
	const b3hBvhSubtreeInfo* subtrees;
	int capacity; /* child pairs kept */
} b3oCompoundTables;
/* as b3o_narrowphase, then the compound child pairs: their contacts follow the
 * pair contacts, in child-pair order (pairs in order, each pair's child pairs
 * sorted by (childA, childB) as unsigned); childPairsOut (optional, capacity
 * entries) receives the child pairs and *numChildPairs their number (all
 * found, also beyond the capacity) */
int b3o_narrowphase_ex(b3hInt4* pairs, int nPairs, const b3hRigidBodyData* bodies, const b3hCollidable* cols,
					   const b3hConvexPolyhedronData* convexes, const b3hFloat4* vertices, const b3hFloat4* uniqueEdges,
					   const b3hGpuFace* faces, const int* indices, b3hContact4Data* contacts, int capacity,
					   int* hasSepOut, b3hFloat4* sepOut, const b3oCompoundTables* compound, b3hInt4* childPairsOut,
					   int* numChildPairs);

/* joint solver (constraints' flags updated by breaking; bodies' velocities in
 * place); returns the number of batches */
int b3o_solve_joints(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias,
					 b3hGpuGenericConstraint* constraints, int numConstraints, float timeStep, int iterations);

/* solver front end + PGS (contacts in, bodies updated in place) */
int b3o_solve_contacts(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias,
					   int numContacts, const b3hContact4Data* contactsIn, const b3hSolverParams* params,
					   int staticIdx, b3hContact4Data* sortedContactsOut, b3hContactConstraint4* constraintsOut,
					   unsigned* cellCountsOut, unsigned* cellOffsetsOut, int* batchSizesOut,
					   const float* initialImpulses, float* residualsOut);

#ifdef __cplusplus
}
#endif

/* ---- whole step (b3GpuRigidBodyPipeline::stepSimulation,
 *      b3GpuRigidBodyPipeline.cpp:221-463) ---- */
typedef struct b3oWorld
{
	int numBodies;
	b3hRigidBodyData* bodies;
	const b3hInertiaData* inertias;
	const b3hCollidable* collidables;
	const b3hConvexPolyhedronData* convexes;
	const b3hFloat4* vertices;
	const b3hFloat4* uniqueEdges;
	const b3hGpuFace* faces;
	const int* indices;
	const b3hAabb* localAabbs;
	const int* smallMap;
	int nSmall;
	const int* largeMap;
	int nLarge;
	int broadphase; /* 0 uniform grid, 1 SAP, 2 parallel linear BVH */
	float cellSize[3];
	int gridDims[3];
	int maxBodiesPerCell;
	int maxPairs;
	int maxContacts;
	b3hSolverParams solver;
	float gravity[3];
	int staticIdx;
	/* caller-allocated scratch: numBodies AABBs, maxPairs pairs, maxContacts contacts */
	b3hAabb* worldAabbs;
	b3hInt4* pairs;
	b3hContact4Data* contacts;
	/* outputs */
	int numPairs;
	int numContacts;
	float* residuals; /* optional, 2*iterations */
	/* joints (optional), solved after the narrowphase, before the contacts */
	b3hGpuGenericConstraint* constraints;
	int numConstraints;
	/* compound tables (capacity 0: none) */
	b3oCompoundTables compound;
} b3oWorld;

#ifdef __cplusplus
extern "C" {
#endif
int b3o_step(b3oWorld* w, float dt);
/* Threads the oracle may use (default 1).  Results do not depend on it: the
 * narrowphase compacts its per-pair results in pair order, and the solver runs
 * at once only cells of one parity class, which share no dynamic body. */
void b3o_set_threads(int n);
int b3o_threads(void);
#ifdef __cplusplus
}
#endif

#endif
