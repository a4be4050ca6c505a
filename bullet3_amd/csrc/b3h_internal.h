/*
 * b3h_internal.h -- library-internal declarations: kernel launchers and the
 * state behind the opaque C-ABI handles.
 */
#ifndef B3H_INTERNAL_H
#define B3H_INTERNAL_H

#include <hip/hip_runtime.h>
#include <vector>
#include <string>
#include "b3h_math.h"
#include "../../include/b3hip.h"

#define B3H_CHECK(x)                                                      \
	do                                                                    \
	{                                                                     \
		hipError_t e_ = (x);                                              \
		if (e_ != hipSuccess)                                             \
		{                                                                 \
			b3h_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
			return B3H_ERR_HIP;                                           \
		}                                                                 \
	} while (0)

void b3h_set_error(const char* fmt, ...);

/* ---------------- primitives (k_prims.hip) ---------------- */
void b3h_radix_sort(unsigned* keys, unsigned* vals, unsigned* tmpKeys, unsigned* tmpVals, unsigned* hist, int n, int bits,
					hipStream_t s);
size_t b3h_radix_hist_size(int n);
void b3h_exclusive_scan(const int* in, int* out, const int* nDev, int maxN, unsigned* tileSums, hipStream_t s);
size_t b3h_scan_tmp_size(int maxN);
void b3h_fill_i32(int* p, int v, int n, hipStream_t s);

/* ---------------- per-body kernels (k_body.hip) ---------------- */
void b3h_launch_world_aabbs(int n, const b3hRigidBodyData* bodies, const b3hCollidable* cols, const b3hAabb* localAabbs,
							b3hAabb* world, hipStream_t s);
/* skip (optional, per body): bodies with skip[i] >= 0 were integrated by the solver */
void b3h_launch_integrate(int n, b3hRigidBodyData* bodies, float dt, float damping, F4 gravity, hipStream_t s,
						  const int* skip = nullptr);

/* ---------------- broadphase (k_broadphase.hip) ---------------- */
struct b3hGridParams
{
	float invCell[3];
	int dims[3];
	int maxBodiesPerCell;
};
struct b3hBroadphaseScratch
{
	unsigned *keys, *vals, *tmpKeys, *tmpVals, *hist;
	int* cellStart; /* dims product */
	int* counts;    /* 2 * nSmall + 1 (large-small counts, then grid/sap counts) */
	int* offsets;   /* 2 * nSmall + 2 */
	unsigned* scanTmp;
	b3hAabb* sortedAabbs;
	double* sapPartials;
	int* axis;
	/* large proxies sorted by min.x (used when there are many of them) */
	unsigned *lkeys, *lvals, *ltmpKeys, *ltmpVals, *lhist;
	unsigned* lmaxExt; /* float bits of the largest x-extent */
	float* lminx;      /* min.x in sorted order */
	const int* sceneOf; /* per proxy handle: its scene (NULL: one scene); pairs across scenes are dropped */
	int2* stash;        /* grid and BVH pairs: the first B3H_GRID_STASH pairs of each small proxy (count pass) */
	/* parallel linear BVH over the small proxies (BVH broadphase only) */
	b3hAabb* bvhMerge;  /* per-block merged boxes, then the merged box */
	int2* bvhChild;     /* internal nodes: (left, right), internal children marked by the sign bit */
	b3hFloat4* bvhBox;  /* internal nodes: (min.xyz, highest leaf), (max.xyz, lowest leaf) */
	int* bvhFlag;       /* internal nodes: children arrived (build) */
	int* bvhMeta;       /* root, traversal stack overflow */
};
#define B3H_GRID_STASH 64
#define B3H_LARGE_BRUTE_MAX 32
void b3h_grid_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
					const b3hGridParams& gp, int numCells, b3hInt4* pairs, int maxPairs, int* pairCount,
					int* pairTotal, b3hBroadphaseScratch& sc, hipStream_t s);
void b3h_sap_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
				   b3hInt4* pairs, int maxPairs, int* pairCount, int* pairTotal, b3hBroadphaseScratch& sc,
				   hipStream_t s);
int b3h_sap_num_partials(int nSmall);
int b3h_plbvh_reduce_blocks(int nSmall);
void b3h_plbvh_build(const b3hAabb* all, const int* smallMap, int nSmall, b3hBroadphaseScratch& sc, hipStream_t s);
void b3h_plbvh_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
					 b3hInt4* pairs, int maxPairs, int* pairCount, int* pairTotal, b3hBroadphaseScratch& sc,
					 hipStream_t s);
void b3h_scene_contact_counts(const int* nPairs, int maxPairs, const b3hInt4* pairs, const int* sceneOf, int* counts,
							  int numScenes, hipStream_t s);

/* ---------------- host geometry (b3h_api.cpp, compound.cpp) ---------------- */
/* b3Matrix3x3::setRotation of (x, y, z, w) */
void b3h_quat_to_basis(const float q[4], float m[3][3]);
/* a compound's local AABB and quantized child tree (offsets in info left 0) */
int b3h_build_compound(int n, const b3hGpuChildShape* children, const b3hAabb* childLocal, b3hAabb* compoundAabb,
					   b3hBvhInfo* info, std::vector<b3hQuantizedBvhNode>& nodes,
					   std::vector<b3hBvhSubtreeInfo>& subtrees);

/* compute units of the current device (cached per device) */
int b3h_cu_count();

/* ---------------- narrowphase (k_narrowphase.hip) ---------------- */
struct b3hShapeTables
{
	const b3hCollidable* collidables;
	const b3hConvexPolyhedronData* convexes;
	const b3hFloat4* vertices;
	const b3hFloat4* uniqueEdges;
	const b3hGpuFace* faces;
	const int* indices;
	/* compounds: children, and the quantized child trees */
	const b3hGpuChildShape* children;
	const b3hBvhInfo* bvhInfos;
	const b3hQuantizedBvhNode* bvhNodes;
	const b3hBvhSubtreeInfo* bvhSubtrees;
};
/* Per-body face records, rebuilt every step before the SAT kernel: for face f
 * of body b's hull, the world face normal w = qrot(orn_b, plane_f) and the
 * interval [mn, mx] of b's own hull projected on w (the projection the SAT
 * would otherwise repeat for every pair b is in).  Records of body b sit at
 * [base[b], base[b+1]); owner[r] is the body of record r.  A body whose range
 * does not match its hull's face count (device bodies changed behind the host
 * mirror) is treated as uncached. */
struct b3hSatFace
{
	b3hFloat4 w;
	float mn, mx, pad0, pad1;
};
#define B3H_SAT_DIRS 162 /* fixed directions of the edge-edge approximation */
/* Per-body SAT descriptor, rebuilt every step: the pose, the hull centre in
 * the SAT frame (position with w = 0) and the table offsets the SAT passes
 * read, so a pair reaches both hulls with one load per body. */
#define B3H_SB_HULL 1   /* the body's collidable is a convex hull */
#define B3H_SB_STATIC 2 /* inverse mass 0 */
struct b3hSatBody
{
	b3hFloat4 pos, quat;
	b3hFloat4 center;
	int vOff, nV, fOff, nF;
	int eOff, nE, rec, dir; /* face-record base and fixed-direction slot, -1 when absent */
	int flags, col, pad0, pad1;
};
struct b3hSatCache
{
	b3hSatFace* recs;
	const int* base;  /* numBodies + 1 */
	const int* owner; /* total */
	int numBodies;    /* bodies covered; 0 disables the cache */
	int total;
	/* intervals of a body's hull on the 162 fixed directions (the edge-edge
	 * approximation): body b's 162 (mn, mx) at dirs[162 dslot[b]]; bodies
	 * that cannot meet a pair taking that branch have dslot -1 */
	float2* dirs;
	const int* dslot;  /* numBodies */
	const int* downer; /* numDirBodies */
	int numDirBodies;
	b3hSatBody* desc; /* numBodies */
};
/* per-pair buffers between the SAT passes (maxPairs entries each; flag and
 * scan one and two more) */
struct b3hSatWork
{
	float* dmin; /* the face families' minimum overlap of an open pair */
	int* flag;   /* pass 1: pair left open; pass 2: pair to clip */
	int* scan;   /* exclusive prefix sum of flag */
	int* list;   /* the flagged pairs in pair order */
	int* count;  /* their number */
};
/* compound child pairs (enabled when a compound shape is registered): the
 * child pairs go to cpairs, their contacts to staging[stagingBase + k] */
struct b3hCompoundWork
{
	int cap;             /* child pairs kept (0: no compounds) */
	int stagingBase;     /* first staging record of the child pairs */
	int* counts;         /* per pair: its child pairs (maxPairs + 1) */
	int* offsets;        /* exclusive prefix sum of counts (maxPairs + 2) */
	b3hInt4* cpairs;     /* cap */
	int* num;            /* [0] child pairs kept, [1] found */
	int* flags;          /* cap + 1 */
	int* scan;           /* cap + 2 */
	unsigned* scanTmp;   /* b3h_scan_tmp_size(max(maxPairs, cap)) */
};
/* sepState and sepNormal (per pair) are required: they carry the SAT state
 * between the passes */
void b3h_narrowphase(b3hInt4* pairs, const int* numPairs, int maxPairs, const b3hRigidBodyData* bodies,
					 const b3hShapeTables& sh, b3hContact4Data* staging, int* flags, int* flagScan,
					 unsigned* scanTmp, int* contactPair, int maxContacts, int* numContacts, int* sepState,
					 b3hFloat4* sepNormal, int pairsPerGroup, hipStream_t s, hipEvent_t evAfterKernel,
					 b3hFloat4* mprScratch, hipEvent_t evAfterMpr, bool hasPrimitives, const b3hSatCache* satCache,
					 const b3hSatWork& work, const b3hCompoundWork& cw);
void b3h_gather_contacts(const int* contactPair, const int* numContacts, int maxContacts,
						 const b3hContact4Data* staging, b3hContact4Data* out, hipStream_t s);

/* ---------------- joints (k_joints.hip) ---------------- */
struct b3hJointRow
{
	F4 r1xn, normal, r2xn, angA, angB;
	float applied, jacInv, rhs, cfm, lo, hi, invMassA, invMassB;
	int slotA, slotB, pad0, pad1;
};
struct b3hJointParams
{
	float timeStep, erp, globalCfm, damping;
	int iterations;
};
struct b3hJointBuffers
{
	int numConstraints;
	b3hGpuGenericConstraint* constraints; /* device copy (flags updated by breaking) */
	int* rowOffset;  /* per constraint */
	int* rowCount;   /* per constraint: 3, 6 or 0 */
	int* slotOf;     /* per constraint: joint-body slots of A and B */
	b3hJointRow* rows;
	int numBatches;
	int* batchStart; /* numBatches + 1 */
	int* order;      /* constraint index per batched position */
	F4* delta;       /* 2 per joint body: linear, angular velocity deltas */
	int numSlots;
	int* bodySlot;   /* per body: its joint-body slot or -1 */
	int numBodies;
};
void b3h_solve_joints(const b3hJointBuffers& jb, b3hRigidBodyData* bodies, const b3hInertiaData* inertias,
					  const b3hJointParams& jp, hipStream_t s);

/* ---------------- ray casts (k_raycast.hip) ---------------- */
void b3h_cast_rays(int numRays, const b3hRayInfo* rays, b3hRayHit* hits, const b3hAabb* all, const int* smallMap,
				   int nSmall, const int* largeMap, int nLarge, const b3hRigidBodyData* bodies,
				   const b3hShapeTables& sh, b3hBroadphaseScratch& sc, hipStream_t s);

/* ---------------- solver (k_solver.hip) ---------------- */
#define B3H_N_SPLIT_X 8
#define B3H_N_SPLIT_Y 4
#define B3H_N_SPLIT_Z 8
#define B3H_N_CELLS (B3H_N_SPLIT_X * B3H_N_SPLIT_Y * B3H_N_SPLIT_Z)
#define B3H_N_CELL_BATCHES 8
/* work-groups a cell's sweep may be split into: the cell's copies in space
 * (cells wrap with the split counts) share no body and are solved apart */
#define B3H_PGS_PARTS 8

/* integration fused into the last friction sweep */
struct b3hIntegrateParams
{
	float dt, damping;
	F4 gravity;
};

struct b3hSolverBuffers
{
	int* cellOf;          /* per contact slot */
	int* cellCounts;      /* 256 */
	int* cellOffsets;     /* 257 */
	int* cellCursor;      /* 256 */
	int* cellSlots;       /* per position: contact slot (unsorted within cell) */
	int* perm;            /* per final position: contact slot */
	int* batchIdx;        /* per final position */
	int* batchSizes;      /* 256 */
	unsigned long long* sortKeys; /* per position (global fallback) */
	b3hContactConstraint4* constraints;
	b3hContact4Data* sortedContacts; /* optional */
	float* resid;         /* 2 * maxIters * 2 (maxDelta, maxLambda) */
	float* warmImpulses;  /* optional, 6 per constraint */
	int* cellBodies;      /* per cell at 2*cellOffsets[cell]: its distinct dynamic bodies, ascending */
	int* cellBodyCount;   /* 256 */
	int* cslot;           /* per constraint: local slots of bodyA, bodyB in its cell's list (-1: static) */
	int pgsBodyCap;       /* < 0: the LDS budget; otherwise a lower cap on cached bodies per cell */
	int pgsMaxParts;       /* work-groups a cell may be split into (0: B3H_PGS_PARTS) */
	b3hFloat4* recN;      /* 10 fields x maxContacts float4 of normal-row invariants, by level position */
	b3hFloat4* recF;      /* 8 fields x maxContacts float4 of friction-row invariants, by level position */
	b3hFloat4* lamN;      /* per level position: normal impulses during the sweeps */
	b3hFloat4* lamF;      /* per level position: friction impulses (x, y) during the sweeps */
	int* posOf;           /* per constraint: its level position (-1: cell on the global walk) */
	int* levelOrder;      /* per cell region [off, off+n): per part, constraints grouped by dependency level */
	int* levelStart;      /* per cell region [off+PARTS*cell, off+PARTS*cell+n+PARTS): per part its level starts */
	int* cellLevels;      /* 256: levels per cell, -1 = global walk */
	b3hInt4* partMeta;    /* 2 x 256 x PARTS: (first position, count, level-start offset, levels), (first body entry, bodies) */
	int* partBodies;      /* per cell at 2*cellOffsets[cell]: (slot, body) pairs grouped by part */
	int* lastClass;       /* per body: the last cell-parity class whose cells touch it (-1: none) */
	int maxContacts;
	float* residSnap;     /* per constraint: 6 impulses before the sweep (residual tracking) */
	double* residL2;      /* per sweep: sum of squared impulse changes, sum of squared impulses */
	const int* numContacts;
};
#define B3H_MAX_ITERATIONS 1024
int b3h_solver_module_init();
int b3h_solver_lds_bytes();
/* set-up (ordering, batching, constraints, PGS records and levels), then the sweeps */
void b3h_solve_setup(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias, const int* numContacts,
					 int maxContacts, const b3hContact4Data* staging, const int* contactPair,
					 const b3hSolverParams& prm, int staticIdx, b3hSolverBuffers& sb, bool writeSortedContacts,
					 hipStream_t s);
void b3h_solve_sweeps(b3hRigidBodyData* bodies, const b3hInertiaData* inertias, const b3hSolverParams& prm,
					  b3hSolverBuffers& sb, bool trackResidual, hipStream_t s, hipEvent_t evNormalDone = nullptr,
					  const b3hIntegrateParams* integrate = nullptr);
void b3h_solve_contacts(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias,
						const int* numContacts, int maxContacts, const b3hContact4Data* staging,
						const int* contactPair, const b3hSolverParams& prm, int staticIdx, b3hSolverBuffers& sb,
						bool writeSortedContacts, bool trackResidual, hipStream_t s, hipEvent_t evSetupDone = nullptr,
						hipEvent_t evNormalDone = nullptr, const b3hIntegrateParams* integrate = nullptr);

/* ---------------- warm start (k_warm.hip) ---------------- */
#define B3H_WARM_FEATURES 24
#define B3H_WARM_OUTPUTS 3
struct b3hWarmHist
{
	F4 pos[4]; /* xyz: contact point, w: final normal impulse */
	F4 fric;   /* xyz: final friction impulse as a world vector, w: number of points */
};
struct b3hWarmParams
{
	float velUnit;  /* |g| dt */
	float matchTol; /* contact point matching distance */
	int mode;       /* 1 previous frame, 2 learned */
};
struct b3hWarmBuffers
{
	int capC;                  /* constraints */
	int htCap;                 /* hash slots, a power of two >= 2 capC */
	unsigned long long* htKeys;
	int* htVals;
	b3hWarmHist* hist;         /* [capC] */
	unsigned long long* histKeys; /* [capC]: (bodyA, bodyB) of each history entry */
	int* histValid;            /* 1 int: number of history entries (0: none) */
	float* prev;               /* [capC][8]: matched normal impulses 4, friction 2, pair matched, points */
	float* warm;               /* [capC][6]: initial impulses in solve order */
	float* X;                  /* [4 capC][24] features (learned mode, logging) */
	float* Y;                  /* [4 capC][3] network outputs */
	float* T;                  /* [4 capC][3] targets (logging) */
	int* rowCount;             /* [capC] */
	int* rowScan;              /* [capC + 1] */
	unsigned* scanTmp;
	float* logX;               /* caller's log: [logCap][24] */
	float* logT;               /* [logCap][3] */
	int logCap;
	int* logCount;             /* device int */
};
struct b3hMlp;
void b3h_warm_prepare(const int* numContacts, int maxContacts, const b3hContactConstraint4* cs,
					  const b3hRigidBodyData* bodies, b3hWarmBuffers& wb, const b3hWarmParams& wp, const b3hMlp* net,
					  bool features, hipStream_t s);
void b3h_warm_finish(const int* numContacts, int maxContacts, const b3hContactConstraint4* cs, b3hWarmBuffers& wb,
					 const b3hWarmParams& wp, bool logRows, hipStream_t s);
/* rebuilds the hash table over entries [0, n) of wb.histKeys / wb.hist (after
 * a state load) and sets the entry count */
void b3h_warm_rebuild(int n, b3hWarmBuffers& wb, hipStream_t s);

#endif
