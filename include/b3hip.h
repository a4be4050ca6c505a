/*
 * b3hip.h -- C ABI of the MI355X-native Bullet3 GPU rigid-body step.
 *
 * The three handle types stand in for the reference's three classes and every
 * entry point names the method it replaces:
 *   b3hNarrowPhase  <- b3GpuNarrowPhase          src/Bullet3OpenCL/RigidBody/b3GpuNarrowPhase.h:9-99
 *   b3hBroadphase   <- b3GpuBroadphaseInterface  src/Bullet3OpenCL/BroadphaseCollision/b3GpuBroadphaseInterface.h:12-40
 *                      (b3GpuGridBroadphase / b3GpuSapBroadphase)
 *   b3hPipeline     <- b3GpuRigidBodyPipeline    src/Bullet3OpenCL/RigidBody/b3GpuRigidBodyPipeline.h:25-68
 *
 * Conventions: plain pointers and sizes only.  Pointers documented as "device"
 * are HIP device pointers; all others are host pointers.  Every call returns an
 * int status (B3H_OK = 0, negative on error; b3h_last_error() describes it).
 * Where the reference returns an index (or -1 on failure) we return the status
 * and write the index through an out-parameter.  Work is queued on the HIP
 * stream given at creation (NULL = the library creates its own stream of
 * default priority); calls that return data to the host synchronise that
 * stream.  stepSimulation queues the whole step without host synchronisation.
 * Not thread-safe per handle, like the reference.
 */
#ifndef B3HIP_H
#define B3HIP_H

#include <stddef.h>
#include "b3hip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define B3H_OK 0
#define B3H_ERR_INVALID -1
#define B3H_ERR_CAPACITY -2
#define B3H_ERR_HIP -3
#define B3H_ERR_NO_DEVICE -4
#define B3H_ERR_STATE -5

typedef struct b3hNarrowPhase b3hNarrowPhase;
typedef struct b3hBroadphase b3hBroadphase;
typedef struct b3hPipeline b3hPipeline;
typedef struct b3hMlp b3hMlp;

#define B3H_BROADPHASE_GRID 0
#define B3H_BROADPHASE_SAP 1
#define B3H_BROADPHASE_PLBVH 2

/* Text of the last error on this thread ("" if none). */
const char* b3h_last_error(void);
/* Library version string. */
const char* b3h_version(void);
/* Default capacities: b3Config::b3Config, src/Bullet3Collision/NarrowPhaseCollision/b3Config.h:21-36 */
int b3h_config_default(b3hConfig* cfg);
/* Default solver parameters (the reference's hard-coded values, see b3hip_types.h) */
int b3h_solver_params_default(b3hSolverParams* p);

/* ---------------- b3GpuNarrowPhase ---------------- */
/* b3GpuNarrowPhase::b3GpuNarrowPhase(ctx, device, queue, config), b3GpuNarrowPhase.cpp:18-93 */
int b3h_np_create(const b3hConfig* cfg, void* hipStream, b3hNarrowPhase** out);
/* b3GpuNarrowPhase::~b3GpuNarrowPhase, b3GpuNarrowPhase.cpp:95-140 */
int b3h_np_destroy(b3hNarrowPhase* np);
/* b3GpuNarrowPhase::registerConvexHullShape(const float* vertices, int strideInBytes,
 * int numVertices, const float* scaling), b3GpuNarrowPhase.cpp:298-319; builds
 * the hull, merges coplanar faces (b3ConvexUtility.cpp:26-248) and writes the
 * new collidable index. */
int b3h_np_register_convex_hull_shape(b3hNarrowPhase* np, const float* vertices, int strideInBytes, int numVertices,
									  const float* scaling3, int* outCollidableIndex);
/* b3GpuNarrowPhase::registerConvexHullShape(b3ConvexUtility* utilPtr),
 * b3GpuNarrowPhase.h:39, b3GpuNarrowPhase.cpp:321-368: registers a polyhedron
 * whose vertices, faces and unique edges the caller computed (the reference
 * demo's route, GpuConvexScene.cpp:231,243), tables taken as given. */
int b3h_np_register_convex_polyhedron(b3hNarrowPhase* np, const b3hConvexUtility* convex, int* outCollidableIndex);
/* b3GpuNarrowPhase::registerSphereShape(radius) / registerPlaneShape(normal,
 * constant), b3GpuNarrowPhase.h:29-30, b3GpuNarrowPhase.cpp:159-232: a sphere
 * collidable (radius in the collidable), a plane n.x = constant (stored as a
 * face).  Pairs with a sphere or a plane get the reference's primitive
 * contacts (primitiveContactsKernel, primitiveContacts.cl:793-977: plane-hull,
 * plane-sphere, sphere-hull, sphere-sphere). */
int b3h_np_register_sphere_shape(b3hNarrowPhase* np, float radius, int* outCollidableIndex);
int b3h_np_register_plane_shape(b3hNarrowPhase* np, const float* normal3, float planeConstant, int* outCollidableIndex);
/* b3GpuNarrowPhase::registerCompoundShape(b3AlignedObjectArray<b3GpuChildShape>*),
 * b3GpuNarrowPhase.cpp:370-517 (b3GpuNarrowPhase.h:33): a compound of
 * registered convex hulls (children[i].m_shapeIndex: the child's collidable)
 * placed by child transforms.  Its local AABB bounds the transformed child
 * AABBs; a quantized AABB tree over the children serves the compound-compound
 * pair search (findCompoundPairsKernel, sat.cl:945-1130).  Compound pairs get
 * their contacts from the compound SAT and clipping (processCompoundPairsKernel,
 * sat.cl:802-925; clipCompoundsHullHullKernel, satClipHullContacts.cl:1003-1130)
 * and, against a plane or a sphere, from the primitive routines
 * (processCompoundPairsPrimitivesKernel, primitiveContacts.cl:980-1090). */
int b3h_np_register_compound_shape(b3hNarrowPhase* np, const b3hGpuChildShape* children, int numChildren,
								   int* outCollidableIndex);
/* b3GpuNarrowPhase::registerRigidBody(collidableIndex, mass, position, orientation,
 * aabbMin, aabbMax, writeToGpu), b3GpuNarrowPhase.cpp:816-908 */
int b3h_np_register_rigid_body(b3hNarrowPhase* np, int collidableIndex, float mass, const float* position3,
							   const float* orientation4, const float* aabbMin3, const float* aabbMax3, int writeToGpu,
							   int* outBodyIndex);
/* b3GpuNarrowPhase::writeAllBodiesToGpu, b3GpuNarrowPhase.cpp:915-944 */
int b3h_np_write_all_bodies_to_gpu(b3hNarrowPhase* np);
/* b3GpuNarrowPhase::reset, b3GpuNarrowPhase.cpp:946-963 */
int b3h_np_reset(b3hNarrowPhase* np);
/* b3GpuNarrowPhase::readbackAllBodiesToCpu, b3GpuNarrowPhase.cpp:965-968; copies
 * the device bodies into the host mirror and, if dst != NULL, into dst. */
int b3h_np_readback_all_bodies_to_cpu(b3hNarrowPhase* np, b3hRigidBodyData* dst);
/* b3GpuNarrowPhase::setObjectTransformCpu / setObjectVelocityCpu /
 * getObjectTransformFromCpu, b3GpuNarrowPhase.cpp:969-1013 (host mirror) */
int b3h_np_set_object_transform_cpu(b3hNarrowPhase* np, const float* position3, const float* orientation4, int bodyIndex);
int b3h_np_set_object_velocity_cpu(b3hNarrowPhase* np, const float* linVel3, const float* angVel3, int bodyIndex);
int b3h_np_get_object_transform_from_cpu(b3hNarrowPhase* np, float* position4, float* orientation4, int bodyIndex);
/* b3GpuNarrowPhase::computeContacts(pairs, numPairs, aabbsWorldSpace, numObjects),
 * b3GpuNarrowPhase.cpp:751-809.  pairs: device b3hInt4[numPairs]; .z receives the
 * contact slot or -1. */
int b3h_np_compute_contacts(b3hNarrowPhase* np, b3hInt4* pairs, int numPairs, const b3hAabb* aabbsWorldSpace,
							int numObjects);
/* getBodiesGpu/getNumRigidBodies, getBodyInertiasGpu, getCollidablesGpu,
 * getContactsGpu/getNumContactsGpu, getStatic0Index: b3GpuNarrowPhase.h:55-83 */
int b3h_np_get_bodies_gpu(b3hNarrowPhase* np, b3hRigidBodyData** dev, int* numBodies);
int b3h_np_get_body_inertias_gpu(b3hNarrowPhase* np, b3hInertiaData** dev, int* numBodies);
int b3h_np_get_collidables_gpu(b3hNarrowPhase* np, b3hCollidable** dev, int* numCollidables);
int b3h_np_get_contacts_gpu(b3hNarrowPhase* np, b3hContact4Data** dev, int* numContacts);
int b3h_np_get_static0_index(b3hNarrowPhase* np, int* index);
int b3h_np_get_num_rigid_bodies(b3hNarrowPhase* np, int* n);
/* getLocalSpaceAabb(collidableIndex), b3GpuNarrowPhase.cpp:811-814 (host) */
int b3h_np_get_local_space_aabb(b3hNarrowPhase* np, int collidableIndex, b3hAabb* out);
/* Host copies of the registered shape tables (collidables, hull records, hull
 * vertices, unique edges, faces, face indices, local AABBs): counts[7] are
 * written; arrays are filled when non-NULL (sized by a first call). */
int b3h_np_get_shape_tables(b3hNarrowPhase* np, int* counts7, b3hCollidable* collidables,
							b3hConvexPolyhedronData* convexes, b3hFloat4* vertices, b3hFloat4* uniqueEdges,
							b3hGpuFace* faces, int* indices, b3hAabb* localAabbs);
/* Host copies of the compound tables (children, child-tree records, tree
 * nodes, subtree headers): counts[4] are written; arrays are filled when
 * non-NULL.  The tables of registerCompoundShape, b3GpuNarrowPhase.cpp:381-513. */
int b3h_np_get_compound_tables(b3hNarrowPhase* np, int* counts4, b3hGpuChildShape* children, b3hBvhInfo* infos,
							   b3hQuantizedBvhNode* nodes, b3hBvhSubtreeInfo* subtrees);
/* Replace the shape tables wholesale (same arrays and counts as
 * b3h_np_get_shape_tables; counts7[6] must equal counts7[0]).  Every hull's
 * offsets and face indices are checked against the tables.  No counterpart in
 * the reference, whose kernel tests hand recorded tables straight to a kernel
 * (test/OpenCL/AllBullet3Kernels/testExecuteBullet3NarrowphaseKernels.cpp:
 * 380-424); the tables reach the device with writeAllBodiesToGpu. */
int b3h_np_load_shape_tables(b3hNarrowPhase* np, const int* counts7, const b3hCollidable* collidables,
							 const b3hConvexPolyhedronData* convexes, const b3hFloat4* vertices,
							 const b3hFloat4* uniqueEdges, const b3hGpuFace* faces, const int* indices,
							 const b3hAabb* localAabbs);
/* Host copy of the registered inertias. */
int b3h_np_get_inertias_cpu(b3hNarrowPhase* np, b3hInertiaData* dst, int n);
/* Narrowphase tuning: lanes cooperating on one pair in the SAT face pass: 32
 * (default, two pairs per wavefront), 64 or 16 (one or four pairs per
 * wavefront); other powers of two up to 64 select 64.  Results do not depend
 * on it. */
int b3h_np_set_pairs_per_group(b3hNarrowPhase* np, int lanesPerPair);
/* Per-pair MPR result of the last computeContacts (device b3hFloat4[3 * numPairs]:
 * direction, point, (depth, 1 if MPR found penetration else 0, 0, 0)); the
 * reference's mprPenetrationKernel appends a contact for exactly the pairs
 * marked 1 (mpr.cl:14-90). */
int b3h_np_get_mpr_debug(b3hNarrowPhase* np, b3hFloat4** mprDev);
/* Per-pair separating-axis state and normal of the last computeContacts (device). */
int b3h_np_get_sat_debug(b3hNarrowPhase* np, int** sepStateDev, b3hFloat4** sepNormalDev);

/* ---------------- b3GpuBroadphaseInterface ---------------- */
/* b3GpuGridBroadphase / b3GpuSapBroadphase / b3GpuParallelLinearBvhBroadphase
 * constructors (b3GpuGridBroadphase.cpp:24-81, b3GpuSapBroadphase.cpp:20-170,
 * b3GpuParallelLinearBvhBroadphase.cpp:16-24), kind B3H_BROADPHASE_GRID / _SAP /
 * _PLBVH.  The BVH broadphase rebuilds a linear BVH over the small proxies on
 * every calculateOverlappingPairs and pairs each small proxy with the later
 * ones in Morton order as (earlier, later), then with the large proxies as
 * (large, small), unnormalised as the reference writes them
 * (b3GpuParallelLinearBvh.cpp:120-355). */
int b3h_bp_create(int kind, const b3hConfig* cfg, void* hipStream, b3hBroadphase** out);
int b3h_bp_destroy(b3hBroadphase* bp);
/* Grid cell size, grid dimensions (powers of two) and bucket cap; the reference
 * hard-codes 3 m, 128^3 and 256 (b3GpuGridBroadphase.cpp:22,38-49). */
int b3h_bp_set_grid_params(b3hBroadphase* bp, const float* cellSize3, const int* dims3, int maxBodiesPerCell);
/* createProxy / createLargeProxy(aabbMin, aabbMax, userPtr, group, mask),
 * b3GpuBroadphaseInterface.h:20-21 */
int b3h_bp_create_proxy(b3hBroadphase* bp, const float* aabbMin3, const float* aabbMax3, int userPtr, int group,
						int mask);
int b3h_bp_create_large_proxy(b3hBroadphase* bp, const float* aabbMin3, const float* aabbMax3, int userPtr,
							  int group, int mask);
/* writeAabbsToGpu, b3GpuBroadphaseInterface.h:27 */
/* createProxy for n AABBs at once (host arrays of n*3 floats; userPtr[i] or i
 * when userPtrs is NULL).  No counterpart in the reference, which registers
 * one proxy per call (b3GpuBroadphaseInterface.h:23). */
int b3h_bp_create_proxies(b3hBroadphase* bp, int n, const float* aabbMins3, const float* aabbMaxs3, const int* userPtrs);
int b3h_bp_write_aabbs_to_gpu(b3hBroadphase* bp);
/* calculateOverlappingPairs(maxPairs), b3GpuBroadphaseInterface.h:24 (no host sync) */
int b3h_bp_calculate_overlapping_pairs(b3hBroadphase* bp, int maxPairs);
/* calculateOverlappingPairsHost(maxPairs), b3GpuBroadphaseInterface.h:25: the
 * same pair set by brute force on the host from the device AABBs
 * (b3GpuGridBroadphase.cpp:260-292: every proxy against every later one;
 * b3GpuSapBroadphase.cpp:862-981: small x small, then small x large), pairs
 * normalised to (min, max), the first maxPairs kept, written to the device
 * pair buffer (syncs). */
int b3h_bp_calculate_overlapping_pairs_host(b3hBroadphase* bp, int maxPairs);
/* getAllAabbsGPU (device b3hAabb[n]) / getAllAabbsCPU (host copy of the proxy
 * AABBs, capacity entries; NULL host only reports n) / getSmallAabbIndicesGPU /
 * getLargeAabbIndicesGPU (device int[n]), b3GpuBroadphaseInterface.h:34-39 */
int b3h_bp_get_all_aabbs_gpu(b3hBroadphase* bp, b3hAabb** dev, int* n);
int b3h_bp_get_all_aabbs_cpu(b3hBroadphase* bp, b3hAabb* host, int capacity, int* n);
int b3h_bp_get_small_aabb_indices_gpu(b3hBroadphase* bp, int** dev, int* n);
int b3h_bp_get_large_aabb_indices_gpu(b3hBroadphase* bp, int** dev, int* n);
/* getAabbBufferWS (device b3hAabb[numAabbs]), getNumOverlap (syncs),
 * getOverlappingPairBuffer (device b3hInt4[]): b3GpuBroadphaseInterface.h:30-34 */
int b3h_bp_get_aabb_buffer_ws(b3hBroadphase* bp, b3hAabb** dev, int* numAabbs);
int b3h_bp_get_num_overlap(b3hBroadphase* bp, int* n);
int b3h_bp_get_overlapping_pair_buffer(b3hBroadphase* bp, b3hInt4** dev);
/* Number of pairs found before clamping to maxPairs (syncs). */
int b3h_bp_get_total_found(b3hBroadphase* bp, int* n);
/* SAP axis chosen by the last calculateOverlappingPairs (syncs). */
int b3h_bp_get_sap_axis(b3hBroadphase* bp, int* axis);
/* BVH broadphase: root node of the last build (internal nodes carry the sign
 * bit, as m_rootNodeIndex, b3GpuParallelLinearBvh.h:70) and whether a
 * traversal ran out of stack (0: never; the stack covers the deepest tree 64-bit
 * keys allow) -- syncs. */
int b3h_bp_get_bvh_info(b3hBroadphase* bp, int* rootNode, int* stackOverflow);
/* Host access to the proxy lists (copies). */
int b3h_bp_get_proxy_maps(b3hBroadphase* bp, int* counts2, int* smallMap, int* largeMap);
/* Debug view of the grid's last (cell hash, small-proxy index) arrays after the
 * sort, in device memory (n entries each) -- b3GpuGridBroadphase::m_hashGpu,
 * b3GpuGridBroadphase.h:21; no public accessor there. */
int b3h_bp_get_grid_sort(b3hBroadphase* bp, const unsigned** keysDev, const unsigned** valsDev, int* n);

/* ---------------- b3GpuRigidBodyPipeline ---------------- */
/* b3GpuRigidBodyPipeline(ctx, device, queue, narrowphase, broadphaseSap,
 * broadphaseDbvt, config), b3GpuRigidBodyPipeline.cpp:65-114.  np and bp stay
 * owned by the caller. */
int b3h_pipeline_create(b3hNarrowPhase* np, b3hBroadphase* bp, const b3hConfig* cfg, void* hipStream,
						b3hPipeline** out);
int b3h_pipeline_destroy(b3hPipeline* p);
/* stepSimulation(deltaTime), b3GpuRigidBodyPipeline.cpp:221-463 */
int b3h_pipeline_step_simulation(b3hPipeline* p, float deltaTime);
/* integrate(timeStep), b3GpuRigidBodyPipeline.cpp:465-498 */
int b3h_pipeline_integrate(b3hPipeline* p, float timeStep);
/* setupGpuAabbsFull, b3GpuRigidBodyPipeline.cpp:500-575 */
int b3h_pipeline_setup_gpu_aabbs_full(b3hPipeline* p);
/* b3GpuPgsContactSolver::solveContacts on the narrowphase's current contacts
 * (b3GpuPgsContactSolver.cpp:568-1103) */
int b3h_pipeline_solve_contacts(b3hPipeline* p);
/* registerPhysicsInstance(mass, position, orientation, collidableIndex, userIndex,
 * writeInstanceToGpu), b3GpuRigidBodyPipeline.cpp:603-669 */
int b3h_pipeline_register_physics_instance(b3hPipeline* p, float mass, const float* position3,
										   const float* orientation4, int collidableIndex, int userIndex,
										   int writeInstanceToGpu, int* outBodyIndex);
/* registerConvexPolyhedron(b3ConvexUtility*), b3GpuRigidBodyPipeline.h:40 (declared
 * there without a definition): registers through the narrowphase as
 * b3h_np_register_convex_polyhedron does. */
int b3h_pipeline_register_convex_polyhedron(b3hPipeline* p, const b3hConvexUtility* convex, int* outCollidableIndex);
/* writeAllInstancesToGpu, b3GpuRigidBodyPipeline.cpp:597-601 */
int b3h_pipeline_write_all_instances_to_gpu(b3hPipeline* p);
/* setGravity, b3GpuRigidBodyPipeline.cpp:589-592 */
int b3h_pipeline_set_gravity(b3hPipeline* p, const float* gravity3);
/* reset, b3GpuRigidBodyPipeline.cpp:141-147 */
int b3h_pipeline_reset(b3hPipeline* p);
/* getBodyBuffer / getNumBodies, b3GpuRigidBodyPipeline.cpp:577-587 */
int b3h_pipeline_get_body_buffer(b3hPipeline* p, b3hRigidBodyData** dev);
int b3h_pipeline_get_num_bodies(b3hPipeline* p, int* n);
/* castRays(rays, hitResults), b3GpuRigidBodyPipeline.h:63,
 * b3GpuRigidBodyPipeline.cpp:671-677 -> b3GpuRaycast::castRays, b3GpuRaycast.cpp:
 * 245-373: host arrays of numRays rays and hit records (in: m_hitResult2 = a body
 * the ray ignores, -1 for none; out: m_hitFraction (1 if nothing is hit),
 * m_hitBody, m_hitPoint, m_hitNormal of the nearest hit, the last three left as
 * given when nothing is hit; a hull's normal is its face normal in the body's
 * own frame, as the reference leaves it, rayCastKernels.cl:191-193,383-389).  Candidates are the bodies whose broadphase AABB
 * (as the last step computed it, before its integration) the segment crosses,
 * found through a linear BVH rebuilt per call; convex hulls and spheres are
 * tested (rayCastKernels.cl:155-240).  Syncs. */
int b3h_pipeline_cast_rays(b3hPipeline* p, const b3hRayInfo* rays, b3hRayHit* hitResults, int numRays);
/* The same on device arrays, queued on the pipeline stream without a host
 * sync.  No counterpart in the reference, whose castRays copies host arrays. */
int b3h_pipeline_cast_rays_gpu(b3hPipeline* p, const b3hRayInfo* raysDev, b3hRayHit* hitResultsDev, int numRays);
/* Root of the ray caster's last BVH and whether a traversal ran out of stack
 * (0: never) -- syncs.  For tests. */
int b3h_pipeline_get_ray_bvh_info(b3hPipeline* p, int* rootNode, int* stackOverflow);
/* Joints.  createPoint2PointConstraint(bodyA, bodyB, pivotInA, pivotInB,
 * breakingThreshold) / createFixedConstraint(..., relTargetAB, ...),
 * b3GpuRigidBodyPipeline.h:56-57, .cpp:186-219: append an enabled constraint to
 * the pipeline's list and write its uid.  removeConstraintByUid (:58,
 * .cpp:159-184): the last constraint takes the removed one's place.
 * copyConstraintsToHost (:52, .cpp:594-596): the device's enabled flags (a
 * constraint is disabled when a row impulse reaches its breaking threshold)
 * into the host list; get_constraints copies that list out (capacity entries;
 * host NULL reports the count).  The list reaches the device at
 * writeAllInstancesToGpu or, when it changed since, at the next step; every
 * step then solves the joints (b3GpuPgsConstraintSolver::solveJoints: 4
 * iterations, dt 1/60, erp 0.2, b3GpuPgsConstraintSolver.cpp:927-937) after the
 * narrowphase and before the contact solve, as stepSimulation does
 * (b3GpuRigidBodyPipeline.cpp:361-372).  Changes to the list pull the device
 * flags first, so breaking survives them. */
int b3h_pipeline_create_point2point_constraint(b3hPipeline* p, int bodyA, int bodyB, const float* pivotInA3,
											   const float* pivotInB3, float breakingThreshold, int* uidOut);
int b3h_pipeline_create_fixed_constraint(b3hPipeline* p, int bodyA, int bodyB, const float* pivotInA3,
										 const float* pivotInB3, const float* relTargetAB4, float breakingThreshold,
										 int* uidOut);
int b3h_pipeline_remove_constraint_by_uid(b3hPipeline* p, int uid);
int b3h_pipeline_copy_constraints_to_host(b3hPipeline* p);
int b3h_pipeline_get_constraints(b3hPipeline* p, b3hGpuGenericConstraint* host, int capacity, int* n);
/* Solver parameters (iterations, friction, dt, Baumgarte, drift, cell size,
 * damping, warm start); the reference hard-codes them. */
int b3h_pipeline_set_solver_params(b3hPipeline* p, const b3hSolverParams* prm);
int b3h_pipeline_get_solver_params(b3hPipeline* p, b3hSolverParams* prm);
/* Solver internals of the last solve (device pointers): contacts in solve
 * order, constraints, per-cell counts / offsets / batch counts. */
int b3h_pipeline_get_solver_debug(b3hPipeline* p, b3hContact4Data** sortedContacts,
								  b3hContactConstraint4** constraints, int** cellCounts, int** cellOffsets,
								  int** batchSizes, int* numConstraints);
/* Upper bound on the bodies whose velocities one PGS work-group (one part of
 * a solver cell) keeps in LDS; a cell with a larger part takes the global
 * read-modify-write walk (negative: the device's LDS budget; 0: every cell
 * takes the global walk).  No counterpart in the reference (the reference
 * always walks global memory); for tests and tuning. */
int b3h_pipeline_set_pgs_body_cache(b3hPipeline* p, int maxBodiesPerCell);
/* Dependency levels per solver cell of the last solve (host int[256]; -1: the
 * cell took the global walk) -- syncs.  No counterpart in the reference. */
int b3h_pipeline_get_pgs_levels(b3hPipeline* p, int* levels256);
/* Work-groups each solver cell's sweep was split into in the last solve (host
 * int[256]: the cell's copies in space that share no body are solved apart, up
 * to 8 parts) and, if largest256 != NULL, the constraint count of its largest
 * part -- syncs.  No counterpart in the reference (its sweep solves a cell on
 * one work-group). */
int b3h_pipeline_get_pgs_parts(b3hPipeline* p, int* parts256, int* largest256);
/* Upper bound on those parts (1..8, default 8; 1 solves every cell on one
 * work-group, as the reference does).  Results do not depend on it. */
int b3h_pipeline_set_pgs_max_parts(b3hPipeline* p, int maxParts);
/* Autotuner for the two schedule knobs above: times `steps` eager steps of
 * length dt from the current state for each narrowphase lanes-per-pair (64,
 * 32, 16) and PGS parts cap (8, 4, 1), restoring the state (b3h_pipeline_
 * save_state image) before each trial and at the end, and keeps the fastest;
 * writes the choice and its ms per step (host pointers, each may be NULL).
 * Results do not depend on the knobs, so stepping continues bit for bit as
 * without the trials.  No counterpart in the reference. */
int b3h_pipeline_autotune(b3hPipeline* p, float dt, int steps, int* lanesOut, int* partsOut, float* msOut);
/* Per-sweep residuals of the last solve (host, 2*iterations floats: normal sweeps
 * then friction sweeps); requires b3h_pipeline_set_track_residual(p, 1). */
int b3h_pipeline_set_track_residual(b3hPipeline* p, int enable);
int b3h_pipeline_get_residuals(b3hPipeline* p, float* out, int n);
/* Per-sweep relative L2 residuals of the last solve: sqrt(sum (impulse change)^2 /
 * sum impulse^2) over the rows the sweep updates (host, 2*iterations floats,
 * normal sweeps then friction sweeps); requires residual tracking.  No
 * counterpart in the reference, which computes no residual. */
int b3h_pipeline_get_residuals_l2(b3hPipeline* p, float* out, int n);
/* Initial impulses for the next solve (device float[6*numContacts] in solve
 * order: 4 normal rows then 2 friction rows per constraint); NULL clears.
 * Takes precedence over the warm-start mode. */
int b3h_pipeline_set_warm_start_impulses(b3hPipeline* p, const float* dev, int numConstraints);
/* Warm start (b3hSolverParams::m_warmStartMode): 0 starts every solve from
 * zero impulses as the reference does (b3ConvertConstraint4.h:73,107,132);
 * 1 starts from the previous step's final impulses of the same body pair
 * (normal rows by nearest contact point within 0.1, friction as a world
 * vector projected on the new tangents); 2 corrects those with the network set
 * here (24 features -> 3 outputs, see b3h_mlp_create).  Only the initial
 * iterate changes; the solve's fixed point does not.  The network must outlive
 * its use by the pipeline. */
int b3h_pipeline_set_warm_start_model(b3hPipeline* p, b3hMlp* net);
/* Device buffers of the last warm-started solve, in solve order (NULL until a
 * warm-start mode or the contact log is used): initial impulses float[6] per
 * constraint; matched previous impulses float[8] per constraint (4 normal, 2
 * friction, pair matched, points); feature rows float[4][24] per constraint
 * (learned mode or logging).  For tests. */
int b3h_pipeline_get_warm_start_debug(b3hPipeline* p, float** warmDev, float** prevDev, float** featuresDev);
/* Forget the impulse history (the next solve finds no previous frame). */
int b3h_pipeline_clear_warm_start_history(b3hPipeline* p);
/* State dump / load.  The image holds the body states and inertias, gravity,
 * the solver parameters and the warm-start history of the last solve, so that
 * stepping after a load repeats bit for bit what stepping after the save did.
 * Shapes and registrations are not part of it: load into a pipeline with the
 * same bodies registered (same count and collidable per body).  save_state
 * with host NULL only reports the size in *bytesOut; otherwise the image is
 * written if capacityBytes suffices (else B3H_ERR_CAPACITY).  Both synchronise
 * the pipeline stream.  The reference keeps no state image; its nearest calls
 * are b3GpuNarrowPhase::readbackAllBodiesToCpu / writeAllBodiesToGpu
 * (b3GpuNarrowPhase.h:45-47), which move the bodies only. */
int b3h_pipeline_save_state(b3hPipeline* p, void* host, size_t capacityBytes, size_t* bytesOut);
int b3h_pipeline_load_state(b3hPipeline* p, const void* host, size_t bytes);
/* Batched scenes: many independent scenes in one pipeline, stepped by the same
 * launches.  Scene k owns the bodies [bodyOffsets[k], bodyOffsets[k+1]) (host
 * array of numScenes + 1 ints from 0 to the body count); the broadphase then
 * drops every pair whose bodies belong to different scenes, so scenes never
 * interact even where they overlap in space (keep them apart anyway: the grid's
 * cells fill with every scene's bodies).  numScenes 0 returns to one world.
 * Registering bodies afterwards requires setting the scenes again (the step
 * reports B3H_ERR_STATE otherwise).  No counterpart in the reference, whose
 * batched demos place scenes apart in one world. */
int b3h_pipeline_set_scenes(b3hPipeline* p, int numScenes, const int* bodyOffsets);
/* Contact manifolds of the last step per scene (host int[numScenes]) -- syncs. */
int b3h_pipeline_get_scene_contact_counts(b3hPipeline* p, int* counts, int numScenes);
/* Contact logging for training the network: after every solve, each contact
 * point appends its 24 features (as computed before the solve) to
 * featuresDev[rows][24] and its 3 targets (the solve's final impulses relative
 * to the previous-frame guess, in the network's output units) to
 * targetsDev[rows][3], until capacityRows rows are held; capacityRows 0 turns
 * logging off.  Setting the log resets its row count. */
int b3h_pipeline_set_contact_log(b3hPipeline* p, float* featuresDev, float* targetsDev, int capacityRows);
int b3h_pipeline_get_contact_log_count(b3hPipeline* p, int* rows);
/* Stage timings of the last eager step in ms (enable first; adds HIP events on
 * the pipeline stream).  B3H_NUM_STAGES values, in order: world AABBs,
 * broadphase, narrowphase MPR, narrowphase SAT and clipping, narrowphase
 * compaction, solver set-up (ordering, batching, constraints, warm start),
 * normal sweeps, friction sweeps, integrate.  No counterpart in the reference. */
#define B3H_NUM_STAGES 9
int b3h_pipeline_enable_timing(b3hPipeline* p, int enable);
int b3h_pipeline_get_stage_times(b3hPipeline* p, float* msOut);
/* Overflow flags of the last step (pairs, contacts) -- syncs. */
int b3h_pipeline_get_overflow(b3hPipeline* p, int* pairsOverflow, int* contactsOverflow);
/* Capture stepSimulation(dt) into a HIP graph and replay it `count` times. */
int b3h_pipeline_step_graph(b3hPipeline* p, float deltaTime, int count);

/* ---------------- parallel primitives ---------------- */
/* Stable sort of n (key, value) pairs (b3hSortData, device) by the low
 * sortBits bits of the key; b3RadixSort32CL::execute(keyValues, sortBits),
 * src/Bullet3OpenCL/ParallelPrimitives/b3RadixSort32CL.h:84.  Keys and values
 * are taken as separate device arrays of uint32. */
int b3h_radix_sort32(unsigned* keysDev, unsigned* valuesDev, int n, int sortBits);
/* Allocation-free variant of b3h_radix_sort32: the same stable sort (same
 * kernels, same result), with its scratch in a fixed device array inside the
 * library instead of hipMalloc'd buffers, and no allocating call of any kind.
 * No counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  n must not exceed B3H_NOALLOC_SORT_MAX (B3H_ERR_CAPACITY
 * otherwise); each device holds about 8.5 MB of scratch until the process
 * ends.  Calls on one device share that scratch: calls that overlap on one
 * device race.  Synchronises the device like b3h_radix_sort32. */
#define B3H_NOALLOC_SORT_MAX (1 << 20)
int b3h_radix_sort32_noalloc(unsigned* keysDev, unsigned* valuesDev, int n, int sortBits);
/* Exclusive prefix sum of n ints (device) into dst (n+1 ints, dst[n] = total);
 * b3PrefixScanCL::execute(src, dst, n, sum), b3PrefixScanCL.h:24. */
int b3h_prefix_scan(const int* srcDev, int* dstDev, int n, int* sumOut);

/* ---------------- learned warm start: the network ----------------
 * No counterpart in the reference: its solver starts every step from zero
 * impulses (src/Bullet3OpenCL/RigidBody/kernels/solverSetup.cl via
 * b3ConvertConstraint4.h:73,107,132).  A 3-layer MLP inDim -> 128 -> 128 ->
 * outDim (inDim, outDim <= 32), GELU (tanh form), bf16 weights and activations
 * with fp32 accumulation on the matrix cores; fp32 master weights and Adam
 * (beta 0.9/0.999, eps 1e-8) on the device.  Parameters cross the boundary in
 * one packed fp32 array: w1[in][128] b1[128] w2[128][128] b2[128]
 * w3[128][out] b3[out] (row = input unit, column = output unit).  Device
 * pointers are row-major with leading dimensions ldx/ldy/ldt (in floats). */
#define B3H_MLP_HIDDEN 128
int b3h_mlp_create(int inDim, int outDim, int maxTrainRows, unsigned seed, void* hipStream, b3hMlp** out);
int b3h_mlp_destroy(b3hMlp* m);
int b3h_mlp_get_dims(const b3hMlp* m, int* inDim, int* hidden, int* outDim);
int b3h_mlp_num_params(const b3hMlp* m, int* n);
/* host float[num_params]; set_params also resets the Adam state */
int b3h_mlp_set_params(b3hMlp* m, const float* host);
int b3h_mlp_get_params(b3hMlp* m, float* host);
/* Input normalisation, part of the model: every input feature is mapped to
 * (x - shift) * scale before the first layer (host float[inDim] each; the
 * default is shift 0, scale 1).  Used in inference and training alike. */
int b3h_mlp_set_input_normalization(b3hMlp* m, const float* shift, const float* scale);
int b3h_mlp_get_input_normalization(b3hMlp* m, float* shift, float* scale);
/* y[rows][outDim] = net(x[rows][inDim]), asynchronous on the network's stream */
int b3h_mlp_forward(b3hMlp* m, const float* xDev, int ldx, int rows, float* yDev, int ldy);
/* mean squared error over rows x outDim and its gradient (host float[num_params],
 * may be NULL) -- syncs; rows <= maxTrainRows */
int b3h_mlp_loss_gradients(b3hMlp* m, const float* xDev, int ldx, const float* tDev, int ldt, int rows,
						   float* gradsHost, float* lossOut);
/* `steps` Adam steps on minibatches of batchRows rows drawn with replacement
 * from the rows x / t (a hash of seed, step and row picks each one); the loss
 * of every step goes to lossHistory (host float[steps], may be NULL) -- syncs */
int b3h_mlp_train(b3hMlp* m, const float* xDev, int ldx, const float* tDev, int ldt, int rows, int batchRows, int steps,
				  float learningRate, unsigned seed, float* lossHistory);

/* ---------------- host-only utility ---------------- */
/* Polyhedral features of a convex point cloud exactly as registerConvexHullShape
 * stores them (b3ConvexUtility::initializePolyhedralFeatures,
 * src/Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.cpp:26-248); no
 * device needed.  counts4 = {numVertices, numFaces, numIndices, numUniqueEdges};
 * arrays (capacity `cap` entries each) are filled when non-NULL: vertices and
 * planes and edges as float4, faceIndexCounts and indices as int. */
int b3h_compute_polyhedral_features(const float* vertices, int strideInBytes, int numVertices, int cap, int* counts4,
									float* vertices4, float* planes4, int* faceIndexCounts, int* indices, float* edges4);

/* What registerPhysicsInstance computes for one body, on the host (no device
 * needed): its broadphase AABB -- the local AABB grown by the 0.01 margin and
 * transformed (b3GpuRigidBodyPipeline.cpp:605-618, b3TransformAabb) -- and its
 * inertia record from that AABB (b3GpuNarrowPhase::registerRigidBody,
 * b3GpuNarrowPhase.cpp:860-901).  The registration entry points use it. */
int b3h_compute_body_registration(const float* localAabbMin3, const float* localAabbMax3, float mass,
								  const float* position3, const float* orientation4, float* aabbMin3, float* aabbMax3,
								  b3hInertiaData* inertia);
/* The tables registerCompoundShape computes for a compound (b3GpuNarrowPhase.cpp:
 * 391-513: the local AABB over the transformed child AABBs, and the quantized
 * child tree of b3QuantizedBvh), on the host, without registering it:
 * childLocalAabbs are the children's local AABBs; the tree record's node and
 * subtree offsets are 0; 2 * numChildren nodes are written (the last one is
 * unused).  The registration entry uses it. */
int b3h_compute_compound_tables(const b3hGpuChildShape* children, int numChildren, const b3hAabb* childLocalAabbs,
								b3hAabb* aabbOut, b3hBvhInfo* infoOut, b3hQuantizedBvhNode* nodesOut, int nodesCapacity,
								int* numNodesOut, b3hBvhSubtreeInfo* subtreesOut, int subtreesCapacity,
								int* numSubtreesOut);

/* ---------------- device memory helpers for bindings ---------------- */
int b3h_device_count(int* n);
int b3h_malloc(void** dev, size_t bytes);
int b3h_free(void* dev);
int b3h_memcpy_htod(void* dev, const void* host, size_t bytes);
int b3h_memcpy_dtoh(void* host, const void* dev, size_t bytes);
int b3h_memcpy_dtod(void* dst, const void* src, size_t bytes);
int b3h_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* B3HIP_H */
