/*
 * b3hip.hpp -- C++ host classes with the reference's names and methods, over
 * the C ABI of b3hip.h.  A program written against
 *   b3GpuNarrowPhase          src/Bullet3OpenCL/RigidBody/b3GpuNarrowPhase.h:9-99
 *   b3GpuBroadphaseInterface  src/Bullet3OpenCL/BroadphaseCollision/b3GpuBroadphaseInterface.h:12-40
 *   b3GpuGridBroadphase / b3GpuSapBroadphase
 *   b3GpuRigidBodyPipeline    src/Bullet3OpenCL/RigidBody/b3GpuRigidBodyPipeline.h:25-68
 * keeps its calls; only the constructors change (a HIP stream instead of the
 * OpenCL context/device/queue triple) and cl_mem becomes a typed device pointer.
 *
 * Error behaviour follows the reference: registration returns the new index or
 * -1 (b3h_last_error() says why); other failures are reported through
 * b3hip::lastStatus() and the error text instead of asserting.  Ownership also
 * follows the reference: the pipeline does not own the narrowphase or the
 * broadphase (b3GpuRigidBodyPipeline.cpp:116-139); the caller deletes them after
 * the pipeline.
 */
#ifndef B3HIP_HPP
#define B3HIP_HPP

#include <vector>

#include "b3hip.h"

namespace b3hip
{
inline int& lastStatusRef()
{
	static thread_local int s = B3H_OK;
	return s;
}
inline int lastStatus() { return lastStatusRef(); }
inline bool ok(int rc)
{
	lastStatusRef() = rc;
	return rc == B3H_OK;
}
}  // namespace b3hip

class b3GpuNarrowPhase
{
	b3hNarrowPhase* m_np = nullptr;

public:
	/* b3GpuNarrowPhase(cl_context, cl_device_id, cl_command_queue, const b3Config&), b3GpuNarrowPhase.h:25 */
	explicit b3GpuNarrowPhase(const b3hConfig& config, void* hipStream = nullptr)
	{
		b3hip::ok(b3h_np_create(&config, hipStream, &m_np));
	}
	~b3GpuNarrowPhase() { b3h_np_destroy(m_np); }
	b3GpuNarrowPhase(const b3GpuNarrowPhase&) = delete;
	b3GpuNarrowPhase& operator=(const b3GpuNarrowPhase&) = delete;
	bool isValid() const { return m_np != nullptr; }
	b3hNarrowPhase* handle() const { return m_np; }

	/* registerConvexHullShape(const float*, int, int, const float*), b3GpuNarrowPhase.h:40 */
	int registerConvexHullShape(const float* vertices, int strideInBytes, int numVertices, const float* scaling)
	{
		int idx = -1;
		return b3hip::ok(b3h_np_register_convex_hull_shape(m_np, vertices, strideInBytes, numVertices, scaling, &idx))
				   ? idx
				   : -1;
	}
	/* registerRigidBody(...), b3GpuNarrowPhase.h:42 */
	/* registerSphereShape / registerPlaneShape, b3GpuNarrowPhase.h:29-30 */
	int registerSphereShape(float radius)
	{
		int idx = -1;
		return b3hip::ok(b3h_np_register_sphere_shape(m_np, radius, &idx)) ? idx : -1;
	}
	int registerPlaneShape(const float* planeNormal, float planeConstant)
	{
		int idx = -1;
		return b3hip::ok(b3h_np_register_plane_shape(m_np, planeNormal, planeConstant, &idx)) ? idx : -1;
	}
	/* registerCompoundShape(b3AlignedObjectArray<b3GpuChildShape>*), b3GpuNarrowPhase.h:33 */
	int registerCompoundShape(const std::vector<b3hGpuChildShape>& childShapes)
	{
		int idx = -1;
		return b3hip::ok(b3h_np_register_compound_shape(m_np, childShapes.data(), (int)childShapes.size(), &idx)) ? idx
																												   : -1;
	}
	int registerRigidBody(int collidableIndex, float mass, const float* position, const float* orientation,
						  const float* aabbMin, const float* aabbMax, bool writeToGpu)
	{
		int idx = -1;
		return b3hip::ok(b3h_np_register_rigid_body(m_np, collidableIndex, mass, position, orientation, aabbMin,
													 aabbMax, writeToGpu ? 1 : 0, &idx))
				   ? idx
				   : -1;
	}
	void writeAllBodiesToGpu() { b3hip::ok(b3h_np_write_all_bodies_to_gpu(m_np)); }
	void reset() { b3hip::ok(b3h_np_reset(m_np)); }
	/* readbackAllBodiesToCpu(): the reference keeps the copy inside; here it is written to dst */
	void readbackAllBodiesToCpu(b3hRigidBodyData* dst) { b3hip::ok(b3h_np_readback_all_bodies_to_cpu(m_np, dst)); }
	void setObjectTransformCpu(const float* position, const float* orientation, int bodyIndex)
	{
		b3hip::ok(b3h_np_set_object_transform_cpu(m_np, position, orientation, bodyIndex));
	}
	void setObjectVelocityCpu(const float* linVel, const float* angVel, int bodyIndex)
	{
		b3hip::ok(b3h_np_set_object_velocity_cpu(m_np, linVel, angVel, bodyIndex));
	}
	bool getObjectTransformFromCpu(float* position, float* orientation, int bodyIndex) const
	{
		return b3hip::ok(b3h_np_get_object_transform_from_cpu(m_np, position, orientation, bodyIndex));
	}
	/* computeContacts(cl_mem pairs, int numPairs, cl_mem aabbsWorldSpace, int numObjects), b3GpuNarrowPhase.h:53 */
	void computeContacts(b3hInt4* pairs, int numPairs, const b3hAabb* aabbsWorldSpace, int numObjects)
	{
		b3hip::ok(b3h_np_compute_contacts(m_np, pairs, numPairs, aabbsWorldSpace, numObjects));
	}
	b3hRigidBodyData* getBodiesGpu()
	{
		b3hRigidBodyData* d = nullptr;
		b3hip::ok(b3h_np_get_bodies_gpu(m_np, &d, nullptr));
		return d;
	}
	int getNumBodiesGpu() const
	{
		int n = 0;
		b3hip::ok(b3h_np_get_num_rigid_bodies(m_np, &n));
		return n;
	}
	b3hInertiaData* getBodyInertiasGpu()
	{
		b3hInertiaData* d = nullptr;
		b3hip::ok(b3h_np_get_body_inertias_gpu(m_np, &d, nullptr));
		return d;
	}
	int getNumBodyInertiasGpu() const { return getNumBodiesGpu(); }
	b3hCollidable* getCollidablesGpu()
	{
		b3hCollidable* d = nullptr;
		b3hip::ok(b3h_np_get_collidables_gpu(m_np, &d, nullptr));
		return d;
	}
	int getNumCollidablesGpu() const
	{
		b3hCollidable* d = nullptr;
		int n = 0;
		b3hip::ok(b3h_np_get_collidables_gpu(m_np, &d, &n));
		return n;
	}
	b3hContact4Data* getContactsGpu()
	{
		b3hContact4Data* d = nullptr;
		b3hip::ok(b3h_np_get_contacts_gpu(m_np, &d, nullptr));
		return d;
	}
	int getNumContactsGpu() const
	{
		b3hContact4Data* d = nullptr;
		int n = 0;
		b3hip::ok(b3h_np_get_contacts_gpu(m_np, &d, &n));
		return n;
	}
	int getNumRigidBodies() const { return getNumBodiesGpu(); }
	int getStatic0Index() const
	{
		int i = -1;
		b3hip::ok(b3h_np_get_static0_index(m_np, &i));
		return i;
	}
	b3hAabb getLocalSpaceAabb(int collidableIndex) const
	{
		b3hAabb a = {};
		b3hip::ok(b3h_np_get_local_space_aabb(m_np, collidableIndex, &a));
		return a;
	}
};

/* b3GpuBroadphaseInterface, b3GpuBroadphaseInterface.h:12-40 (kind chosen at construction) */
class b3GpuBroadphaseInterface
{
protected:
	b3hBroadphase* m_bp = nullptr;
	explicit b3GpuBroadphaseInterface(int kind, const b3hConfig& config, void* hipStream)
	{
		b3hip::ok(b3h_bp_create(kind, &config, hipStream, &m_bp));
	}

public:
	virtual ~b3GpuBroadphaseInterface() { b3h_bp_destroy(m_bp); }
	b3GpuBroadphaseInterface(const b3GpuBroadphaseInterface&) = delete;
	b3GpuBroadphaseInterface& operator=(const b3GpuBroadphaseInterface&) = delete;
	bool isValid() const { return m_bp != nullptr; }
	b3hBroadphase* handle() const { return m_bp; }

	void createProxy(const float* aabbMin, const float* aabbMax, int userPtr, int collisionFilterGroup,
					 int collisionFilterMask)
	{
		b3hip::ok(b3h_bp_create_proxy(m_bp, aabbMin, aabbMax, userPtr, collisionFilterGroup, collisionFilterMask));
	}
	void createLargeProxy(const float* aabbMin, const float* aabbMax, int userPtr, int collisionFilterGroup,
						  int collisionFilterMask)
	{
		b3hip::ok(
			b3h_bp_create_large_proxy(m_bp, aabbMin, aabbMax, userPtr, collisionFilterGroup, collisionFilterMask));
	}
	void calculateOverlappingPairs(int maxPairs) { b3hip::ok(b3h_bp_calculate_overlapping_pairs(m_bp, maxPairs)); }
	void writeAabbsToGpu() { b3hip::ok(b3h_bp_write_aabbs_to_gpu(m_bp)); }
	b3hAabb* getAabbBufferWS()
	{
		b3hAabb* d = nullptr;
		b3hip::ok(b3h_bp_get_aabb_buffer_ws(m_bp, &d, nullptr));
		return d;
	}
	int getNumOverlap()
	{
		int n = 0;
		b3hip::ok(b3h_bp_get_num_overlap(m_bp, &n));
		return n;
	}
	b3hInt4* getOverlappingPairBuffer()
	{
		b3hInt4* d = nullptr;
		b3hip::ok(b3h_bp_get_overlapping_pair_buffer(m_bp, &d));
		return d;
	}
};

/* b3GpuGridBroadphase (cell size and grid dimensions settable; the reference hard-codes 3 m, 128^3) */
class b3GpuGridBroadphase : public b3GpuBroadphaseInterface
{
public:
	explicit b3GpuGridBroadphase(const b3hConfig& config, void* hipStream = nullptr)
		: b3GpuBroadphaseInterface(B3H_BROADPHASE_GRID, config, hipStream)
	{
	}
	void setGridParams(const float* cellSize3, const int* dims3, int maxBodiesPerCell)
	{
		b3hip::ok(b3h_bp_set_grid_params(m_bp, cellSize3, dims3, maxBodiesPerCell));
	}
};

class b3GpuSapBroadphase : public b3GpuBroadphaseInterface
{
public:
	explicit b3GpuSapBroadphase(const b3hConfig& config, void* hipStream = nullptr)
		: b3GpuBroadphaseInterface(B3H_BROADPHASE_SAP, config, hipStream)
	{
	}
};

/* b3GpuParallelLinearBvhBroadphase, b3GpuParallelLinearBvhBroadphase.h:22-66 */
class b3GpuParallelLinearBvhBroadphase : public b3GpuBroadphaseInterface
{
public:
	explicit b3GpuParallelLinearBvhBroadphase(const b3hConfig& config, void* hipStream = nullptr)
		: b3GpuBroadphaseInterface(B3H_BROADPHASE_PLBVH, config, hipStream)
	{
	}
};

class b3GpuRigidBodyPipeline
{
	b3hPipeline* m_p = nullptr;

public:
	/* b3GpuRigidBodyPipeline(ctx, device, queue, narrowphase, broadphase, dbvt, config),
	 * b3GpuRigidBodyPipeline.h:33 (the DBVT broadphase is not part of this path) */
	b3GpuRigidBodyPipeline(b3GpuNarrowPhase* narrowphase, b3GpuBroadphaseInterface* broadphase,
						   const b3hConfig& config, void* hipStream = nullptr)
	{
		b3hip::ok(b3h_pipeline_create(narrowphase->handle(), broadphase->handle(), &config, hipStream, &m_p));
	}
	virtual ~b3GpuRigidBodyPipeline() { b3h_pipeline_destroy(m_p); }
	b3GpuRigidBodyPipeline(const b3GpuRigidBodyPipeline&) = delete;
	b3GpuRigidBodyPipeline& operator=(const b3GpuRigidBodyPipeline&) = delete;
	bool isValid() const { return m_p != nullptr; }
	b3hPipeline* handle() const { return m_p; }

	void stepSimulation(float deltaTime) { b3hip::ok(b3h_pipeline_step_simulation(m_p, deltaTime)); }
	void integrate(float timeStep) { b3hip::ok(b3h_pipeline_integrate(m_p, timeStep)); }
	void setupGpuAabbsFull() { b3hip::ok(b3h_pipeline_setup_gpu_aabbs_full(m_p)); }
	/* registerPhysicsInstance(float, const float*, const float*, int, int, bool), b3GpuRigidBodyPipeline.h:49 */
	int registerPhysicsInstance(float mass, const float* position, const float* orientation, int collisionShapeIndex,
								int userData, bool writeInstanceToGpu)
	{
		int idx = -1;
		return b3hip::ok(b3h_pipeline_register_physics_instance(m_p, mass, position, orientation, collisionShapeIndex,
																 userData, writeInstanceToGpu ? 1 : 0, &idx))
				   ? idx
				   : -1;
	}
	void writeAllInstancesToGpu() { b3hip::ok(b3h_pipeline_write_all_instances_to_gpu(m_p)); }
	void setGravity(const float* grav) { b3hip::ok(b3h_pipeline_set_gravity(m_p, grav)); }
	void reset() { b3hip::ok(b3h_pipeline_reset(m_p)); }
	b3hRigidBodyData* getBodyBuffer()
	{
		b3hRigidBodyData* d = nullptr;
		b3hip::ok(b3h_pipeline_get_body_buffer(m_p, &d));
		return d;
	}
	int getNumBodies() const
	{
		int n = 0;
		b3hip::ok(b3h_pipeline_get_num_bodies(m_p, &n));
		return n;
	}
	/* castRays(rays, hitResults), b3GpuRigidBodyPipeline.h:63 (host arrays) */
	void castRays(const std::vector<b3hRayInfo>& rays, std::vector<b3hRayHit>& hitResults)
	{
		if (hitResults.size() < rays.size()) hitResults.resize(rays.size());
		b3hip::ok(b3h_pipeline_cast_rays(m_p, rays.data(), hitResults.data(), (int)rays.size()));
	}
	/* createPoint2PointConstraint / createFixedConstraint / removeConstraintByUid /
	 * copyConstraintsToHost, b3GpuRigidBodyPipeline.h:52-58; return the uid or -1 */
	int createPoint2PointConstraint(int bodyA, int bodyB, const float* pivotInA, const float* pivotInB,
									float breakingThreshold)
	{
		int uid = -1;
		return b3hip::ok(b3h_pipeline_create_point2point_constraint(m_p, bodyA, bodyB, pivotInA, pivotInB,
																	 breakingThreshold, &uid))
				   ? uid
				   : -1;
	}
	int createFixedConstraint(int bodyA, int bodyB, const float* pivotInA, const float* pivotInB,
							  const float* relTargetAB, float breakingThreshold)
	{
		int uid = -1;
		return b3hip::ok(b3h_pipeline_create_fixed_constraint(m_p, bodyA, bodyB, pivotInA, pivotInB, relTargetAB,
															   breakingThreshold, &uid))
				   ? uid
				   : -1;
	}
	void removeConstraintByUid(int uid) { b3hip::ok(b3h_pipeline_remove_constraint_by_uid(m_p, uid)); }
	void copyConstraintsToHost() { b3hip::ok(b3h_pipeline_copy_constraints_to_host(m_p)); }
	/* solver parameters the reference hard-codes (iterations, friction, dt, ...) */
	void setSolverParams(const b3hSolverParams& prm) { b3hip::ok(b3h_pipeline_set_solver_params(m_p, &prm)); }
	b3hSolverParams getSolverParams() const
	{
		b3hSolverParams prm = {};
		b3hip::ok(b3h_pipeline_get_solver_params(m_p, &prm));
		return prm;
	}
	/* stepSimulation repeated `count` times as one captured HIP graph launch */
	void stepSimulationGraph(float deltaTime, int count)
	{
		b3hip::ok(b3h_pipeline_step_graph(m_p, deltaTime, count));
	}
	/* warm start (no counterpart in the reference, which starts every solve
	 * from zero impulses): b3hSolverParams::m_warmStartMode 0 cold, 1 previous
	 * frame, 2 learned with the network given here */
	bool setWarmStartModel(b3hMlp* net) { return b3hip::ok(b3h_pipeline_set_warm_start_model(m_p, net)); }
	void clearWarmStartHistory() { b3hip::ok(b3h_pipeline_clear_warm_start_history(m_p)); }
	/* state image: bodies, inertias, gravity, solver parameters and the
	 * warm-start history (no counterpart in the reference) */
	bool saveState(std::vector<unsigned char>& out)
	{
		size_t n = 0;
		if (!b3hip::ok(b3h_pipeline_save_state(m_p, nullptr, 0, &n))) return false;
		out.resize(n);
		return b3hip::ok(b3h_pipeline_save_state(m_p, out.data(), out.size(), &n));
	}
	bool loadState(const std::vector<unsigned char>& in)
	{
		return b3hip::ok(b3h_pipeline_load_state(m_p, in.data(), in.size()));
	}
	/* contact logging for training the network (device rows [cap][24] and [cap][3]) */
	bool setContactLog(float* featuresDev, float* targetsDev, int capacityRows)
	{
		return b3hip::ok(b3h_pipeline_set_contact_log(m_p, featuresDev, targetsDev, capacityRows));
	}
	int getContactLogCount()
	{
		int n = 0;
		b3hip::ok(b3h_pipeline_get_contact_log_count(m_p, &n));
		return n;
	}
	/* PGS residuals per sweep (after setTrackResidual(true)) */
	void setTrackResidual(bool on) { b3hip::ok(b3h_pipeline_set_track_residual(m_p, on ? 1 : 0)); }
	bool getResidualsL2(float* out, int n) { return b3hip::ok(b3h_pipeline_get_residuals_l2(m_p, out, n)); }
};

/* The learned warm start's network (3-layer MLP, GELU, bf16 weights, fp32
 * accumulation on the matrix cores); no counterpart in the reference. */
class b3WarmStartNetwork
{
	b3hMlp* m_m = nullptr;

public:
	explicit b3WarmStartNetwork(int inDim = 24, int outDim = 3, int maxTrainRows = 8192, unsigned seed = 1,
								void* hipStream = nullptr)
	{
		b3hip::ok(b3h_mlp_create(inDim, outDim, maxTrainRows, seed, hipStream, &m_m));
	}
	~b3WarmStartNetwork() { b3h_mlp_destroy(m_m); }
	b3WarmStartNetwork(const b3WarmStartNetwork&) = delete;
	b3WarmStartNetwork& operator=(const b3WarmStartNetwork&) = delete;
	b3hMlp* handle() const { return m_m; }
	int numParams() const
	{
		int n = 0;
		b3hip::ok(b3h_mlp_num_params(m_m, &n));
		return n;
	}
	bool setParams(const float* host) { return b3hip::ok(b3h_mlp_set_params(m_m, host)); }
	bool getParams(float* host) { return b3hip::ok(b3h_mlp_get_params(m_m, host)); }
	bool forward(const float* xDev, int ldx, int rows, float* yDev, int ldy)
	{
		return b3hip::ok(b3h_mlp_forward(m_m, xDev, ldx, rows, yDev, ldy));
	}
	bool train(const float* xDev, int ldx, const float* tDev, int ldt, int rows, int batchRows, int steps,
			   float learningRate, unsigned seed, float* lossHistory)
	{
		return b3hip::ok(b3h_mlp_train(m_m, xDev, ldx, tDev, ldt, rows, batchRows, steps, learningRate, seed,
									   lossHistory));
	}
};

#endif
