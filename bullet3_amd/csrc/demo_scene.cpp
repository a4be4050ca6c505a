/*
 * demo_scene.cpp -- a host program written against the C++ classes of
 * include/b3hip.hpp the way the reference's demo drives its OpenCL backend
 * (GpuRigidBodyDemo::initPhysics / stepSimulation,
 * examples/OpenCL/rigidbody/GpuRigidBodyDemo.cpp:110-190, and the instance
 * registration of GpuConvexScene.cpp:171-304).  Built as a small shared
 * library so that it can be driven with any scene: shapes as packed xyz point
 * clouds, bodies as mass / position / orientation / shape index.
 */
#include "../../include/b3hip.hpp"
#include <vector>

extern "C" int b3h_demo_run_scene(int numShapes, const int* shapeVertexCounts, const float* shapeVertices,
								  int numBodies, const float* mass, const float* pos3, const float* orn4,
								  const int* shapeIndex, float cellSize, int broadphaseKind, int iterations,
								  float friction, int steps, float dt, b3hRigidBodyData* outBodies)
{
	b3hConfig config;
	b3h_config_default(&config);
	if (numBodies + 16 > config.m_maxConvexBodies) config.m_maxConvexBodies = numBodies + 16;
	config.m_maxBroadphasePairs = 16 * config.m_maxConvexBodies;
	config.m_maxContactCapacity = config.m_maxBroadphasePairs;

	b3GpuNarrowPhase np(config);
	b3GpuGridBroadphase grid(config);
	b3GpuSapBroadphase sap(config);
	b3GpuBroadphaseInterface* bp = broadphaseKind == B3H_BROADPHASE_SAP ? (b3GpuBroadphaseInterface*)&sap
																		   : (b3GpuBroadphaseInterface*)&grid;
	if (!np.isValid() || !bp->isValid()) return b3hip::lastStatus();
	const float cs[3] = {cellSize, cellSize, cellSize};
	const int dims[3] = {128, 128, 128};
	grid.setGridParams(cs, dims, 256);

	b3GpuRigidBodyPipeline pipeline(&np, bp, config);
	if (!pipeline.isValid()) return b3hip::lastStatus();
	b3hSolverParams prm = pipeline.getSolverParams();
	prm.m_numIterations = iterations;
	prm.m_frictionCoeff = friction;
	pipeline.setSolverParams(prm);

	std::vector<int> collidable(numShapes);
	const float scaling[3] = {1.f, 1.f, 1.f};
	const float* v = shapeVertices;
	for (int s = 0; s < numShapes; s++)
	{
		collidable[s] = np.registerConvexHullShape(v, 3 * sizeof(float), shapeVertexCounts[s], scaling);
		if (collidable[s] < 0) return b3hip::lastStatus();
		v += 3 * shapeVertexCounts[s];
	}
	for (int i = 0; i < numBodies; i++)
	{
		int body = pipeline.registerPhysicsInstance(mass[i], pos3 + 3 * i, orn4 + 4 * i, collidable[shapeIndex[i]], i,
													false);
		if (body < 0) return b3hip::lastStatus();
	}
	pipeline.writeAllInstancesToGpu();
	for (int s = 0; s < steps; s++) pipeline.stepSimulation(dt);
	if (b3hip::lastStatus() != B3H_OK) return b3hip::lastStatus();
	np.readbackAllBodiesToCpu(outBodies);
	return b3hip::lastStatus();
}
