/*
 * ORACLE -- test infrastructure only.
 * Exposes the reference's host emulation of its GPU contact solver --
 * SolveTask with solveContact<false> / solveFriction,
 * src/Bullet3OpenCL/RigidBody/b3Solver.cpp:187-466 -- compiled from
 * /root/reference by oracle/ref.mk (with the reference's vendored clew, so no
 * OpenCL library is needed and none is called).  b3Solver.cpp's sweep helpers
 * have internal linkage, so this file includes that source file.
 *
 * ref_pgs_host runs the sweep schedule of b3Solver::solveContactConstraintHost
 * (b3Solver.cpp:531-610): all normal iterations, then all friction
 * iterations; each iteration visits the 8 cell-parity batches in order and,
 * within a batch, the 32 work-group cells in wgIdx order; a cell's
 * constraints run batch by batch (SolveTask::run).
 */
#include "Bullet3OpenCL/RigidBody/b3Solver.cpp"
#include <string.h>

extern "C" int ref_pgs_host(int numBodies, void* bodies, const void* inertias, int n, void* constraints,
							const unsigned* cellCounts, const unsigned* cellOffsets, int iterations)
{
	b3AlignedObjectArray<b3RigidBodyData> bodyNative;
	bodyNative.resize(numBodies);
	memcpy(&bodyNative[0], bodies, sizeof(b3RigidBodyData) * numBodies);
	b3AlignedObjectArray<b3InertiaData> shapeNative;
	shapeNative.resize(numBodies);
	memcpy(&shapeNative[0], inertias, sizeof(b3InertiaData) * numBodies);
	b3AlignedObjectArray<b3GpuConstraint4> cs;
	cs.resize(n);
	if (n) memcpy(&cs[0], constraints, sizeof(b3GpuConstraint4) * n);
	/* SolveTask reads the size of every (cell, batch) */
	b3AlignedObjectArray<int> batchSizes;
	batchSizes.resize(B3_SOLVER_N_CELLS * B3_MAX_NUM_BATCHES);
	for (int i = 0; i < batchSizes.size(); i++) batchSizes[i] = 0;
	for (int cell = 0; cell < B3_SOLVER_N_CELLS; cell++)
		for (unsigned k = 0; k < cellCounts[cell]; k++)
		{
			int b = cs[cellOffsets[cell] + k].m_batchIdx;
			if (b < 0 || b >= B3_MAX_NUM_BATCHES) return -1;
			batchSizes[cell * B3_MAX_NUM_BATCHES + b]++;
		}
	const int nSplitX = B3_SOLVER_N_SPLIT_X, nSplitY = B3_SOLVER_N_SPLIT_Y;
	const int numWorkgroups = B3_SOLVER_N_CELLS / B3_SOLVER_N_BATCHES;
	for (int friction = 0; friction < 2; friction++)
		for (int iter = 0; iter < iterations; iter++)
			for (int cellBatch = 0; cellBatch < B3_SOLVER_N_BATCHES; cellBatch++)
				for (int wgIdx = 0; wgIdx < numWorkgroups; wgIdx++)
				{
					int zIdx = (wgIdx / ((nSplitX * nSplitY) / 4)) * 2 + ((cellBatch & 4) >> 2);
					int remain = (wgIdx % ((nSplitX * nSplitY) / 4));
					int yIdx = (remain / (nSplitX / 2)) * 2 + ((cellBatch & 2) >> 1);
					int xIdx = (remain % (nSplitX / 2)) * 2 + (cellBatch & 1);
					int cellIdx = xIdx + yIdx * nSplitX + zIdx * (nSplitX * nSplitY);
					if (cellCounts[cellIdx] == 0) continue;
					SolveTask task(bodyNative, shapeNative, cs, cellOffsets[cellIdx], cellCounts[cellIdx], B3_MAX_NUM_BATCHES,
								   0, wgIdx, &batchSizes, cellIdx);
					task.m_solveFriction = friction != 0;
					task.run(0);
				}
	memcpy(bodies, &bodyNative[0], sizeof(b3RigidBodyData) * numBodies);
	if (n) memcpy(constraints, &cs[0], sizeof(b3GpuConstraint4) * n);
	return 0;
}
