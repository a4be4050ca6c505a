/*
 * k_joints.hip -- the joint solver of the step (b3GpuPgsConstraintSolver::
 * solveJoints, src/Bullet3OpenCL/RigidBody/b3GpuPgsConstraintSolver.cpp:
 * 217-1068, kernels RigidBody/kernels/jointSolver.cl), run between the
 * narrowphase and the contact solver as b3GpuRigidBodyPipeline::stepSimulation
 * runs it (b3GpuRigidBodyPipeline.cpp:361-372).
 *
 *   k_joint_rows   one thread per constraint: its 3 (point-to-point) or 6
 *                  (fixed) solver rows -- Jacobians, limits clamped to the
 *                  breaking threshold, angular components, effective mass
 *                  inverse and right-hand side (getInfo2Kernel, :651-877)
 *   k_joint_solve  ONE work-group for all iterations and batches: the batches
 *                  (host-built greedy batches, no body twice in a batch) run in
 *                  order with a barrier between them, the constraints of a
 *                  batch in parallel, each constraint's rows in order
 *                  (solveJointConstraintRows, :303-329) -- where the reference
 *                  launches a kernel per batch and iteration.  The velocity
 *                  deltas live per joint body (the distinct bodies the
 *                  constraints name), not per body of the world.
 *   k_joint_finish one thread per constraint clears the enabled flag of a
 *                  constraint whose row impulse reached its breaking threshold
 *                  (breakViolatedConstraintsKernel, :352-372), and one thread
 *                  per body adds its deltas to a dynamic body's velocities
 *                  (writeBackVelocitiesKernel, :631-649).
 */
#include "b3h_internal.h"

#define JOINT_INF 1e30f

/* skew-symmetric rows of v (getSkewSymmetricMatrix, jointSolver.cl:476-481) */
__device__ __forceinline__ void joint_skew(F4 v, F4* r0, F4* r1, F4* r2)
{
	*r0 = f4(0.f, -v.z, v.y, 0.f);
	*r1 = f4(v.z, 0.f, -v.x, 0.f);
	*r2 = f4(-v.y, v.x, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_joint_rows(int numC, const b3hGpuGenericConstraint* __restrict__ cs,
													const int* __restrict__ rowOffset,
													const b3hRigidBodyData* __restrict__ bodies,
													const b3hInertiaData* __restrict__ inertias,
													const int* __restrict__ slotOf, b3hJointRow* __restrict__ rows,
													b3hJointParams jp)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= numC) return;
	const b3hGpuGenericConstraint c = cs[i];
	const int type = c.m_constraintType;
	const int n = type == B3H_POINT2POINT_CONSTRAINT_TYPE ? 3 : (type == B3H_FIXED_CONSTRAINT_TYPE ? 6 : 0);
	if (n == 0) return;
	const b3hRigidBodyData& A = bodies[c.m_rbA];
	const b3hRigidBodyData& B = bodies[c.m_rbB];
	F4 r1[6], nrm[6], r2[6];
	float rhs[6];
	for (int j = 0; j < 6; j++)
	{
		r1[j] = nrm[j] = r2[j] = f4(0.f, 0.f, 0.f, 0.f);
		rhs[j] = 0.f;
	}
	const float fps = 1.f / jp.timeStep;
	const float k = fps * jp.erp;
	/* point-to-point rows (getInfo2Point2Point, :483-531) */
	nrm[0].x = 1.f;
	nrm[1].y = 1.f;
	nrm[2].z = 1.f;
	const F4 a1 = qrot(A.m_quat, c.m_pivotInA);
	joint_skew(neg4(a1), &r1[0], &r1[1], &r1[2]);
	const F4 a2 = qrot(B.m_quat, c.m_pivotInB);
	joint_skew(a2, &r2[0], &r2[1], &r2[2]);
	const F4 res = sub4(sub4(add4(a2, B.m_pos), a1), A.m_pos);
	rhs[0] = k * res.x;
	rhs[1] = k * res.y;
	rhs[2] = k * res.z;
	if (n == 6)
	{
		/* three rows that make the orientations equal (getInfo2FixedOrientation,
		 * :576-629, with calculateDiffAxisAngleQuaternion :549-574) */
		r1[3].x = 1.f;
		r1[4].y = 1.f;
		r1[5].z = 1.f;
		r2[3].x = -1.f;
		r2[4].y = -1.f;
		r2[5].z = -1.f;
		const F4 qrel = qmul(A.m_quat, qinv(B.m_quat));
		const F4 t = c.m_relTargetAB;
		const F4 dq = sub4(t, qrel), sq = add4(t, qrel);
		const F4 near = dot4(dq, dq) < dot4(sq, sq) ? qrel : neg4(qrel);
		const F4 dorn = qmul(near, qinv(t));
		float w = dorn.w;
		if (w >= 1.f) w = 1.f;
		if (w < -1.f) w = -1.f;
		const float angle = 2.f * b3h_acos(w);
		F4 axis = f4(dorn.x, dorn.y, dorn.z, 0.f);
		const float len = dot3(axis, axis);
		if (len < FLT_EPSILON * FLT_EPSILON)
			axis = f4(1.f, 0.f, 0.f, 0.f);
		else
		{
			const float sl = sqrtf(len);
			axis = f4(axis.x / sl, axis.y / sl, axis.z / sl, axis.w / sl);
		}
		const F4 diff = mul4(axis, -angle);
		rhs[3] = k * diff.x;
		rhs[4] = k * diff.y;
		rhs[5] = k * diff.z;
	}
	const b3hMat3x3 IA = inertias[c.m_rbA].m_invInertiaWorld, IB = inertias[c.m_rbB].m_invInertiaWorld;
	const int sa = slotOf[2 * i], sb = slotOf[2 * i + 1];
	b3hJointRow* out = rows + rowOffset[i];
	for (int j = 0; j < n; j++)
	{
		b3hJointRow r;
		r.r1xn = r1[j];
		r.normal = nrm[j];
		r.r2xn = r2[j];
		r.lo = -JOINT_INF;
		r.hi = JOINT_INF;
		if (r.hi >= c.m_breakingImpulseThreshold) r.hi = c.m_breakingImpulseThreshold;
		if (r.lo <= -c.m_breakingImpulseThreshold) r.lo = -c.m_breakingImpulseThreshold;
		r.cfm = j == 0 ? jp.globalCfm : 0.f;
		r.angA = mtMul1(IA, r1[j]);
		r.angB = mtMul1(IB, r2[j]);
		const F4 iMJlA = mul4(nrm[j], A.m_invMass);
		const F4 iMJaA = mtMul3(r1[j], IA);
		const F4 iMJlB = mul4(nrm[j], B.m_invMass);
		const F4 iMJaB = mtMul3(r2[j], IB);
		float sum = dot3(iMJlA, nrm[j]);
		sum += dot3(iMJaA, r1[j]);
		sum += dot3(iMJlB, nrm[j]);
		sum += dot3(iMJaB, r2[j]);
		r.jacInv = fabsf(sum) > FLT_EPSILON ? 1.f / sum : 0.f;
		const float v1 = dot3(nrm[j], A.m_linVel) + dot3(r1[j], A.m_angVel);
		const float v2 = -dot3(nrm[j], B.m_linVel) + dot3(r2[j], B.m_angVel);
		const float relVel = v1 + v2;
		const float velErr = 0.f - relVel * jp.damping;
		r.rhs = rhs[j] * r.jacInv + velErr * r.jacInv;
		r.applied = 0.f;
		r.invMassA = A.m_invMass;
		r.invMassB = B.m_invMass;
		r.slotA = sa;
		r.slotB = sb;
		out[j] = r;
	}
}

/* resolveSingleConstraintRowGeneric, jointSolver.cl:260-301 */
__device__ __forceinline__ void joint_resolve(b3hJointRow& c, F4* __restrict__ delta)
{
	F4* dA = delta + 2 * c.slotA;
	F4* dB = delta + 2 * c.slotB;
	float di = c.rhs - c.applied * c.cfm;
	const float d1 = dot3(c.normal, dA[0]) + dot3(c.r1xn, dA[1]);
	const float d2 = -dot3(c.normal, dB[0]) + dot3(c.r2xn, dB[1]);
	di -= d1 * c.jacInv;
	di -= d2 * c.jacInv;
	const float sum = c.applied + di;
	if (sum < c.lo)
	{
		di = c.lo - c.applied;
		c.applied = c.lo;
	}
	else if (sum > c.hi)
	{
		di = c.hi - c.applied;
		c.applied = c.hi;
	}
	else
		c.applied = sum;
	const F4 linA = mul4(c.normal, c.invMassA);
	dA[0] = add4(dA[0], mul4(linA, di));
	dA[1] = add4(dA[1], mul4(c.angA, di));
	const F4 linB = mul4(neg4(c.normal), c.invMassB);
	dB[0] = add4(dB[0], mul4(linB, di));
	dB[1] = add4(dB[1], mul4(c.angB, di));
}

#define JOINT_SOLVE_THREADS 1024
__global__ __launch_bounds__(JOINT_SOLVE_THREADS) void k_joint_solve(
	int numBatches, const int* __restrict__ batchStart, const int* __restrict__ order,
	const b3hGpuGenericConstraint* __restrict__ cs, const int* __restrict__ rowOffset, const int* __restrict__ rowCount,
	b3hJointRow* rows, F4* delta, int numSlots, int iterations)
{
	const int t = threadIdx.x;
	for (int s = t; s < 2 * numSlots; s += JOINT_SOLVE_THREADS) delta[s] = f4(0.f, 0.f, 0.f, 0.f);
	__syncthreads();
	for (int it = 0; it < iterations; it++)
		for (int bb = 0; bb < numBatches; bb++)
		{
			const int b0 = batchStart[bb], b1 = batchStart[bb + 1];
			for (int k = b0 + t; k < b1; k += JOINT_SOLVE_THREADS)
			{
				const int ci = order[k];
				if (!(cs[ci].m_flags & B3H_CONSTRAINT_FLAG_ENABLED)) continue;
				const int ro = rowOffset[ci], nr = rowCount[ci];
				for (int j = 0; j < nr; j++)
				{
					b3hJointRow r = rows[ro + j];
					joint_resolve(r, delta);
					rows[ro + j].applied = r.applied;
				}
			}
			__syncthreads();
		}
}

__global__ __launch_bounds__(256) void k_joint_break(int numC, b3hGpuGenericConstraint* __restrict__ cs,
													 const int* __restrict__ rowOffset,
													 const int* __restrict__ rowCount,
													 const b3hJointRow* __restrict__ rows)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= numC) return;
	const float thr = cs[i].m_breakingImpulseThreshold;
	for (int j = 0; j < rowCount[i]; j++)
		if (fabsf(rows[rowOffset[i] + j].applied) >= thr) cs[i].m_flags = 0;
}

/* every dynamic body takes its delta, zero for a body no joint names (as the
 * reference's pass over all bodies, which turns a -0 velocity lane into +0) */
__global__ __launch_bounds__(256) void k_joint_writeback(int numBodies, const int* __restrict__ bodySlot,
														 const F4* __restrict__ delta,
														 b3hRigidBodyData* __restrict__ bodies)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= numBodies) return;
	b3hRigidBodyData& b = bodies[i];
	if (b.m_invMass != 0.f)
	{
		const int s = bodySlot[i];
		const F4 z = f4(0.f, 0.f, 0.f, 0.f);
		b.m_linVel = add4(b.m_linVel, s >= 0 ? delta[2 * s] : z);
		b.m_angVel = add4(b.m_angVel, s >= 0 ? delta[2 * s + 1] : z);
	}
}

void b3h_solve_joints(const b3hJointBuffers& jb, b3hRigidBodyData* bodies, const b3hInertiaData* inertias,
					  const b3hJointParams& jp, hipStream_t s)
{
	if (jb.numConstraints <= 0) return;
	const int nb = (jb.numConstraints + 255) / 256;
	hipLaunchKernelGGL(k_joint_rows, dim3(nb), dim3(256), 0, s, jb.numConstraints,
					   (const b3hGpuGenericConstraint*)jb.constraints, (const int*)jb.rowOffset,
					   (const b3hRigidBodyData*)bodies, inertias, (const int*)jb.slotOf, jb.rows, jp);
	hipLaunchKernelGGL(k_joint_solve, dim3(1), dim3(JOINT_SOLVE_THREADS), 0, s, jb.numBatches,
					   (const int*)jb.batchStart, (const int*)jb.order,
					   (const b3hGpuGenericConstraint*)jb.constraints, (const int*)jb.rowOffset,
					   (const int*)jb.rowCount, jb.rows, jb.delta, jb.numSlots, jp.iterations);
	hipLaunchKernelGGL(k_joint_break, dim3(nb), dim3(256), 0, s, jb.numConstraints, jb.constraints,
					   (const int*)jb.rowOffset, (const int*)jb.rowCount, (const b3hJointRow*)jb.rows);
	hipLaunchKernelGGL(k_joint_writeback, dim3((jb.numBodies + 255) / 256), dim3(256), 0, s, jb.numBodies,
					   (const int*)jb.bodySlot, (const F4*)jb.delta, bodies);
}
