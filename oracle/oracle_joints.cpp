/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * The joint solver of the step, b3GpuPgsConstraintSolver::solveJoints
 * (src/Bullet3OpenCL/RigidBody/b3GpuPgsConstraintSolver.cpp:217-1068) with the
 * kernels of RigidBody/kernels/jointSolver.cl, restated in its own order:
 * solver bodies for every body, rows per constraint (getInfo2Kernel), the
 * host batching (sortConstraintByBatch3), the batched iterations run
 * constraint after constraint, breaking, then the velocity write-back over
 * every body.
 */
#include "oracle.h"
#include <vector>
#include <algorithm>
#include <stdlib.h>

namespace
{
/* b3GpuSolverBody (jointSolver.cl:72-90): the deltas the rows accumulate */
struct SolverBody
{
	F4 dLin, dAng, invMass;
};
/* b3SolverConstraint (jointSolver.cl:127-152), the fields the solve reads */
struct Row
{
	F4 r1xn, n, r2xn, angA, angB;
	float applied, jacInv, rhs, cfm, lo, hi;
	int a, b;
};
void skew(F4 v, F4* r0, F4* r1, F4* r2)
{
	*r0 = f4(0.f, -v.z, v.y, 0.f);
	*r1 = f4(v.z, 0.f, -v.x, 0.f);
	*r2 = f4(-v.y, v.x, 0.f, 0.f);
}
} // namespace

extern "C" int b3o_solve_joints(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias,
								 b3hGpuGenericConstraint* cs, int numC, float timeStep, int iterations)
{
	if (numC <= 0) return 0;
	const float erp = 0.2f, globalCfm = 0.f, damping = 1.f; /* b3ContactSolverInfo.h:60-80 */
	/* initSolverBodies, :331-350 */
	std::vector<SolverBody> sb((size_t)numBodies);
	for (int i = 0; i < numBodies; i++)
	{
		sb[i].dLin = sb[i].dAng = f4(0.f, 0.f, 0.f, 0.f);
		float m = bodies[i].m_invMass;
		sb[i].invMass = f4(m, m, m, 0.f);
	}
	/* getInfo1Kernel, :374-398, and the row offsets (prefix scan) */
	std::vector<int> info1(numC), rowOff(numC);
	int total = 0;
	for (int i = 0; i < numC; i++)
	{
		int t = cs[i].m_constraintType;
		info1[i] = t == 3 ? 3 : (t == 4 ? 6 : 0);
		rowOff[i] = total;
		total += info1[i];
	}
	/* getInfo2Kernel, :651-877 */
	std::vector<Row> rows((size_t)total);
	const float fps = 1.f / timeStep;
	const float k = fps * erp;
	for (int i = 0; i < numC; i++)
	{
		if (!info1[i]) continue;
		const b3hGpuGenericConstraint& c = cs[i];
		const b3hRigidBodyData& A = bodies[c.m_rbA];
		const b3hRigidBodyData& B = bodies[c.m_rbB];
		Row* r = &rows[rowOff[i]];
		for (int j = 0; j < info1[i]; j++)
		{
			memset(&r[j], 0, sizeof(Row));
			r[j].lo = -1e30f;
			r[j].hi = 1e30f;
			r[j].a = c.m_rbA;
			r[j].b = c.m_rbB;
		}
		r[0].cfm = globalCfm;
		/* getInfo2Point2Point, :483-531 */
		r[0].n.x = 1.f;
		r[1].n.y = 1.f;
		r[2].n.z = 1.f;
		F4 a1 = qrot(A.m_quat, c.m_pivotInA);
		skew(neg4(a1), &r[0].r1xn, &r[1].r1xn, &r[2].r1xn);
		F4 a2 = qrot(B.m_quat, c.m_pivotInB);
		skew(a2, &r[0].r2xn, &r[1].r2xn, &r[2].r2xn);
		F4 res = sub4(sub4(add4(a2, B.m_pos), a1), A.m_pos);
		r[0].rhs = k * res.x;
		r[1].rhs = k * res.y;
		r[2].rhs = k * res.z;
		if (c.m_constraintType == 4)
		{
			/* getInfo2FixedOrientation, :576-629 */
			r[3].r1xn.x = 1.f;
			r[4].r1xn.y = 1.f;
			r[5].r1xn.z = 1.f;
			r[3].r2xn.x = -1.f;
			r[4].r2xn.y = -1.f;
			r[5].r2xn.z = -1.f;
			F4 qrel = qmul(A.m_quat, qinv(B.m_quat));
			/* calculateDiffAxisAngleQuaternion(relTargetAB, qrel), :549-574 */
			F4 orn0 = c.m_relTargetAB;
			F4 diffq = sub4(orn0, qrel), sumq = add4(orn0, qrel);
			F4 orn1 = dot4(diffq, diffq) < dot4(sumq, sumq) ? qrel : neg4(qrel);
			F4 dorn = qmul(orn1, qinv(orn0));
			float w = dorn.w;
			if (w >= 1.f) w = 1.f; /* getAngle */
			if (w < -1.f) w = -1.f; /* b3Acos */
			if (w > 1.f) w = 1.f;
			float angle = 2.f * b3h_acos(w);
			F4 axis = f4(dorn.x, dorn.y, dorn.z, 0.f);
			float len = dot3(axis, axis);
			if (len < FLT_EPSILON * FLT_EPSILON)
				axis = f4(1.f, 0.f, 0.f, 0.f);
			else
			{
				float sl = sqrtf(len);
				axis = f4(axis.x / sl, axis.y / sl, axis.z / sl, axis.w / sl);
			}
			F4 d = mul4(axis, -angle);
			r[3].rhs = k * d.x;
			r[4].rhs = k * d.y;
			r[5].rhs = k * d.z;
		}
		/* finalize the rows */
		const b3hMat3x3& IA = inertias[c.m_rbA].m_invInertiaWorld;
		const b3hMat3x3& IB = inertias[c.m_rbB].m_invInertiaWorld;
		for (int j = 0; j < info1[i]; j++)
		{
			Row& q = r[j];
			if (q.hi >= c.m_breakingImpulseThreshold) q.hi = c.m_breakingImpulseThreshold;
			if (q.lo <= -c.m_breakingImpulseThreshold) q.lo = -c.m_breakingImpulseThreshold;
			q.angA = mtMul1(IA, q.r1xn);
			q.angB = mtMul1(IB, q.r2xn);
			F4 iMJlA = mul4(q.n, A.m_invMass), iMJaA = mtMul3(q.r1xn, IA);
			F4 iMJlB = mul4(q.n, B.m_invMass), iMJaB = mtMul3(q.r2xn, IB);
			float sum = dot3(iMJlA, q.n);
			sum += dot3(iMJaA, q.r1xn);
			sum += dot3(iMJlB, q.n);
			sum += dot3(iMJaB, q.r2xn);
			q.jacInv = fabsf(sum) > FLT_EPSILON ? 1.f / sum : 0.f;
			float v1 = dot3(q.n, A.m_linVel) + dot3(q.r1xn, A.m_angVel);
			float v2 = -dot3(q.n, B.m_linVel) + dot3(q.r2xn, B.m_angVel);
			float relVel = v1 + v2;
			float velErr = 0.f - relVel * damping;
			float pen = q.rhs * q.jacInv;
			float vel = velErr * q.jacInv;
			q.rhs = pen + vel;
			q.applied = 0.f;
		}
	}
	/* initBatchConstraintsKernel (:400-416) + sortConstraintByBatch3
	 * (b3GpuPgsConstraintSolver.cpp:862-948), simdWidth numC + 1, staticIdx -1 */
	struct BC
	{
		int a, b, orig, batch;
	};
	std::vector<BC> bc(numC);
	for (int i = 0; i < numC; i++)
	{
		int ra = cs[i].m_rbA, rb = cs[i].m_rbB;
		bc[i].a = bodies[ra].m_invMass != 0.f ? ra : -ra;
		bc[i].b = bodies[rb].m_invMass != 0.f ? rb : -rb;
		bc[i].orig = i;
		bc[i].batch = -1;
	}
	std::vector<int> batchSizes;
	{
		const int simdWidth = numC + 1, staticIdx = -1;
		std::vector<int> bodyUsed(numBodies / 32 + 1, 0), curUsed(2 * simdWidth);
		int curBodyUsed = 0, numValid = 0, batchIdx = 0;
		while (numValid < numC)
		{
			int nCur = 0;
			for (int i = 0; i < curBodyUsed; i++) bodyUsed[curUsed[i] / 32] = 0;
			curBodyUsed = 0;
			for (int i = numValid; i < numC; i++)
			{
				int as = bc[i].a, bs = bc[i].b;
				int a = abs(as), b = abs(bs);
				bool aS = as < 0 || as == staticIdx, bS = bs < 0 || bs == staticIdx;
				int aU = 0, bU = 0;
				if (!aS) aU = bodyUsed[a / 32] & (1 << (a & 31));
				if (!aU)
					if (!bS) bU = bodyUsed[b / 32] & (1 << (b & 31));
				if (aU == 0 && bU == 0)
				{
					if (!aS)
					{
						bodyUsed[a / 32] |= (1 << (a & 31));
						curUsed[curBodyUsed++] = a;
					}
					if (!bS)
					{
						bodyUsed[b / 32] |= (1 << (b & 31));
						curUsed[curBodyUsed++] = b;
					}
					bc[i].batch = batchIdx;
					if (i != numValid) std::swap(bc[i], bc[numValid]);
					numValid++;
					nCur++;
					if (nCur == simdWidth)
					{
						nCur = 0;
						for (int u = 0; u < curBodyUsed; u++) bodyUsed[curUsed[u] / 32] = 0;
						curBodyUsed = 0;
					}
				}
			}
			batchSizes.push_back(nCur);
			batchIdx++;
		}
	}
	/* solveGroupCacheFriendlyIterations (:586-700) -> solveJointConstraintRows
	 * (jointSolver.cl:303-329) -> resolveSingleConstraintRowGeneric (:260-301) */
	for (int it = 0; it < iterations; it++)
	{
		int off = 0;
		for (size_t bb = 0; bb < batchSizes.size(); bb++)
		{
			for (int q = off; q < off + batchSizes[bb]; q++)
			{
				int ci = bc[q].orig;
				if (!(cs[ci].m_flags & 1)) continue;
				for (int j = 0; j < info1[ci]; j++)
				{
					Row& c = rows[rowOff[ci] + j];
					SolverBody& b1 = sb[c.a];
					SolverBody& b2 = sb[c.b];
					float di = c.rhs - c.applied * c.cfm;
					float d1 = dot3(c.n, b1.dLin) + dot3(c.r1xn, b1.dAng);
					float d2 = -dot3(c.n, b2.dLin) + dot3(c.r2xn, b2.dAng);
					di -= d1 * c.jacInv;
					di -= d2 * c.jacInv;
					float sum = c.applied + di;
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
					/* internalApplyImpulse: linear component n * invMass (float4),
					 * factors (1,1,1,0) */
					F4 l1 = f4(c.n.x * b1.invMass.x, c.n.y * b1.invMass.y, c.n.z * b1.invMass.z, c.n.w * b1.invMass.w);
					b1.dLin = add4(b1.dLin, mul4(l1, di));
					b1.dAng = add4(b1.dAng, mul4(c.angA, di));
					F4 nn = neg4(c.n);
					F4 l2 = f4(nn.x * b2.invMass.x, nn.y * b2.invMass.y, nn.z * b2.invMass.z, nn.w * b2.invMass.w);
					b2.dLin = add4(b2.dLin, mul4(l2, di));
					b2.dAng = add4(b2.dAng, mul4(c.angB, di));
				}
			}
			off += batchSizes[bb];
		}
	}
	/* breakViolatedConstraintsKernel, :352-372 */
	for (int i = 0; i < numC; i++)
		for (int j = 0; j < info1[i]; j++)
			if (fabsf(rows[rowOff[i] + j].applied) >= cs[i].m_breakingImpulseThreshold) cs[i].m_flags = 0;
	/* writeBackVelocitiesKernel, :631-649 */
	for (int i = 0; i < numBodies; i++)
		if (bodies[i].m_invMass != 0.f)
		{
			bodies[i].m_linVel = add4(bodies[i].m_linVel, sb[i].dLin);
			bodies[i].m_angVel = add4(bodies[i].m_angVel, sb[i].dAng);
		}
	return (int)batchSizes.size();
}

/* the arc cosine the fixed joint uses, for tests */
extern "C" float b3o_acos(float x) { return b3h_acos(x); }
