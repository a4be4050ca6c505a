/*
 * k_solver.hip -- batched projected Gauss-Seidel contact solver
 * (b3GpuPgsContactSolver::solveContacts, src/Bullet3OpenCL/RigidBody/b3GpuPgsContactSolver.cpp:568-1103).
 *
 * Contact order.  The reference sorts contacts by (bodyA, bodyB, childA, childB)
 * as uint32 (four radix sorts, :575-637), then stably by the 8x4x8 spatial cell
 * of the dynamic body (SetSortDataKernel, solverSetup2.cl:540-599), and batches
 * every cell greedily (CreateBatchesNew, batchingKernelsNew.cl:144-231).  Since
 * (bodyA, bodyB) is unique per contact (one manifold per pair; the child indices
 * are -1 for convex hulls), that order equals: cell, then (uint bodyA, uint
 * bodyB).  Here: one atomic histogram + scatter puts every contact into its
 * cell, and one work-group per cell sorts its (bodyA,bodyB) keys in LDS (bitonic)
 * and runs the greedy batching there -- no global radix sorts.
 *
 * Solve order.  BatchSolveKernelContact / BatchSolveKernelFriction
 * (solveContact.cl:397-482, solveFriction.cl:426-503) walk a cell's batches in
 * order, 64 lanes per row, each row aligned to its batch start and free of
 * shared dynamic bodies; 8 launches per sweep cover the 2x2x2 cell parities.
 * The same walk is kept here (one wave per cell, rows within a batch), so the
 * result equals the reference's sequential order bit for bit.
 *
 * Bytes per sweep (HBM): constraint 176 B read + 16 B impulse write, and per
 * touched body 80 B + 48 B inverse inertia read, 32 B velocity write.
 */
#include "b3h_internal.h"

#define CHECK_SIZE 64
#define SIMD_WIDTH 64

__device__ __forceinline__ int cell_index(F4 p, float scale)
{
	int xIdx = (int)((p.x - ((p.x < 0.f) ? 1.f : 0.f)) * scale) & (B3H_N_SPLIT_X - 1);
	int yIdx = (int)((p.y - ((p.y < 0.f) ? 1.f : 0.f)) * scale) & (B3H_N_SPLIT_Y - 1);
	int zIdx = (int)((p.z - ((p.z < 0.f) ? 1.f : 0.f)) * scale) & (B3H_N_SPLIT_Z - 1);
	return xIdx + yIdx * B3H_N_SPLIT_X + zIdx * B3H_N_SPLIT_X * B3H_N_SPLIT_Y;
}

__device__ __forceinline__ int iabs(int x) { return x < 0 ? -x : x; }

__global__ __launch_bounds__(256) void k_solver_keys(const int* __restrict__ nPtr, const b3hContact4Data* __restrict__ staging,
													 const int* __restrict__ contactPair,
													 const b3hRigidBodyData* __restrict__ bodies, int staticIdx,
													 float scale, int* __restrict__ cellOf, int* __restrict__ cellCounts)
{
	int n = *nPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const b3hContact4Data& c = staging[contactPair[i]];
		int aS = c.m_bodyAPtrAndSignBit, bS = c.m_bodyBPtrAndSignBit;
		bool aStatic = (aS < 0) || (aS == staticIdx);
		int idx = aStatic ? iabs(bS) : iabs(aS);
		int cell = cell_index(bodies[idx].m_pos, scale);
		cellOf[i] = cell;
		atomicAdd(&cellCounts[cell], 1);
	}
}

__global__ __launch_bounds__(B3H_N_CELLS) void k_cell_scan(const int* __restrict__ counts, int* __restrict__ offsets,
														   int* __restrict__ cursor)
{
	__shared__ int s[B3H_N_CELLS];
	int t = threadIdx.x;
	int v = counts[t];
	s[t] = v;
	__syncthreads();
	for (int off = 1; off < B3H_N_CELLS; off <<= 1)
	{
		int x = t >= off ? s[t - off] : 0;
		__syncthreads();
		s[t] += x;
		__syncthreads();
	}
	int ex = s[t] - v;
	offsets[t] = ex;
	cursor[t] = ex;
	if (t == B3H_N_CELLS - 1) offsets[B3H_N_CELLS] = s[t];
}

__global__ __launch_bounds__(256) void k_cell_scatter(const int* __restrict__ nPtr, const int* __restrict__ cellOf,
													  int* __restrict__ cursor, int* __restrict__ cellSlots)
{
	int n = *nPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		int pos = atomicAdd(&cursor[cellOf[i]], 1);
		cellSlots[pos] = i;
	}
}

/* One work-group per cell: sort the cell's (uint bodyA, uint bodyB) keys, then
 * CreateBatchesNew's greedy batching.  Arrays live in LDS when the cell fits
 * (ldsCap entries), otherwise in global scratch (sortKeys, 2 entries per
 * contact so that the power-of-two padding fits). */
__global__ __launch_bounds__(256) void k_batch_cells(const int* __restrict__ cellCounts, const int* __restrict__ cellOffsets,
													 const int* __restrict__ cellSlots,
													 const b3hContact4Data* __restrict__ staging,
													 const int* __restrict__ contactPair, int staticIdx, int ldsCap,
													 unsigned long long* __restrict__ gKeys, int* __restrict__ perm,
													 int* __restrict__ batchIdxOut, int* __restrict__ batchSizes)
{
	extern __shared__ unsigned long long smem[];
	__shared__ unsigned bitmap[CHECK_SIZE];
	int cell = blockIdx.x;
	int n = cellCounts[cell];
	int off = cellOffsets[cell];
	int t = threadIdx.x;
	if (n == 0)
	{
		if (t == 0) batchSizes[cell] = 0;
		return;
	}
	int np2 = 1;
	while (np2 < n) np2 <<= 1;
	bool inLds = np2 <= ldsCap;
	/* key = uint bodyA << 32 | uint bodyB ; payload = contact slot (parallel array) */
	unsigned long long* key = inLds ? smem : gKeys + 2 * (size_t)off;
	unsigned long long* pay = inLds ? smem + ldsCap : gKeys + 2 * (size_t)(off + 0) + np2; /* second half */
	if (!inLds)
	{
		/* global layout: keys at [2*off, 2*off+np2), payload right after; the
		 * 2*count region of this cell holds 2*np2 > 2*n entries only if np2 <= n,
		 * so use the padded region of the whole scratch instead */
		key = gKeys + 4 * (size_t)off;
		pay = key + np2;
	}
	for (int i = t; i < np2; i += blockDim.x)
	{
		if (i < n)
		{
			int slot = cellSlots[off + i];
			const b3hContact4Data& c = staging[contactPair[slot]];
			key[i] = ((unsigned long long)(unsigned)c.m_bodyAPtrAndSignBit << 32) | (unsigned)c.m_bodyBPtrAndSignBit;
			pay[i] = (unsigned long long)slot;
		}
		else
		{
			key[i] = ~0ull;
			pay[i] = ~0ull;
		}
	}
	__syncthreads();
	/* bitonic sort ascending */
	for (int k = 2; k <= np2; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1)
		{
			for (int i = t; i < np2; i += blockDim.x)
			{
				int ixj = i ^ j;
				if (ixj > i)
				{
					unsigned long long ki = key[i], kj = key[ixj];
					bool up = (i & k) == 0;
					if ((ki > kj) == up)
					{
						key[i] = kj;
						key[ixj] = ki;
						unsigned long long p = pay[i];
						pay[i] = pay[ixj];
						pay[ixj] = p;
					}
				}
			}
			__syncthreads();
		}
	/* greedy batching (serial, as the reference): batch ids go into the high
	 * half of pay[] */
	if (t == 0)
	{
		int numValid = 0, batch = 0;
		while (numValid < n)
		{
			int nCur = 0;
			for (int w = 0; w < CHECK_SIZE; w++) bitmap[w] = 0;
			for (int i = numValid; i < n; i++)
			{
				unsigned long long kk = key[i];
				int aS = (int)(unsigned)(kk >> 32), bS = (int)(unsigned)(kk & 0xffffffffu);
				int a = iabs(aS), b = iabs(bS);
				bool aStatic = (aS < 0) || aS == staticIdx;
				bool bStatic = (bS < 0) || bS == staticIdx;
				int ia = a % (32 * CHECK_SIZE), ib = b % (32 * CHECK_SIZE);
				unsigned aU = aStatic ? 0u : (bitmap[ia >> 5] & (1u << (ia & 31)));
				unsigned bU = bStatic ? 0u : (bitmap[ib >> 5] & (1u << (ib & 31)));
				if (aU == 0 && bU == 0)
				{
					if (!aStatic) bitmap[ia >> 5] |= (1u << (ia & 31));
					if (!bStatic) bitmap[ib >> 5] |= (1u << (ib & 31));
					unsigned long long pi = (pay[i] & 0xffffffffull) | ((unsigned long long)(unsigned)batch << 32);
					if (i != numValid)
					{
						key[i] = key[numValid];
						pay[i] = pay[numValid];
						key[numValid] = kk;
					}
					pay[numValid] = pi;
					numValid++;
					nCur++;
					if (nCur == SIMD_WIDTH)
					{
						nCur = 0;
						for (int w = 0; w < CHECK_SIZE; w++) bitmap[w] = 0;
					}
				}
			}
			batch++;
		}
		batchSizes[cell] = batch;
	}
	__syncthreads();
	for (int i = t; i < n; i += blockDim.x)
	{
		unsigned long long p = pay[i];
		perm[off + i] = (int)(unsigned)(p & 0xffffffffull);
		batchIdxOut[off + i] = (int)(unsigned)(p >> 32);
	}
}

/* ---------------- contact -> constraint (setConstraint4, b3ConvertConstraint4.h:62-148) ---------------- */
__device__ void plane_space(F4 n, F4* p, F4* q)
{
	if (fabsf(n.z) > 0.70710678f)
	{
		float a = n.y * n.y + n.z * n.z;
		float k = 1.f / sqrtf(a);
		p->x = 0;
		p->y = -n.z * k;
		p->z = n.y * k;
		q->x = a * k;
		q->y = -n.x * p->z;
		q->z = n.x * p->y;
	}
	else
	{
		float a = n.x * n.x + n.y * n.y;
		float k = 1.f / sqrtf(a);
		p->x = -n.y * k;
		p->y = n.x * k;
		p->z = 0;
		q->x = -n.z * p->y;
		q->y = n.z * p->x;
		q->z = a * k;
	}
	p->w = 0.f;
	q->w = 0.f;
}
__device__ __forceinline__ float calc_jac_coeff(F4 ang0, F4 ang1, float im0, const b3hMat3x3& ii0, float im1,
												 const b3hMat3x3& ii1)
{
	float j0 = im0;
	float j1 = dot3(mtMul3(ang0, ii0), ang0);
	float j2 = im1;
	float j3 = dot3(mtMul3(ang1, ii1), ang1);
	return -1.f / (j0 + j1 + j2 + j3);
}
__device__ __forceinline__ float calc_rel_vel(F4 l0, F4 l1, F4 a0, F4 a1, F4 v0, F4 w0, F4 v1, F4 w1)
{
	return dot3(l0, v0) + dot3(a0, w0) + dot3(l1, v1) + dot3(a1, w1);
}

__global__ __launch_bounds__(256) void k_to_constraints(const int* __restrict__ nPtr, const int* __restrict__ perm,
														const int* __restrict__ batchIdx,
														const b3hContact4Data* __restrict__ staging,
														const int* __restrict__ contactPair,
														const b3hRigidBodyData* __restrict__ bodies,
														const b3hInertiaData* __restrict__ inertias, b3hSolverParams prm,
														b3hContactConstraint4* __restrict__ out,
														b3hContact4Data* __restrict__ sortedOut)
{
	int n = *nPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		b3hContact4Data src = staging[contactPair[perm[i]]];
		src.m_batchIdx = batchIdx[i];
		if (sortedOut) sortedOut[i] = src;
		int a = iabs(src.m_bodyAPtrAndSignBit), b = iabs(src.m_bodyBPtrAndSignBit);
		const b3hRigidBodyData& A = bodies[a];
		const b3hRigidBodyData& B = bodies[b];
		b3hMat3x3 iiA = inertias[a].m_initInvInertia;
		b3hMat3x3 iiB = inertias[b].m_initInvInertia;
		F4 posA = A.m_pos, posB = B.m_pos;
		b3hContactConstraint4 d;
		d.m_bodyA = (unsigned)a;
		d.m_bodyB = (unsigned)b;
		float dtInv = 1.f / prm.m_solverDt;
		d.m_linear = src.m_worldNormalOnB;
		d.m_linear.w = prm.m_frictionCoeff;
		float np = src.m_worldNormalOnB.w;
		F4 nrm = src.m_worldNormalOnB;
		for (int ic = 0; ic < 4; ic++)
		{
			d.m_appliedRambdaDt[ic] = 0.f;
			if (ic >= np)
			{
				d.m_jacCoeffInv[ic] = 0.f;
				d.m_b[ic] = 0.f;
				continue;
			}
			F4 r0 = sub4(src.m_worldPosB[ic], posA);
			F4 r1 = sub4(src.m_worldPosB[ic], posB);
			F4 linear = f4(nrm.x, nrm.y, nrm.z, 0.f);
			F4 ang0 = cross3(r0, nrm);
			F4 ang1 = neg4(cross3(r1, nrm));
			d.m_jacCoeffInv[ic] = calc_jac_coeff(ang0, ang1, A.m_invMass, iiA, B.m_invMass, iiB);
			float relVelN = calc_rel_vel(linear, neg4(linear), ang0, ang1, A.m_linVel, A.m_angVel, B.m_linVel, B.m_angVel);
			float e = 0.f;
			if (relVelN * relVelN < 0.004f) e = 0.f;
			float bb = e * relVelN;
			bb += (src.m_worldPosB[ic].w + prm.m_positionDrift) * prm.m_positionConstraintCoeff * dtInv;
			d.m_b[ic] = bb;
		}
		d.m_center = f4(0, 0, 0, 0);
		d.m_fJacCoeffInv[0] = d.m_fJacCoeffInv[1] = 0.f;
		d.m_fAppliedRambdaDt[0] = d.m_fAppliedRambdaDt[1] = 0.f;
		if (np > 0)
		{
			F4 center = f4(0.f, 0.f, 0.f, 0.f);
			for (int k = 0; k < np; k++) center = add4(center, src.m_worldPosB[k]);
			float fnp = np;
			center = f4(center.x / fnp, center.y / fnp, center.z / fnp, center.w / fnp);
			F4 tg[2];
			plane_space(nrm, &tg[0], &tg[1]);
			F4 r0 = sub4(center, posA), r1 = sub4(center, posB);
			for (int k = 0; k < 2; k++)
			{
				F4 ang0 = cross3(r0, tg[k]);
				F4 ang1 = neg4(cross3(r1, tg[k]));
				d.m_fJacCoeffInv[k] = calc_jac_coeff(ang0, ang1, A.m_invMass, iiA, B.m_invMass, iiB);
			}
			d.m_center = center;
		}
		for (int k = 0; k < 4; k++) d.m_worldPos[k] = k < np ? src.m_worldPosB[k] : f4(0.f, 0.f, 0.f, 0.f);
		d.m_batchIdx = src.m_batchIdx;
		d.m_paddings = 0;
		out[i] = d;
	}
}

/* ---------------- PGS rows ---------------- */
struct Vel
{
	F4 lv, av;
};
__device__ __forceinline__ void apply_impulse(F4 linear, F4 ang0, F4 ang1, float rl, float imA, const b3hMat3x3& iiA,
											  float imB, const b3hMat3x3& iiB, Vel& a, Vel& b)
{
	F4 linImp0 = mul4(smul4(imA, linear), rl);
	F4 linImp1 = mul4(smul4(imB, neg4(linear)), rl);
	F4 angImp0 = mul4(mtMul1(iiA, ang0), rl);
	F4 angImp1 = mul4(mtMul1(iiB, ang1), rl);
	a.lv = add4(a.lv, linImp0);
	a.av = add4(a.av, angImp0);
	b.lv = add4(b.lv, linImp1);
	b.av = add4(b.av, angImp1);
}
__device__ __forceinline__ void atomic_max_pos(float* addr, float v)
{
	/* v >= 0: the IEEE bits of non-negative floats order like unsigned ints */
	atomicMax((unsigned*)addr, f2u(v));
}
__device__ __forceinline__ void write_back(b3hRigidBodyData* body, float invMass, const Vel& v)
{
	if (invMass)
	{
		body->m_linVel = v.lv;
		body->m_angVel = v.av;
	}
	else
	{
		body->m_linVel = f4(0, 0, 0, 0);
		body->m_angVel = f4(0, 0, 0, 0);
	}
}

/* phase 0: normal rows (solveContact.cl:259-377); 1: friction rows
 * (solveFriction.cl:288-408); 2: warm start (apply initial impulses once) */
__device__ void solve_one(b3hContactConstraint4* __restrict__ csp, b3hRigidBodyData* __restrict__ bodies,
						  const b3hInertiaData* __restrict__ inertias, int phase, float friction, const float* warm,
						  float* resid)
{
	b3hContactConstraint4 cs = *csp;
	int aIdx = (int)cs.m_bodyA, bIdx = (int)cs.m_bodyB;
	F4 posA = bodies[aIdx].m_pos, posB = bodies[bIdx].m_pos;
	float imA = bodies[aIdx].m_invMass, imB = bodies[bIdx].m_invMass;
	Vel va = {bodies[aIdx].m_linVel, bodies[aIdx].m_angVel};
	Vel vb = {bodies[bIdx].m_linVel, bodies[bIdx].m_angVel};
	const b3hMat3x3 iiA = inertias[aIdx].m_invInertiaWorld;
	const b3hMat3x3 iiB = inertias[bIdx].m_invInertiaWorld;
	float maxD = 0.f, maxL = 0.f;
	if (phase == 0 || phase == 2)
	{
		for (int ic = 0; ic < 4; ic++)
		{
			if (cs.m_jacCoeffInv[ic] == 0.f) continue;
			F4 r0 = sub4(cs.m_worldPos[ic], posA);
			F4 r1 = sub4(cs.m_worldPos[ic], posB);
			F4 n = neg4(cs.m_linear);
			F4 linear = f4(-n.x, -n.y, -n.z, 0.f);
			F4 ang0 = neg4(cross3(r0, n));
			F4 ang1 = cross3(r1, n);
			float rl;
			if (phase == 0)
			{
				rl = calc_rel_vel(cs.m_linear, neg4(cs.m_linear), ang0, ang1, va.lv, va.av, vb.lv, vb.av) + cs.m_b[ic];
				rl *= cs.m_jacCoeffInv[ic];
				float prev = cs.m_appliedRambdaDt[ic];
				float upd = prev;
				upd += rl;
				upd = cl_max(upd, 0.f);
				upd = cl_min(upd, FLT_MAX);
				rl = upd - prev;
				cs.m_appliedRambdaDt[ic] = upd;
				maxD = fmaxf(maxD, fabsf(rl));
				maxL = fmaxf(maxL, fabsf(upd));
			}
			else
			{
				rl = cl_max(warm[ic], 0.f);
				cs.m_appliedRambdaDt[ic] = rl;
			}
			apply_impulse(linear, ang0, ang1, rl, imA, iiA, imB, iiB, va, vb);
		}
		if (phase == 2 && !(cs.m_fJacCoeffInv[0] == 0))
		{
			float sum = 0;
			for (int j = 0; j < 4; j++) sum += cs.m_appliedRambdaDt[j];
			float maxR = friction * sum;
			F4 n = neg4(cs.m_linear);
			F4 t[2];
			plane_space(n, &t[0], &t[1]);
			F4 r0 = sub4(cs.m_center, posA), r1 = sub4(cs.m_center, posB);
			for (int i = 0; i < 2; i++)
			{
				F4 linear = f4(-t[i].x, -t[i].y, -t[i].z, 0.f);
				F4 ang0 = neg4(cross3(r0, t[i]));
				F4 ang1 = cross3(r1, t[i]);
				float l = cl_min(cl_max(warm[4 + i], -maxR), maxR);
				cs.m_fAppliedRambdaDt[i] = l;
				apply_impulse(linear, ang0, ang1, l, imA, iiA, imB, iiB, va, vb);
			}
		}
	}
	else
	{
		float sum = 0;
		for (int j = 0; j < 4; j++) sum += cs.m_appliedRambdaDt[j];
		float maxR = friction * sum;
		float minR = -maxR;
		if (cs.m_fJacCoeffInv[0] == 0 && cs.m_fJacCoeffInv[0] == 0) return;
		F4 center = cs.m_center;
		F4 n = neg4(cs.m_linear);
		F4 t[2];
		plane_space(n, &t[0], &t[1]);
		F4 r0 = sub4(center, posA), r1 = sub4(center, posB);
		for (int i = 0; i < 2; i++)
		{
			F4 linear = f4(-t[i].x, -t[i].y, -t[i].z, 0.f);
			F4 ang0 = neg4(cross3(r0, t[i]));
			F4 ang1 = cross3(r1, t[i]);
			float rl = calc_rel_vel(linear, neg4(linear), ang0, ang1, va.lv, va.av, vb.lv, vb.av);
			rl *= cs.m_fJacCoeffInv[i];
			float prev = cs.m_fAppliedRambdaDt[i];
			float upd = prev;
			upd += rl;
			upd = cl_max(upd, minR);
			upd = cl_min(upd, maxR);
			rl = upd - prev;
			cs.m_fAppliedRambdaDt[i] = upd;
			maxD = fmaxf(maxD, fabsf(rl));
			maxL = fmaxf(maxL, fabsf(upd));
			apply_impulse(linear, ang0, ang1, rl, imA, iiA, imB, iiB, va, vb);
		}
		F4 ab = nrm3(sub4(posB, posA));
		F4 ac = nrm3(sub4(center, posA));
		if (dot3(ab, ac) > 0.95f || (imA == 0.f || imB == 0.f))
		{
			float angNA = dot3(n, va.av);
			float angNB = dot3(n, vb.av);
			va.av = sub4(va.av, smul4(angNA * 0.1f, n));
			vb.av = sub4(vb.av, smul4(angNB * 0.1f, n));
		}
	}
	if (phase == 0)
		for (int ic = 0; ic < 4; ic++) csp->m_appliedRambdaDt[ic] = cs.m_appliedRambdaDt[ic];
	else if (phase == 1)
		for (int i = 0; i < 2; i++) csp->m_fAppliedRambdaDt[i] = cs.m_fAppliedRambdaDt[i];
	else
	{
		for (int ic = 0; ic < 4; ic++) csp->m_appliedRambdaDt[ic] = cs.m_appliedRambdaDt[ic];
		for (int i = 0; i < 2; i++) csp->m_fAppliedRambdaDt[i] = cs.m_fAppliedRambdaDt[i];
	}
	write_back(bodies + aIdx, imA, va);
	write_back(bodies + bIdx, imB, vb);
	if (resid)
	{
		if (maxD > 0.f) atomic_max_pos(resid, maxD);
		if (maxL > 0.f) atomic_max_pos(resid + 1, maxL);
	}
}

__device__ __forceinline__ int cell_of(int wgIdx, int ib)
{
	int zIdx = (wgIdx / ((B3H_N_SPLIT_X * B3H_N_SPLIT_Y) / 4)) * 2 + ((ib & 4) >> 2);
	int remain = (wgIdx % ((B3H_N_SPLIT_X * B3H_N_SPLIT_Y) / 4));
	int yIdx = (remain / (B3H_N_SPLIT_X / 2)) * 2 + ((ib & 2) >> 1);
	int xIdx = (remain % (B3H_N_SPLIT_X / 2)) * 2 + (ib & 1);
	return xIdx + yIdx * B3H_N_SPLIT_X + zIdx * (B3H_N_SPLIT_X * B3H_N_SPLIT_Y);
}

/* one wave per cell; BatchSolveKernelContact's batch walk (solveContact.cl:457-479) */
__global__ __launch_bounds__(64) void k_pgs(b3hRigidBodyData* __restrict__ bodies, const b3hInertiaData* __restrict__ inertias,
											b3hContactConstraint4* __restrict__ cs, const int* __restrict__ cellCounts,
											const int* __restrict__ cellOffsets, const int* __restrict__ batchSizes,
											const int* __restrict__ batchIdx, int ib, int phase, float friction,
											const float* __restrict__ warm, float* __restrict__ resid)
{
	int cell = cell_of(blockIdx.x, ib);
	int n = cellCounts[cell];
	if (n == 0) return;
	int maxBatch = batchSizes[cell];
	int start = cellOffsets[cell];
	int end = start + n;
	int lane = threadIdx.x;
	int idx = start + lane;
	for (int cur = 0; cur < maxBatch; cur++)
	{
		while (idx < end && batchIdx[idx] == cur)
		{
			solve_one(cs + idx, bodies, inertias, phase, friction, warm ? warm + 6 * (size_t)idx : nullptr, resid);
			idx += 64;
		}
		__syncthreads();
	}
}

static int s_ldsCapBytes = -1;

/* one-time set-up (outside any graph capture): LDS budget of the per-cell
 * batching kernel */
int b3h_solver_module_init()
{
	if (s_ldsCapBytes > 0) return B3H_OK;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) return B3H_ERR_HIP;
	int v = 0;
	if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
	if (v > 160 * 1024) v = 160 * 1024;
	int want = v - 1024;
	if (hipFuncSetAttribute((const void*)k_batch_cells, hipFuncAttributeMaxDynamicSharedMemorySize, want) != hipSuccess)
		want = 65536 - 1024;
	s_ldsCapBytes = want;
	return B3H_OK;
}
int b3h_solver_lds_bytes() { return s_ldsCapBytes; }

static inline int gblocks(int maxN)
{
	int b = (maxN + 255) / 256;
	if (b > 8192) b = 8192;
	return b < 1 ? 1 : b;
}

void b3h_solve_contacts(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias, const int* numContacts,
						int maxContacts, const b3hContact4Data* staging, const int* contactPair,
						const b3hSolverParams& prm, int staticIdx, b3hSolverBuffers& sb, bool writeSortedContacts,
						bool trackResidual, hipStream_t s)
{
	(void)numBodies;
	int ldsCapBytes = s_ldsCapBytes > 0 ? s_ldsCapBytes : 32768;
	int ldsCap = 1;
	while ((ldsCap * 2) * 16 <= ldsCapBytes) ldsCap *= 2;
	b3h_fill_i32(sb.cellCounts, 0, B3H_N_CELLS, s);
	int gb = gblocks(maxContacts);
	hipLaunchKernelGGL(k_solver_keys, dim3(gb), dim3(256), 0, s, numContacts, staging, contactPair, bodies, staticIdx,
					   1.f / prm.m_batchCellSize, sb.cellOf, sb.cellCounts);
	hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(B3H_N_CELLS), 0, s, sb.cellCounts, sb.cellOffsets, sb.cellCursor);
	hipLaunchKernelGGL(k_cell_scatter, dim3(gb), dim3(256), 0, s, numContacts, sb.cellOf, sb.cellCursor, sb.cellSlots);
	hipLaunchKernelGGL(k_batch_cells, dim3(B3H_N_CELLS), dim3(256), (size_t)ldsCap * 16, s, sb.cellCounts, sb.cellOffsets,
					   sb.cellSlots, staging, contactPair, staticIdx, ldsCap, sb.sortKeys, sb.perm, sb.batchIdx,
					   sb.batchSizes);
	hipLaunchKernelGGL(k_to_constraints, dim3(gb), dim3(256), 0, s, numContacts, sb.perm, sb.batchIdx, staging,
					   contactPair, bodies, inertias, prm, sb.constraints,
					   writeSortedContacts ? sb.sortedContacts : (b3hContact4Data*)nullptr);
	int K = prm.m_numIterations;
	if (trackResidual) b3h_fill_i32((int*)sb.resid, 0, 4 * K, s);
	if (sb.warmImpulses)
		for (int ib = 0; ib < B3H_N_CELL_BATCHES; ib++)
			hipLaunchKernelGGL(k_pgs, dim3(B3H_N_CELLS / B3H_N_CELL_BATCHES), dim3(64), 0, s, bodies, inertias,
							   sb.constraints, sb.cellCounts, sb.cellOffsets, sb.batchSizes, sb.batchIdx, ib, 2,
							   prm.m_frictionCoeff, (const float*)sb.warmImpulses, (float*)nullptr);
	for (int phase = 0; phase < 2; phase++)
		for (int it = 0; it < K; it++)
			for (int ib = 0; ib < B3H_N_CELL_BATCHES; ib++)
				hipLaunchKernelGGL(k_pgs, dim3(B3H_N_CELLS / B3H_N_CELL_BATCHES), dim3(64), 0, s, bodies, inertias,
								   sb.constraints, sb.cellCounts, sb.cellOffsets, sb.batchSizes, sb.batchIdx, ib, phase,
								   prm.m_frictionCoeff, (const float*)nullptr,
								   trackResidual ? sb.resid + 2 * (phase * K + it) : (float*)nullptr);
}
