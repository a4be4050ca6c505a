/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * Contact solver of the Bullet3 GPU pipeline restated on the CPU
 * (b3GpuPgsContactSolver::solveContacts, b3GpuPgsContactSolver.cpp:568-1103):
 *   determinism order   b3GpuPgsContactSolver.cpp:575-637, solverSetup2.cl:437-499
 *   spatial cells       SetSortDataKernel solverSetup2.cl:540-599 (8x4x8 cells, b3Solver.h:33-41)
 *   greedy batching     CreateBatchesNew batchingKernelsNew.cl:144-231
 *   constraint setup    ContactToConstraintKernel solverSetup.cl:234-277 -> setConstraint4
 *                       (b3ConvertConstraint4.h:62-148)
 *   PGS                 BatchSolveKernelContact solveContact.cl:259-482 and
 *                       BatchSolveKernelFriction solveFriction.cl:288-503, driven by
 *                       b3GpuPgsContactSolver::solveContactConstraint (:315-468)
 * The GPU kernels walk each cell's constraints batch by batch in rows of 64
 * lanes; every row is free of shared dynamic bodies, so that walk equals a
 * sequential pass over the cell in position order, cells of one parity class
 * being independent.  That sequential order is what is restated here (it is
 * also the order of the reference's host emulation, b3Solver.cpp:468-637).
 */
#include "oracle.h"
#include <vector>
#include <algorithm>
#include <stdlib.h>

#define N_SPLIT_X 8
#define N_SPLIT_Y 4
#define N_SPLIT_Z 8
#define N_CELLS (N_SPLIT_X * N_SPLIT_Y * N_SPLIT_Z)
#define N_CELL_BATCHES 8
#define CHECK_SIZE 64
#define SIMD_WIDTH 64

static int cell_index(F4 p, float scale)
{
	int xIdx = (int)((p.x - ((p.x < 0.f) ? 1.f : 0.f)) * scale) & (N_SPLIT_X - 1);
	int yIdx = (int)((p.y - ((p.y < 0.f) ? 1.f : 0.f)) * scale) & (N_SPLIT_Y - 1);
	int zIdx = (int)((p.z - ((p.z < 0.f) ? 1.f : 0.f)) * scale) & (N_SPLIT_Z - 1);
	return xIdx + yIdx * N_SPLIT_X + zIdx * N_SPLIT_X * N_SPLIT_Y;
}

/* CreateBatchesNew for one cell (batchingKernelsNew.cl:144-231) */
static int create_batches(b3hContact4Data* cs, int n, int staticIdx)
{
	unsigned buf[CHECK_SIZE];
	int numValid = 0, batchIdx = 0;
	while (numValid < n)
	{
		int nCur = 0;
		memset(buf, 0, sizeof(buf));
		for (int i = numValid; i < n; i++)
		{
			int aS = cs[i].m_bodyAPtrAndSignBit, bS = cs[i].m_bodyBPtrAndSignBit;
			int a = abs(aS), b = abs(bS);
			bool aStatic = (aS < 0) || aS == staticIdx;
			bool bStatic = (bS < 0) || bS == staticIdx;
			int ia = a % (32 * CHECK_SIZE), ib = b % (32 * CHECK_SIZE);
			unsigned aU = aStatic ? 0 : (buf[ia / 32] & (1u << (ia % 32)));
			unsigned bU = bStatic ? 0 : (buf[ib / 32] & (1u << (ib % 32)));
			if (aU == 0 && bU == 0)
			{
				if (!aStatic) buf[ia / 32] |= (1u << (ia % 32));
				if (!bStatic) buf[ib / 32] |= (1u << (ib % 32));
				cs[i].m_batchIdx = batchIdx;
				if (i != numValid)
				{
					b3hContact4Data t = cs[i];
					cs[i] = cs[numValid];
					cs[numValid] = t;
				}
				numValid++;
				nCur++;
				if (nCur == SIMD_WIDTH)
				{
					nCur = 0;
					memset(buf, 0, sizeof(buf));
				}
			}
		}
		batchIdx++;
	}
	return batchIdx;
}

/* b3PlaneSpace1 (b3ConvertConstraint4.h:6-35) / btPlaneSpace1 (solveFriction.cl:258-284) */
static void plane_space(F4 n, F4* p, F4* q)
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

static float calc_jac_coeff(F4 ang0, F4 ang1, float im0, const b3hMat3x3& ii0, float im1, const b3hMat3x3& ii1)
{
	float j0 = im0;
	float j1 = dot3(mtMul3(ang0, ii0), ang0);
	float j2 = im1;
	float j3 = dot3(mtMul3(ang1, ii1), ang1);
	return -1.f / (j0 + j1 + j2 + j3);
}
static inline float calc_rel_vel(F4 l0, F4 l1, F4 a0, F4 a1, F4 v0, F4 w0, F4 v1, F4 w1)
{
	return dot3(l0, v0) + dot3(a0, w0) + dot3(l1, v1) + dot3(a1, w1);
}

/* setConstraint4 (b3ConvertConstraint4.h:62-148) */
static void set_constraint4(const b3hRigidBodyData& A, const b3hMat3x3& iiA, const b3hRigidBodyData& B, const b3hMat3x3& iiB,
							const b3hContact4Data* src, float dt, float drift, float coeff, float friction,
							b3hContactConstraint4* dst)
{
	F4 posA = A.m_pos, posB = B.m_pos;
	dst->m_bodyA = (unsigned)abs(src->m_bodyAPtrAndSignBit);
	dst->m_bodyB = (unsigned)abs(src->m_bodyBPtrAndSignBit);
	float dtInv = 1.f / dt;
	for (int ic = 0; ic < 4; ic++) dst->m_appliedRambdaDt[ic] = 0.f;
	dst->m_fJacCoeffInv[0] = dst->m_fJacCoeffInv[1] = 0.f;
	dst->m_linear = src->m_worldNormalOnB;
	dst->m_linear.w = friction;
	float np = src->m_worldNormalOnB.w;
	for (int ic = 0; ic < 4; ic++)
	{
		F4 r0 = sub4(src->m_worldPosB[ic], posA);
		F4 r1 = sub4(src->m_worldPosB[ic], posB);
		if (ic >= np)
		{
			dst->m_jacCoeffInv[ic] = 0.f;
			dst->m_b[ic] = 0.f;
			continue;
		}
		F4 n = src->m_worldNormalOnB;
		F4 linear = f4(n.x, n.y, n.z, 0.f);
		F4 ang0 = cross3(r0, n);
		F4 ang1 = neg4(cross3(r1, n));
		dst->m_jacCoeffInv[ic] = calc_jac_coeff(ang0, ang1, A.m_invMass, iiA, B.m_invMass, iiB);
		float relVelN = calc_rel_vel(linear, neg4(linear), ang0, ang1, A.m_linVel, A.m_angVel, B.m_linVel, B.m_angVel);
		float e = 0.f;
		if (relVelN * relVelN < 0.004f) e = 0.f;
		dst->m_b[ic] = e * relVelN;
		dst->m_b[ic] += (src->m_worldPosB[ic].w + drift) * coeff * dtInv;
		dst->m_appliedRambdaDt[ic] = 0.f;
	}
	dst->m_center = f4(0, 0, 0, 0);
	dst->m_fAppliedRambdaDt[0] = dst->m_fAppliedRambdaDt[1] = 0.f;
	if (np > 0)
	{
		F4 center = f4(0.f, 0.f, 0.f, 0.f);
		for (int i = 0; i < np; i++) center = add4(center, src->m_worldPosB[i]);
		float fnp = np;
		center = f4(center.x / fnp, center.y / fnp, center.z / fnp, center.w / fnp);
		F4 t[2];
		plane_space(src->m_worldNormalOnB, &t[0], &t[1]);
		F4 r0 = sub4(center, posA), r1 = sub4(center, posB);
		for (int i = 0; i < 2; i++)
		{
			F4 ang0 = cross3(r0, t[i]);
			F4 ang1 = neg4(cross3(r1, t[i]));
			dst->m_fJacCoeffInv[i] = calc_jac_coeff(ang0, ang1, A.m_invMass, iiA, B.m_invMass, iiB);
			dst->m_fAppliedRambdaDt[i] = 0.f;
		}
		dst->m_center = center;
	}
	for (int i = 0; i < 4; i++)
	{
		if (i < np)
			dst->m_worldPos[i] = src->m_worldPosB[i];
		else
			dst->m_worldPos[i] = f4(0.f, 0.f, 0.f, 0.f);
	}
}

struct Vel { F4 lv, av; };

static void apply_impulse(F4 linear, F4 ang0, F4 ang1, float rambdaDt, float imA, const b3hMat3x3& iiA, float imB,
						  const b3hMat3x3& iiB, Vel& a, Vel& b)
{
	F4 linImp0 = mul4(smul4(imA, linear), rambdaDt);
	F4 linImp1 = mul4(smul4(imB, neg4(linear)), rambdaDt);
	F4 angImp0 = mul4(mtMul1(iiA, ang0), rambdaDt);
	F4 angImp1 = mul4(mtMul1(iiB, ang1), rambdaDt);
	a.lv = add4(a.lv, linImp0);
	a.av = add4(a.av, angImp0);
	b.lv = add4(b.lv, linImp1);
	b.av = add4(b.av, angImp1);
}

struct Resid { float maxDelta, maxLambda; };

/* solveContact (solveContact.cl:259-303) */
static void solve_contact(b3hContactConstraint4& cs, const b3hRigidBodyData& A, const b3hMat3x3& iiA,
						  const b3hRigidBodyData& B, const b3hMat3x3& iiB, Vel& va, Vel& vb, Resid& res)
{
	for (int ic = 0; ic < 4; ic++)
	{
		if (cs.m_jacCoeffInv[ic] == 0.f) continue;
		F4 r0 = sub4(cs.m_worldPos[ic], A.m_pos);
		F4 r1 = sub4(cs.m_worldPos[ic], B.m_pos);
		F4 n = neg4(cs.m_linear);
		F4 linear = f4(-n.x, -n.y, -n.z, 0.f);
		F4 ang0 = neg4(cross3(r0, n));
		F4 ang1 = cross3(r1, n);
		float rl = calc_rel_vel(cs.m_linear, neg4(cs.m_linear), ang0, ang1, va.lv, va.av, vb.lv, vb.av) + cs.m_b[ic];
		rl *= cs.m_jacCoeffInv[ic];
		float prev = cs.m_appliedRambdaDt[ic];
		float upd = prev;
		upd += rl;
		upd = cl_max(upd, 0.f);
		upd = cl_min(upd, FLT_MAX);
		rl = upd - prev;
		cs.m_appliedRambdaDt[ic] = upd;
		float ad = fabsf(rl), al = fabsf(upd);
		if (ad > res.maxDelta) res.maxDelta = ad;
		if (al > res.maxLambda) res.maxLambda = al;
		apply_impulse(linear, ang0, ang1, rl, A.m_invMass, iiA, B.m_invMass, iiB, va, vb);
	}
}

/* solveFrictionConstraint body (solveFriction.cl:288-408); returns false when the
 * constraint has no friction rows (then the bodies are not written). */
static bool solve_friction(b3hContactConstraint4& cs, float friction, const b3hRigidBodyData& A, const b3hMat3x3& iiA,
						   const b3hRigidBodyData& B, const b3hMat3x3& iiB, Vel& va, Vel& vb, Resid& res)
{
	float sum = 0;
	for (int j = 0; j < 4; j++) sum += cs.m_appliedRambdaDt[j];
	float maxR = friction * sum;
	float minR = -maxR;
	if (cs.m_fJacCoeffInv[0] == 0 && cs.m_fJacCoeffInv[0] == 0) return false;
	F4 center = cs.m_center;
	F4 n = neg4(cs.m_linear);
	F4 t[2];
	plane_space(n, &t[0], &t[1]);
	F4 r0 = sub4(center, A.m_pos), r1 = sub4(center, B.m_pos);
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
		float ad = fabsf(rl), al = fabsf(upd);
		if (ad > res.maxDelta) res.maxDelta = ad;
		if (al > res.maxLambda) res.maxLambda = al;
		apply_impulse(linear, ang0, ang1, rl, A.m_invMass, iiA, B.m_invMass, iiB, va, vb);
	}
	{
		F4 ab = nrm3(sub4(B.m_pos, A.m_pos));
		F4 ac = nrm3(sub4(center, A.m_pos));
		if (dot3(ab, ac) > 0.95f || (A.m_invMass == 0.f || B.m_invMass == 0.f))
		{
			float angNA = dot3(n, va.av);
			float angNB = dot3(n, vb.av);
			va.av = sub4(va.av, smul4(angNA * 0.1f, n));
			vb.av = sub4(vb.av, smul4(angNB * 0.1f, n));
		}
	}
	return true;
}

static void write_back(b3hRigidBodyData& body, const Vel& v)
{
	if (body.m_invMass)
	{
		body.m_linVel = v.lv;
		body.m_angVel = v.av;
	}
	else
	{
		body.m_linVel = f4(0, 0, 0, 0);
		body.m_angVel = f4(0, 0, 0, 0);
	}
}

/* cell visited by work-group wgIdx of cell batch ib (solveContact.cl:425-429) */
static int cell_of(int wgIdx, int ib)
{
	int zIdx = (wgIdx / ((N_SPLIT_X * N_SPLIT_Y) / 4)) * 2 + ((ib & 4) >> 2);
	int remain = (wgIdx % ((N_SPLIT_X * N_SPLIT_Y) / 4));
	int yIdx = (remain / (N_SPLIT_X / 2)) * 2 + ((ib & 2) >> 1);
	int xIdx = (remain % (N_SPLIT_X / 2)) * 2 + (ib & 1);
	return xIdx + yIdx * N_SPLIT_X + zIdx * (N_SPLIT_X * N_SPLIT_Y);
}

/* Warm start: constraint impulses start from initialImpulses (6 floats per
 * constraint: 4 normal rows, 2 friction rows), applied to the bodies once in
 * solve order: the normal rows before the first normal sweep (part 0), the
 * friction rows -- clamped by the normal impulses the normal sweeps left --
 * before the first friction sweep (part 1).  As the normal sweeps all run
 * before the friction sweeps, each phase then starts from its own guess and
 * keeps the fixed point of the cold start. */
static void warm_start(b3hContactConstraint4& cs, const float* l0, float friction, b3hRigidBodyData* bodies,
					   const b3hInertiaData* inertias, int part)
{
	b3hRigidBodyData& A = bodies[cs.m_bodyA];
	b3hRigidBodyData& B = bodies[cs.m_bodyB];
	const b3hMat3x3& iiA = inertias[cs.m_bodyA].m_invInertiaWorld;
	const b3hMat3x3& iiB = inertias[cs.m_bodyB].m_invInertiaWorld;
	Vel va = {A.m_linVel, A.m_angVel}, vb = {B.m_linVel, B.m_angVel};
	for (int ic = 0; ic < 4 && part == 0; ic++)
	{
		if (cs.m_jacCoeffInv[ic] == 0.f) continue;
		F4 r0 = sub4(cs.m_worldPos[ic], A.m_pos);
		F4 r1 = sub4(cs.m_worldPos[ic], B.m_pos);
		F4 n = neg4(cs.m_linear);
		F4 linear = f4(-n.x, -n.y, -n.z, 0.f);
		F4 ang0 = neg4(cross3(r0, n));
		F4 ang1 = cross3(r1, n);
		float l = cl_max(l0[ic], 0.f);
		cs.m_appliedRambdaDt[ic] = l;
		apply_impulse(linear, ang0, ang1, l, A.m_invMass, iiA, B.m_invMass, iiB, va, vb);
	}
	if (part == 1 && !(cs.m_fJacCoeffInv[0] == 0))
	{
		float sum = 0;
		for (int j = 0; j < 4; j++) sum += cs.m_appliedRambdaDt[j];
		float maxR = friction * sum;
		F4 n = neg4(cs.m_linear);
		F4 t[2];
		plane_space(n, &t[0], &t[1]);
		F4 r0 = sub4(cs.m_center, A.m_pos), r1 = sub4(cs.m_center, B.m_pos);
		for (int i = 0; i < 2; i++)
		{
			F4 linear = f4(-t[i].x, -t[i].y, -t[i].z, 0.f);
			F4 ang0 = neg4(cross3(r0, t[i]));
			F4 ang1 = cross3(r1, t[i]);
			float l = cl_min(cl_max(l0[4 + i], -maxR), maxR);
			cs.m_fAppliedRambdaDt[i] = l;
			apply_impulse(linear, ang0, ang1, l, A.m_invMass, iiA, B.m_invMass, iiB, va, vb);
		}
	}
	write_back(A, va);
	write_back(B, vb);
}

extern "C" int b3o_solve_contacts(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias,
								   int numContacts, const b3hContact4Data* contactsIn, const b3hSolverParams* prm,
								   int staticIdx, b3hContact4Data* sortedContactsOut, b3hContactConstraint4* constraintsOut,
								   unsigned* cellCountsOut, unsigned* cellOffsetsOut, int* batchSizesOut,
								   const float* initialImpulses, float* residualsOut)
{
	(void)numBodies;
	int n = numContacts;
	/* 1. determinism order: (bodyA, bodyB, childA, childB) as uint32, stable */
	std::vector<int> order(n);
	for (int i = 0; i < n; i++) order[i] = i;
	std::stable_sort(order.begin(), order.end(), [&](int i, int j) {
		const b3hContact4Data& a = contactsIn[i];
		const b3hContact4Data& b = contactsIn[j];
		unsigned ka[4] = {(unsigned)a.m_bodyAPtrAndSignBit, (unsigned)a.m_bodyBPtrAndSignBit, (unsigned)a.m_childIndexA, (unsigned)a.m_childIndexB};
		unsigned kb[4] = {(unsigned)b.m_bodyAPtrAndSignBit, (unsigned)b.m_bodyBPtrAndSignBit, (unsigned)b.m_childIndexA, (unsigned)b.m_childIndexB};
		for (int k = 0; k < 4; k++)
			if (ka[k] != kb[k]) return ka[k] < kb[k];
		return false;
	});
	std::vector<b3hContact4Data> det(n);
	for (int i = 0; i < n; i++) det[i] = contactsIn[order[i]];
	/* 2. spatial cells of the dynamic body, stable sort by cell */
	float scale = 1.f / prm->m_batchCellSize;
	std::vector<b3hSortData> kv(n);
	for (int i = 0; i < n; i++)
	{
		int aS = det[i].m_bodyAPtrAndSignBit, bS = det[i].m_bodyBPtrAndSignBit;
		int aIdx = abs(aS), bIdx = abs(bS);
		bool aStatic = (aS < 0) || (aS == staticIdx);
		int idx = aStatic ? bIdx : aIdx;
		kv[i].m_key = (unsigned)cell_index(bodies[idx].m_pos, scale);
		kv[i].m_value = (unsigned)i;
	}
	b3o_sort_keys(n, kv.data());
	unsigned counts[N_CELLS], offsets[N_CELLS];
	memset(counts, 0, sizeof(counts));
	for (int i = 0; i < n; i++) counts[kv[i].m_key]++;
	unsigned run = 0;
	for (int c = 0; c < N_CELLS; c++) { offsets[c] = run; run += counts[c]; }
	std::vector<b3hContact4Data> cs(n);
	for (int i = 0; i < n; i++) cs[i] = det[kv[i].m_value];
	/* 3. batching per cell */
	int batchSizes[N_CELLS];
	const int threads = b3o_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
	for (int c = 0; c < N_CELLS; c++) batchSizes[c] = counts[c] ? create_batches(&cs[offsets[c]], (int)counts[c], staticIdx) : 0;
	/* 4. constraints (initial local inverse inertia, solverSetup.cl:254,260) */
	std::vector<b3hContactConstraint4> con(n);
#pragma omp parallel for num_threads(threads) schedule(static)
	for (int i = 0; i < n; i++)
	{
		int a = abs(cs[i].m_bodyAPtrAndSignBit), b = abs(cs[i].m_bodyBPtrAndSignBit);
		set_constraint4(bodies[a], inertias[a].m_initInvInertia, bodies[b], inertias[b].m_initInvInertia, &cs[i],
						prm->m_solverDt, prm->m_positionDrift, prm->m_positionConstraintCoeff, prm->m_frictionCoeff, &con[i]);
		con[i].m_batchIdx = cs[i].m_batchIdx;
		con[i].m_paddings = 0;
	}
	/* 5. PGS: all normal sweeps, then all friction sweeps, each phase after
	 * its part of the warm start */
	int K = prm->m_numIterations;
	for (int phase = 0; phase < 2; phase++)
	{
		/* the cells of one parity class share no dynamic body (they are a cell
		 * apart and bodies are smaller than a cell), so they may run at once */
		if (initialImpulses)
			for (int ib = 0; ib < N_CELL_BATCHES; ib++)
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
				for (int wg = 0; wg < N_CELLS / N_CELL_BATCHES; wg++)
				{
					int cell = cell_of(wg, ib);
					for (unsigned k = 0; k < counts[cell]; k++)
					{
						int i = (int)(offsets[cell] + k);
						warm_start(con[i], initialImpulses + 6 * i, prm->m_frictionCoeff, bodies, inertias, phase);
					}
				}
		for (int it = 0; it < K; it++)
		{
			float maxDelta = 0.f, maxLambda = 0.f;
			for (int ib = 0; ib < N_CELL_BATCHES; ib++)
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(max : maxDelta, maxLambda)
				for (int wg = 0; wg < N_CELLS / N_CELL_BATCHES; wg++)
				{
					Resid res = {0.f, 0.f};
					int cell = cell_of(wg, ib);
					for (unsigned k = 0; k < counts[cell]; k++)
					{
						b3hContactConstraint4& c = con[offsets[cell] + k];
						b3hRigidBodyData& A = bodies[c.m_bodyA];
						b3hRigidBodyData& B = bodies[c.m_bodyB];
						const b3hMat3x3& iiA = inertias[c.m_bodyA].m_invInertiaWorld;
						const b3hMat3x3& iiB = inertias[c.m_bodyB].m_invInertiaWorld;
						Vel va = {A.m_linVel, A.m_angVel}, vb = {B.m_linVel, B.m_angVel};
						if (phase == 0)
							solve_contact(c, A, iiA, B, iiB, va, vb, res);
						else if (!solve_friction(c, prm->m_frictionCoeff, A, iiA, B, iiB, va, vb, res))
							continue;
						write_back(A, va);
						write_back(B, vb);
					}
					maxDelta = res.maxDelta > maxDelta ? res.maxDelta : maxDelta;
					maxLambda = res.maxLambda > maxLambda ? res.maxLambda : maxLambda;
				}
			if (residualsOut)
			{
				float denom = maxLambda > 1e-9f ? maxLambda : 1e-9f;
				residualsOut[phase * K + it] = maxDelta / denom;
			}
		}
	}
	if (sortedContactsOut) memcpy(sortedContactsOut, cs.data(), sizeof(b3hContact4Data) * n);
	if (constraintsOut) memcpy(constraintsOut, con.data(), sizeof(b3hContactConstraint4) * n);
	if (cellCountsOut) memcpy(cellCountsOut, counts, sizeof(counts));
	if (cellOffsetsOut) memcpy(cellOffsetsOut, offsets, sizeof(offsets));
	if (batchSizesOut) memcpy(batchSizesOut, batchSizes, sizeof(batchSizes));
	return n;
}
