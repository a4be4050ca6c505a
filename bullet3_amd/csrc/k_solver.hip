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

/* LDS ordering between lanes of one wavefront */
__device__ __forceinline__ void wave_sync_solver()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
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
	/* per-block histogram in LDS, one global add per (block, cell) */
	__shared__ int hist[B3H_N_CELLS];
	int t = threadIdx.x;
	hist[t] = 0;
	__syncthreads();
	int n = *nPtr;
	for (int i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x)
	{
		const b3hContact4Data& c = staging[contactPair[i]];
		int aS = c.m_bodyAPtrAndSignBit, bS = c.m_bodyBPtrAndSignBit;
		bool aStatic = (aS < 0) || (aS == staticIdx);
		int idx = aStatic ? iabs(bS) : iabs(aS);
		int cell = cell_index(bodies[idx].m_pos, scale);
		cellOf[i] = cell;
		atomicAdd(&hist[cell], 1);
	}
	__syncthreads();
	if (hist[t]) atomicAdd(&cellCounts[t], hist[t]);
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
	/* each block reserves one range per cell (one global add per (block, cell))
	 * and places its contacts inside it by LDS ranks; the order inside a cell is
	 * arbitrary here and fixed by the per-cell sort that follows */
	__shared__ int cnt[B3H_N_CELLS];
	__shared__ int base[B3H_N_CELLS];
	int t = threadIdx.x;
	int n = *nPtr;
	const int stride = gridDim.x * blockDim.x;
	for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += stride)
	{
		cnt[t] = 0;
		__syncthreads();
		int i = i0 + t;
		int cell = -1, rank = 0;
		if (i < n)
		{
			cell = cellOf[i];
			rank = atomicAdd(&cnt[cell], 1);
		}
		__syncthreads();
		if (cnt[t]) base[t] = atomicAdd(&cursor[t], cnt[t]);
		__syncthreads();
		if (cell >= 0) cellSlots[base[cell] + rank] = i;
		__syncthreads();
	}
}

#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) unsigned long long LdsU64;
typedef __attribute__((address_space(3))) unsigned LdsU32;
#define LDS_U64(p) ((LdsU64*)(p))
#define LDS_U32(p) ((LdsU32*)(p))
#else
#define LDS_U64(p) (p)
#define LDS_U32(p) (p)
#endif
/* ascending sort of n keys with their payloads (KP, PP: LDS or global
 * pointers), with no padding: the bitonic network of the next power of two in
 * its all-ascending form (the first step of every merge compares mirrored
 * positions), where the absent positions n.. act as the largest keys and are
 * never moved, so they are never stored */
template <typename KP, typename PP>
__device__ __forceinline__ void bitonic_sort_n(KP key, PP pay, int n, int t)
{
	int np2 = 1;
	while (np2 < n) np2 <<= 1;
	for (int k = 2; k <= np2; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1)
		{
			const int mask = j == (k >> 1) ? k - 1 : j;
			for (int i = t; i < n; i += blockDim.x)
			{
				int l = i ^ mask;
				if (l > i && l < n)
				{
					unsigned long long ki = key[i], kl = key[l];
					/* equal keys (a compound's child contacts) keep slot order */
					if (ki > kl || (ki == kl && pay[i] > pay[l]))
					{
						key[i] = kl;
						key[l] = ki;
						unsigned p = pay[i];
						pay[i] = pay[l];
						pay[l] = p;
					}
				}
			}
			__syncthreads();
		}
}
/* the greedy pass of k_batch_cells on one wavefront (KP, PP: LDS or global
 * pointers to the keys and payloads); a contact's batch is stored (batchOut)
 * when it is accepted, at its final position; returns the number of batches */
template <typename KP, typename PP>
__device__ __forceinline__ int batch_greedy(KP key, PP pay, int n, int staticIdx, int* __restrict__ batchOut)
{
	const int lane = threadIdx.x & 63;
	int numValid = 0, batch = 0;
	while (numValid < n)
	{
		int nCur = 0;
		unsigned word = 0u; /* bitmap word `lane` */
		/* the next chunk's keys are read while this one is resolved: the swaps
		 * write only positions below the end of the current chunk */
		unsigned long long kkNext = (numValid + lane < n) ? key[numValid + lane] : 0ull;
		for (int c0 = numValid; c0 < n; c0 += 64)
		{
			int i = c0 + lane;
			bool valid = i < n;
			unsigned long long kk = kkNext;
			if (i + 64 < n) kkNext = key[i + 64];
			int aS = (int)(unsigned)(kk >> 32), bS = (int)(unsigned)(kk & 0xffffffffu);
			int a = iabs(aS), b = iabs(bS);
			bool aStatic = (aS < 0) || aS == staticIdx;
			bool bStatic = (bS < 0) || bS == staticIdx;
			/* body ids are below 2^31, so the remainder is the low 11 bits */
			int ia = (int)((unsigned)a & (32u * CHECK_SIZE - 1u)), ib = (int)((unsigned)b & (32u * CHECK_SIZE - 1u));
			int wa = ia >> 5, wb = ib >> 5;
			unsigned ma = aStatic ? 0u : (1u << (ia & 31));
			unsigned mb = bStatic ? 0u : (1u << (ib & 31));
			unsigned wA = __shfl(word, wa), wB = __shfl(word, wb);
			bool pre = valid && (wA & ma) == 0u && (wB & mb) == 0u;
			unsigned long long cand = __ballot(pre);
			unsigned long long acc = 0ull;
			const unsigned long long validMask = __ballot(valid);
			/* bitmap indices for the in-chunk conflict test: two contacts
			 * conflict when a dynamic body of one has the index of a dynamic
			 * body of the other; static bodies take sentinels that match
			 * nothing (4096 on this lane, 8192 as the accepted one) */
			const unsigned myA = aStatic ? 4096u : (unsigned)ia, myB = bStatic ? 4096u : (unsigned)ib;
			const unsigned pk = (aStatic ? 8192u : (unsigned)ia) | ((bStatic ? 8192u : (unsigned)ib) << 16);
			while (cand)
			{
				/* j is the same on every lane: read its indices with one readlane */
				int j = __ffsll((long long)cand) - 1;
				acc |= 1ull << j;
				const unsigned jp = (unsigned)__builtin_amdgcn_readlane((int)pk, j);
				const unsigned jA = jp & 0xffffu, jB = jp >> 16;
				/* a sentinel's word (256) is no lane */
				if (lane == (int)(jA >> 5)) word |= 1u << (jA & 31u);
				if (lane == (int)(jB >> 5)) word |= 1u << (jB & 31u);
				nCur++;
				unsigned long long later = (j == 63) ? 0ull : (~0ull << (j + 1));
				if (nCur == SIMD_WIDTH)
				{
					nCur = 0;
					word = 0u;
					cand = validMask & later;
				}
				else
				{
					/* bitwise, without short-circuit branches */
					bool hit = (myA == jA) | (myA == jB) | (myB == jA) | (myB == jB);
					cand &= later & ~__ballot(hit);
				}
			}
			/* move the accepted contacts to the front, in order.  When at least
			 * as many rejected contacts lie before the chunk as the chunk
			 * accepts, the swaps (accepted m <-> numValid + m) are disjoint and
			 * run in parallel; otherwise one lane performs them in sequence. */
			int kAcc = __popcll(acc);
			if (c0 - numValid >= kAcc)
			{
				if ((acc >> lane) & 1ull)
				{
					int m = __popcll(acc & ((1ull << lane) - 1ull));
					int dst = numValid + m;
					int ii = c0 + lane;
					unsigned pj = pay[ii];
					unsigned long long kd = key[dst];
					unsigned pd = pay[dst];
					key[ii] = kd;
					pay[ii] = pd;
					key[dst] = kk;
					pay[dst] = pj;
					batchOut[dst] = batch;
				}
				numValid += kAcc;
			}
			else if (lane == 0)
			{
				unsigned long long m = acc;
				while (m)
				{
					int j = __ffsll((long long)m) - 1;
					m &= m - 1;
					int ii = c0 + j;
					if (ii != numValid)
					{
						unsigned long long kj = key[ii];
						unsigned pj = pay[ii];
						key[ii] = key[numValid];
						pay[ii] = pay[numValid];
						key[numValid] = kj;
						pay[numValid] = pj;
					}
					batchOut[numValid] = batch;
					numValid++;
				}
			}
			numValid = __builtin_amdgcn_readfirstlane(numValid);
			wave_sync_solver();
		}
		batch++;
	}
	return batch;
}

/* One work-group per cell: sort the cell's (uint bodyA, uint bodyB) keys, then
 * CreateBatchesNew's greedy batching.  Keys (8 bytes) and contact slots (4
 * bytes) live in LDS when the cell fits (ldsBytes), otherwise in the cell's
 * region of the global scratch (4 keys per contact). */
__global__ __launch_bounds__(256) void k_batch_cells(const int* __restrict__ cellCounts, const int* __restrict__ cellOffsets,
													 const int* __restrict__ cellSlots,
													 const b3hContact4Data* __restrict__ staging,
													 const int* __restrict__ contactPair, int staticIdx, int ldsBytes,
													 unsigned long long* __restrict__ gKeys, int* __restrict__ perm,
													 int* __restrict__ batchIdxOut, int* __restrict__ batchSizes)
{
	extern __shared__ unsigned long long smem[];
	int cell = blockIdx.x;
	int n = cellCounts[cell];
	int off = cellOffsets[cell];
	int t = threadIdx.x;
	if (n == 0)
	{
		if (t == 0) batchSizes[cell] = 0;
		return;
	}
	bool inLds = (size_t)n * 12 <= (size_t)ldsBytes;
	/* key = uint bodyA << 32 | uint bodyB ; payload = contact slot (parallel array) */
	unsigned long long* key = inLds ? smem : gKeys + 4 * (size_t)off;
	unsigned* pay = (unsigned*)(key + n);
	for (int i = t; i < n; i += blockDim.x)
	{
		int slot = cellSlots[off + i];
		const b3hContact4Data& c = staging[contactPair[slot]];
		key[i] = ((unsigned long long)(unsigned)c.m_bodyAPtrAndSignBit << 32) | (unsigned)c.m_bodyBPtrAndSignBit;
		pay[i] = (unsigned)slot;
	}
	__syncthreads();
	if (inLds)
		bitonic_sort_n(LDS_U64(key), LDS_U32(pay), n, t);
	else
		bitonic_sort_n(key, pay, n, t);
	/* greedy batching, wave-parallel with the sequential semantics of
	 * CreateBatchesNew: the 2048-bit bitmap is held one 32-bit word per lane;
	 * each chunk of 64 candidates is pre-checked against the bitmap at chunk
	 * start, then acceptances are resolved in candidate order (lowest pending
	 * lane first; later candidates that hit its bits drop out; after every 64th
	 * accepted contact the bitmap is cleared and the later lanes are
	 * re-examined against the empty map).  Accepted contacts are then swapped
	 * to the front one by one, exactly as the reference moves them (that
	 * permutes the rejected contacts, and so the next pass's scan order).
	 * Positions not yet examined are never written by the swaps, so each chunk
	 * can be read up front. */
	if (t < 64)
	{
		int batch;
		if (inLds)
			batch = batch_greedy(LDS_U64(key), LDS_U32(pay), n, staticIdx, batchIdxOut + off);
		else
			batch = batch_greedy(key, pay, n, staticIdx, batchIdxOut + off);
		if (t == 0) batchSizes[cell] = batch;
	}
	__syncthreads();
	for (int i = t; i < n; i += blockDim.x) perm[off + i] = (int)pay[i];
}

/* ---------------- contact -> constraint (setConstraint4, b3ConvertConstraint4.h:62-148) ---------------- */
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
/* v + d*s on x, y, z (multiply, then add: the order of mul4 then add4).  The w
 * lane is carried unchanged: the reference adds d.w*s = (+-0)*s to it, which
 * leaves every velocity w lane at the +0 that registration and integration
 * give it, so skipping that add changes no bit. */
__device__ __forceinline__ F4 add_scaled3(F4 v, F4 d, float s)
{
	return f4(v.x + d.x * s, v.y + d.y * s, v.z + d.z * s, v.w);
}
__device__ __forceinline__ void apply_impulse(F4 linear, F4 ang0, F4 ang1, float rl, float imA, const b3hMat3x3& iiA,
											  float imB, const b3hMat3x3& iiB, Vel& a, Vel& b)
{
	a.lv = add_scaled3(a.lv, smul4(imA, linear), rl);
	a.av = add_scaled3(a.av, mtMul1(iiA, ang0), rl);
	b.lv = add_scaled3(b.lv, smul4(imB, neg4(linear)), rl);
	b.av = add_scaled3(b.av, mtMul1(iiB, ang1), rl);
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
 * (solveFriction.cl:288-408); 2 / 3: warm start of the normal / friction rows
 * (apply the initial impulses once) */
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
	}
	else if (phase == 3)
	{
		if (!(cs.m_fJacCoeffInv[0] == 0))
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
	if (phase == 0 || phase == 2)
		for (int ic = 0; ic < 4; ic++) csp->m_appliedRambdaDt[ic] = cs.m_appliedRambdaDt[ic];
	else
		for (int i = 0; i < 2; i++) csp->m_fAppliedRambdaDt[i] = cs.m_fAppliedRambdaDt[i];
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

/* ---------------- per-cell body lists (for the LDS velocity cache) ----------------
 * For every cell: its distinct dynamic bodies in ascending order (at
 * cellBodies + 2*cellOffsets[cell]) and, per constraint, the positions of
 * bodyA / bodyB in that list (-1 for a static body).  One work-group per cell:
 * bitonic sort of the 2n body references in LDS (global scratch when the cell
 * is larger than the LDS budget), a block scan of the run heads, then a binary
 * search per reference. */
__global__ __launch_bounds__(256) void k_cell_bodies(const int* __restrict__ cellCounts, const int* __restrict__ cellOffsets,
													 const b3hContactConstraint4* __restrict__ cs,
													 const b3hRigidBodyData* __restrict__ bodies, int numBodies, int ldsCapKeys,
													 unsigned* __restrict__ gScratch, int* __restrict__ cellBodies,
													 int* __restrict__ cellBodyCount, int* __restrict__ cslot,
													 int* __restrict__ lastClass)
{
	extern __shared__ unsigned long long smem[];
	__shared__ int waveSums[4];
	__shared__ int runTotal;
	int cell = blockIdx.x;
	int n = cellCounts[cell];
	int off = cellOffsets[cell];
	int t = threadIdx.x;
	/* launch order of the cell within a sweep: its 2x2x2 parity class */
	const int cellClass = (cell & 1) | (((cell / B3H_N_SPLIT_X) & 1) << 1) |
						  (((cell / (B3H_N_SPLIT_X * B3H_N_SPLIT_Y)) & 1) << 2);
	if (n == 0)
	{
		if (t == 0) cellBodyCount[cell] = 0;
		return;
	}
	int m = 2 * n;
	int words = (numBodies + 31) >> 5;
	if (2 * words <= ldsCapKeys)
	{
		/* a bitmap over all body indices: mark the cell's dynamic bodies, then
		 * the rank of a body among them is the popcount of the bits before it
		 * (ascending order, no sort) */
		unsigned* bm = (unsigned*)smem;
		int* wp = (int*)(bm + words); /* bits set before each word */
		for (int i = t; i < words; i += blockDim.x) bm[i] = 0u;
		__syncthreads();
		for (int i = t; i < m; i += blockDim.x)
		{
			const b3hContactConstraint4& c = cs[off + (i >> 1)];
			unsigned body = (i & 1) ? c.m_bodyB : c.m_bodyA;
			if (bodies[body].m_invMass != 0.f) atomicOr(&bm[body >> 5], 1u << (body & 31));
		}
		__syncthreads();
		int per = (words + (int)blockDim.x - 1) / (int)blockDim.x;
		int w0 = t * per, w1 = w0 + per < words ? w0 + per : words;
		int cnt = 0;
		for (int w = w0; w < w1; w++) cnt += __popc(bm[w]);
		/* exclusive scan of the per-thread counts over the block */
		int lane = t & 63, wave = t >> 6;
		int incl = cnt;
		for (int o = 1; o < 64; o <<= 1)
		{
			int v = __shfl_up(incl, o);
			if (lane >= o) incl += v;
		}
		if (lane == 63) waveSums[wave] = incl;
		__syncthreads();
		int base = incl - cnt;
		for (int w = 0; w < wave; w++) base += waveSums[w];
		int total = waveSums[0] + waveSums[1] + waveSums[2] + waveSums[3];
		for (int w = w0; w < w1; w++)
		{
			wp[w] = base;
			unsigned bits = bm[w];
			while (bits)
			{
				int b = __ffs(bits) - 1;
				bits &= bits - 1u;
				int body = w * 32 + b;
				cellBodies[2 * (size_t)off + base] = body;
				if (lastClass) atomicMax(&lastClass[body], cellClass);
				base++;
			}
		}
		if (t == 0) cellBodyCount[cell] = total;
		__syncthreads();
		for (int i = t; i < m; i += blockDim.x)
		{
			const b3hContactConstraint4& c = cs[off + (i >> 1)];
			unsigned body = (i & 1) ? c.m_bodyB : c.m_bodyA;
			int slot = -1;
			if (bodies[body].m_invMass != 0.f)
				slot = wp[body >> 5] + __popc(bm[body >> 5] & ((1u << (body & 31)) - 1u));
			cslot[2 * (size_t)(off + (i >> 1)) + (i & 1)] = slot;
		}
		return;
	}
	int np2 = 1;
	while (np2 < m) np2 <<= 1;
	/* keys[np2] then rank[np2]; global fallback region of 8n >= 2*np2 entries */
	unsigned* key = (2 * np2 <= ldsCapKeys) ? (unsigned*)smem : gScratch + 8 * (size_t)off;
	unsigned* rank = key + np2;
	for (int i = t; i < np2; i += blockDim.x)
	{
		unsigned k = ~0u;
		if (i < m)
		{
			const b3hContactConstraint4& c = cs[off + (i >> 1)];
			unsigned body = (i & 1) ? c.m_bodyB : c.m_bodyA;
			if (bodies[body].m_invMass != 0.f) k = body;
		}
		key[i] = k;
	}
	__syncthreads();
	for (int k = 2; k <= np2; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1)
		{
			for (int i = t; i < np2; i += blockDim.x)
			{
				int ixj = i ^ j;
				if (ixj > i)
				{
					unsigned a = key[i], b = key[ixj];
					if ((a > b) == ((i & k) == 0))
					{
						key[i] = b;
						key[ixj] = a;
					}
				}
			}
			__syncthreads();
		}
	/* rank of each sorted entry among the distinct values (run heads scanned) */
	if (t == 0) runTotal = 0;
	__syncthreads();
	int lane = t & 63, wave = t >> 6;
	for (int c0 = 0; c0 < np2; c0 += blockDim.x)
	{
		int i = c0 + t;
		unsigned k = i < np2 ? key[i] : ~0u;
		bool head = k != ~0u && (i == 0 || key[i - 1] != k);
		unsigned long long bal = __ballot(head);
		int inWave = __popcll(bal & ((1ull << lane) - 1ull));
		if (lane == 0) waveSums[wave] = __popcll(bal);
		__syncthreads();
		int before = runTotal;
		for (int w = 0; w < wave; w++) before += waveSums[w];
		int r = before + inWave + (head ? 0 : -1);
		if (i < np2) rank[i] = (unsigned)r;
		if (head)
		{
			cellBodies[2 * (size_t)off + r] = (int)k;
			if (lastClass) atomicMax(&lastClass[k], cellClass);
		}
		__syncthreads();
		if (t == 0)
			for (int w = 0; w < 4; w++) runTotal += waveSums[w];
		__syncthreads();
	}
	int total = runTotal;
	if (t == 0) cellBodyCount[cell] = total;
	/* slots: first occurrence of the body in the sorted keys -> its rank */
	for (int i = t; i < m; i += blockDim.x)
	{
		const b3hContactConstraint4& c = cs[off + (i >> 1)];
		unsigned body = (i & 1) ? c.m_bodyB : c.m_bodyA;
		int slot = -1;
		if (bodies[body].m_invMass != 0.f)
		{
			int lo = 0, hi = np2;
			while (lo < hi)
			{
				int mid = (lo + hi) >> 1;
				if (key[mid] < body)
					lo = mid + 1;
				else
					hi = mid;
			}
			slot = (int)rank[lo];
		}
		cslot[2 * (size_t)(off + (i >> 1)) + (i & 1)] = slot;
	}
}

/* ---------------- PGS: per-step preparation ----------------
 * During the solve only velocities and impulses change, so every vector of a
 * row that depends on positions, masses and inertias alone is computed once per
 * step with exactly the operations the sweep would use (solveContact.cl:259-303,
 * solveFriction.cl:288-408), packed without padding:
 *   normal record   lin = (n, imA) | 4 points x (ang0.xyz, ang1.xyz) in 6
 *                   float4 | jacCoeffInv | b | meta (imB, slotA, slotB, index)
 *                                                                     10 float4
 *   friction record 2 directions x (lin.xyz, ang0.xyz, ang1.xyz), n.xyz in 6
 *                   float4 | (fJac0, fJac1, twist-damping flag, imA) | meta
 *                                                                      8 float4
 * The sweep evaluates relVel = dot3(lin,vA)+dot3(ang0,wA)+dot3(-lin,vB)+dot3(ang1,wB)
 * and v += imA*lin*rl, w += (iiA*ang0)*rl with the bodies' inverse inertias
 * cached in LDS: the same operations, so the same roundings, as recomputing
 * everything from the constraint. */
#define PGS_NREC 16
#define PGS_FREC 11
#define PGS_GATHER 8
#ifndef PGS_THREADS
#define PGS_THREADS 256
#endif
#define PGS_NSLOT (PGS_NREC + 1) /* record + applied normal impulses */
#define PGS_FSLOT (PGS_FREC + 2) /* record + normal impulses + friction impulses */
__device__ __forceinline__ float ibits(int v) { return __builtin_bit_cast(float, v); }
__device__ __forceinline__ int fbits(float v) { return __builtin_bit_cast(int, v); }
__global__ __launch_bounds__(256) void k_pgs_prep(const int* __restrict__ nPtr, const b3hContactConstraint4* __restrict__ cs,
												  const int* __restrict__ cslot, const b3hRigidBodyData* __restrict__ bodies,
												  const b3hInertiaData* __restrict__ inertias, const int* __restrict__ posOf,
												  int stride, F4* __restrict__ recN, F4* __restrict__ recF,
												  F4* __restrict__ lamN, F4* __restrict__ lamF)
{
	int n = *nPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const int p = posOf[i];
		if (p < 0) continue; /* a cell on the global walk: no records */
		const b3hContactConstraint4 c = cs[i];
		int a = (int)c.m_bodyA, b = (int)c.m_bodyB;
		F4 posA = bodies[a].m_pos, posB = bodies[b].m_pos;
		float imA = bodies[a].m_invMass, imB = bodies[b].m_invMass;
		int sA = cslot[2 * (size_t)i], sB = cslot[2 * (size_t)i + 1];
		F4 meta = f4(imB, ibits(sA), ibits(sB), ibits(p));
		/* the sweep's inverse world inertias: the body's, zero for a static
		 * body (slot -1) */
		const F4 z4 = f4(0.f, 0.f, 0.f, 0.f);
		b3hMat3x3 iiA, iiB;
		if (sA >= 0)
			iiA = inertias[a].m_invInertiaWorld;
		else
			iiA.m_row[0] = iiA.m_row[1] = iiA.m_row[2] = z4;
		if (sB >= 0)
			iiB = inertias[b].m_invInertiaWorld;
		else
			iiB.m_row[0] = iiB.m_row[1] = iiB.m_row[2] = z4;
		F4 n = neg4(c.m_linear);
		recN[p] = f4(c.m_linear.x, c.m_linear.y, c.m_linear.z, imA);
		float pk[48];
		for (int ic = 0; ic < 4; ic++)
		{
			F4 r0 = sub4(c.m_worldPos[ic], posA);
			F4 r1 = sub4(c.m_worldPos[ic], posB);
			F4 ang0 = neg4(cross3(r0, n));
			F4 ang1 = cross3(r1, n);
			F4 q0 = mtMul1(iiA, ang0), q1 = mtMul1(iiB, ang1);
			const float row[12] = {ang0.x, ang0.y, ang0.z, ang1.x, ang1.y, ang1.z, q0.x, q0.y, q0.z, q1.x, q1.y, q1.z};
			for (int j = 0; j < 12; j++) pk[12 * ic + j] = row[j];
		}
		for (int q = 0; q < 12; q++)
			recN[(size_t)(1 + q) * stride + p] = f4(pk[4 * q], pk[4 * q + 1], pk[4 * q + 2], pk[4 * q + 3]);
		recN[(size_t)13 * stride + p] = f4(c.m_jacCoeffInv[0], c.m_jacCoeffInv[1], c.m_jacCoeffInv[2], c.m_jacCoeffInv[3]);
		recN[(size_t)14 * stride + p] = f4(c.m_b[0], c.m_b[1], c.m_b[2], c.m_b[3]);
		recN[(size_t)15 * stride + p] = meta;
		F4 t[2];
		plane_space(n, &t[0], &t[1]);
		F4 r0 = sub4(c.m_center, posA), r1 = sub4(c.m_center, posB);
		float fk[36];
		for (int k = 0; k < 2; k++)
		{
			F4 ang0 = neg4(cross3(r0, t[k]));
			F4 ang1 = cross3(r1, t[k]);
			F4 q0 = mtMul1(iiA, ang0), q1 = mtMul1(iiB, ang1);
			const float dir[15] = {-t[k].x, -t[k].y, -t[k].z, ang0.x, ang0.y, ang0.z, ang1.x, ang1.y,
								   ang1.z,  q0.x,    q0.y,    q0.z,   q1.x,   q1.y,   q1.z};
			for (int j = 0; j < 15; j++) fk[15 * k + j] = dir[j];
		}
		fk[30] = n.x;
		fk[31] = n.y;
		fk[32] = n.z;
		fk[33] = n.w; /* the twist damping subtracts the whole float4 (w included) */
		fk[34] = fk[35] = 0.f;
		for (int q = 0; q < 9; q++)
			recF[(size_t)q * stride + p] = f4(fk[4 * q], fk[4 * q + 1], fk[4 * q + 2], fk[4 * q + 3]);
		F4 ab = nrm3(sub4(posB, posA));
		F4 ac = nrm3(sub4(c.m_center, posA));
		bool damp = dot3(ab, ac) > 0.95f || (imA == 0.f || imB == 0.f);
		recF[(size_t)9 * stride + p] = f4(c.m_fJacCoeffInv[0], c.m_fJacCoeffInv[1], damp ? 1.f : 0.f, imA);
		recF[(size_t)10 * stride + p] = meta;
		lamN[p] = f4(c.m_appliedRambdaDt[0], c.m_appliedRambdaDt[1], c.m_appliedRambdaDt[2], c.m_appliedRambdaDt[3]);
		lamF[p] = f4(c.m_fAppliedRambdaDt[0], c.m_fAppliedRambdaDt[1], 0.f, 0.f);
	}
}

/* the impulses of the level-walked constraints back into the constraints */
__global__ __launch_bounds__(256) void k_lambda_writeback(const int* __restrict__ nPtr, const int* __restrict__ posOf,
														  const F4* __restrict__ lamN, const F4* __restrict__ lamF,
														  b3hContactConstraint4* __restrict__ cs)
{
	int n = *nPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const int p = posOf[i];
		if (p < 0) continue;
		const F4 ln = lamN[p], lf = lamF[p];
		cs[i].m_appliedRambdaDt[0] = ln.x;
		cs[i].m_appliedRambdaDt[1] = ln.y;
		cs[i].m_appliedRambdaDt[2] = ln.z;
		cs[i].m_appliedRambdaDt[3] = ln.w;
		cs[i].m_fAppliedRambdaDt[0] = lf.x;
		cs[i].m_fAppliedRambdaDt[1] = lf.y;
	}
}

/* LDS bytes the level sweep needs for a part of nbp bodies and L levels:
 * velocities (nbp + 2 slots, linear and angular) and the level starts */
__host__ __device__ __forceinline__ size_t pgs_part_lds(int nbp, int L)
{
	return (size_t)(nbp + 2) * 32 + (size_t)(L + 2) * 4;
}

/* ---------------- per-cell dependency levels and parts ----------------
 * Gauss-Seidel results depend only on the order in which each body's updates
 * happen.  Walking the cell in the reference order (rows of up to 64 aligned to
 * batch starts; a row shares no dynamic body), a constraint's level is one more
 * than the last level that touched either of its bodies.  Constraints of one
 * level share no dynamic body, and every body sees its updates in the reference
 * order, so solving level by level reproduces the sequential sweep bit for
 * bit.
 *
 * Parts: cell indices wrap with the split counts, so one cell is several
 * copies in space (tiles B3H_N_SPLIT_X x Y x Z cells apart).  When no dynamic
 * body has constraints in two tiles of the cell (checked here), the tiles
 * share no body and their constraints can be solved by different work-groups
 * at once without changing any body's update order.  The tiles are packed into
 * at most B3H_PGS_PARTS parts, largest first onto the least loaded part.
 * Each part's bodies take part-local velocity slots (cslot is rewritten to
 * them), so a part needs LDS for its own bodies only.
 *
 * Output per cell: the constraints part by part, each part grouped by level
 * (levelOrder, region [off, off+n)); per part its level starts relative to the
 * part's first position, empty levels dropped (levelStart, region starting at
 * off + PARTS*cell); partMeta[2 (cell PARTS + p)] = (first position, count,
 * offset of the level starts, levels) and partMeta[... + 1] = (first entry,
 * count) of the part's bodies in partBodies, which lists (slot in the cell's
 * body list, body) part by part in the cell's region 2 cellOffsets[cell]; per
 * constraint its position in levelOrder (posOf; -1 on the global walk); L
 * (cellLevels; -1 when the cell is too large for LDS: it then takes the global
 * walk, as one part). */
__device__ __forceinline__ int floor_div_pow2(int x, int d) { return x >= 0 ? x / d : -((-x + d - 1) / d); }
__device__ __forceinline__ unsigned tile_key(F4 p, float scale)
{
	int x = (int)((p.x - ((p.x < 0.f) ? 1.f : 0.f)) * scale);
	int y = (int)((p.y - ((p.y < 0.f) ? 1.f : 0.f)) * scale);
	int z = (int)((p.z - ((p.z < 0.f) ? 1.f : 0.f)) * scale);
	unsigned tx = (unsigned)floor_div_pow2(x, B3H_N_SPLIT_X) & 1023u;
	unsigned ty = (unsigned)floor_div_pow2(y, B3H_N_SPLIT_Y) & 1023u;
	unsigned tz = (unsigned)floor_div_pow2(z, B3H_N_SPLIT_Z) & 1023u;
	return tx | (ty << 10) | (tz << 20);
}
/* tile hash table (open addressing) and the sort of the occupied slots */
#define PGS_TILE_SLOTS 2048
#define PGS_TILE_LDS (PGS_TILE_SLOTS * (4 + 4 + 4 + 8) + 256)
/* lev[k]: level in the low 20 bits, tile slot and then part above */
#define LEV_MASK 0xfffff
#define LEV_SHIFT 20
__global__ __launch_bounds__(256) void k_cell_levels(const int* __restrict__ cellCounts,
													 const int* __restrict__ cellOffsets,
													 const int* __restrict__ batchIdx,
													 const int* __restrict__ cellBodyCount, int* __restrict__ cslot,
													 const b3hContactConstraint4* __restrict__ cs,
													 const b3hRigidBodyData* __restrict__ bodies, float scale,
													 int bodyCap, int ldsInts, int pgsLdsBytes, int* __restrict__ levelOrder,
													 int* __restrict__ levelStart, int* __restrict__ cellLevels,
													 const int* __restrict__ cellBodies, b3hInt4* __restrict__ partMeta,
													 int* __restrict__ partBodies, int* __restrict__ posOf, int maxParts)
{
	extern __shared__ int sm[];
	__shared__ int sL, sP, sNumTiles, sOverflow, sFit;
	__shared__ unsigned tKey[PGS_TILE_SLOTS];
	__shared__ int tCount[PGS_TILE_SLOTS], tPart[PGS_TILE_SLOTS];
	__shared__ unsigned long long tSort[PGS_TILE_SLOTS];
	__shared__ int pCount[B3H_PGS_PARTS], pbFirst[B3H_PGS_PARTS], pbCursor[B3H_PGS_PARTS];
	int cell = blockIdx.x;
	int n = cellCounts[cell];
	int off = cellOffsets[cell];
	int nb = cellBodyCount[cell];
	int t = threadIdx.x;
	b3hInt4* meta = partMeta + 2 * cell * B3H_PGS_PARTS; /* meta[2q], meta[2q + 1] */
	if (t < 2 * B3H_PGS_PARTS) meta[t] = b3hInt4{0, 0, 0, 0};
	if (n == 0)
	{
		if (t == 0) cellLevels[cell] = 0;
		return;
	}
	int* lsOut = levelStart + off + B3H_PGS_PARTS * cell;
	/* last[nb] | lev[n] | cnt[P (L + 1)]; a cell whose arrays do not fit
	 * takes the global walk */
	if (nb + n + 1 > ldsInts)
	{
		if (t == 0)
		{
			cellLevels[cell] = -1;
			meta[0] = b3hInt4{0, n, 0, -1};
			meta[1] = b3hInt4{0, nb, 0, 0};
		}
		for (int k = t; k < n; k += blockDim.x) posOf[off + k] = -1;
		return;
	}
	int* last = sm;
	int* lev = sm + nb;
	int* cnt = lev + n;
	for (int k = t; k < nb; k += blockDim.x) last[k] = 0;
	for (int q = t; q < PGS_TILE_SLOTS; q += blockDim.x)
	{
		tKey[q] = 0xffffffffu;
		tCount[q] = 0;
		tPart[q] = 0;
	}
	if (t == 0)
	{
		sNumTiles = 0;
		sOverflow = 0;
		sFit = 1;
	}
	__syncthreads();
	if (t < 64)
	{
		int end = off + n;
		int p = off;
		int L = 0;
		while (p < end)
		{
			int idx = p + t;
			int b0 = batchIdx[p];
			bool active = idx < end && batchIdx[idx] == b0;
			int c = __popcll(__ballot(active));
			int lv = 0;
			int sA = -1, sB = -1;
			if (active)
			{
				sA = cslot[2 * (size_t)idx];
				sB = cslot[2 * (size_t)idx + 1];
				int la = sA >= 0 ? last[sA] : 0;
				int lb = sB >= 0 ? last[sB] : 0;
				lv = la > lb ? la : lb;
				lev[idx - off] = lv;
			}
			wave_sync_solver();
			if (active)
			{
				if (sA >= 0) last[sA] = lv + 1;
				if (sB >= 0) last[sB] = lv + 1;
				L = L > lv + 1 ? L : lv + 1;
			}
			wave_sync_solver();
			p += c;
		}
		for (int m = 1; m < 64; m <<= 1)
		{
			int o = __shfl_xor(L, m);
			L = o > L ? o : L;
		}
		if (t == 0) sL = L;
	}
	__syncthreads();
	int L = sL;
	if (nb + n + L + 1 > ldsInts)
	{
		if (t == 0)
		{
			cellLevels[cell] = -1;
			meta[0] = b3hInt4{0, n, 0, -1};
			meta[1] = b3hInt4{0, nb, 0, 0};
		}
		for (int k = t; k < n; k += blockDim.x) posOf[off + k] = -1;
		return;
	}
	/* tiles: key of each constraint's first dynamic body, counted in a small
	 * open-addressing table; an overflowing table leaves the cell in one part */
	bool split = nb + n + B3H_PGS_PARTS * (L + 1) <= ldsInts;
	for (int k = t; k < nb; k += blockDim.x) last[k] = -1; /* now: the part of each body */
	for (int k = t; k < n && split; k += blockDim.x)
	{
		int sA = cslot[2 * (size_t)(off + k)];
		const b3hContactConstraint4& c = cs[off + k];
		unsigned body = sA >= 0 ? c.m_bodyA : c.m_bodyB;
		unsigned key = tile_key(bodies[body].m_pos, scale);
		int slot = (int)((key * 2654435761u) >> 21); /* 11 bits */
		int probes = 0;
		for (;; slot = (slot + 1) & (PGS_TILE_SLOTS - 1))
		{
			unsigned old = atomicCAS(&tKey[slot], 0xffffffffu, key);
			if (old == 0xffffffffu || old == key) break;
			if (++probes == PGS_TILE_SLOTS)
			{
				slot = -1;
				break;
			}
		}
		if (slot < 0)
			sOverflow = 1;
		else
		{
			if (atomicAdd(&tCount[slot], 1) == 0) atomicAdd(&sNumTiles, 1);
			lev[k] |= slot << LEV_SHIFT;
		}
	}
	__syncthreads();
	/* the occupied slots in assignment order: largest tile first, ties by the
	 * smaller key (count < 2^20 and key < 2^30 pack with the slot) */
	const bool tiled = split && !sOverflow && maxParts > 1 && sNumTiles > 1;
	int nT = 0, np2T = 1;
	if (tiled)
	{
		__shared__ int sCur;
		if (t == 0) sCur = 0;
		__syncthreads();
		for (int q = t; q < PGS_TILE_SLOTS; q += blockDim.x)
			if (tCount[q] > 0)
			{
				int i = atomicAdd(&sCur, 1);
				tSort[i] = ((unsigned long long)(0xfffffu - (unsigned)tCount[q]) << 41) |
						   ((unsigned long long)tKey[q] << 11) | (unsigned long long)q;
			}
		__syncthreads();
		nT = sCur;
		while (np2T < nT) np2T <<= 1;
		for (int i = nT + t; i < np2T; i += blockDim.x) tSort[i] = ~0ull;
		__syncthreads();
		for (int kk = 2; kk <= np2T; kk <<= 1)
			for (int j = kk >> 1; j > 0; j >>= 1)
			{
				for (int i = t; i < np2T; i += blockDim.x)
				{
					int ixj = i ^ j;
					if (ixj > i)
					{
						unsigned long long a = tSort[i], b = tSort[ixj];
						if ((a > b) == ((i & kk) == 0))
						{
							tSort[i] = b;
							tSort[ixj] = a;
						}
					}
				}
				__syncthreads();
			}
	}
	if (t == 0)
	{
		int P = 1;
		if (tiled)
		{
			/* each tile onto the least loaded part (ties: the lower part) */
			int load[B3H_PGS_PARTS];
#pragma unroll
			for (int q = 0; q < B3H_PGS_PARTS; q++) load[q] = 0;
			for (int i = 0; i < nT; i++)
			{
				int q = (int)(tSort[i] & (PGS_TILE_SLOTS - 1));
				int pm = 0, best = load[0];
#pragma unroll
				for (int r = 1; r < B3H_PGS_PARTS; r++)
					if (r < maxParts && load[r] < best)
					{
						best = load[r];
						pm = r;
					}
				load[pm] += tCount[q];
				tPart[q] = pm;
				P = P > pm + 1 ? P : pm + 1;
			}
		}
		sP = P;
	}
	__syncthreads();
	int P = sP;
	/* a body with constraints in two parts: solve the cell as one part */
	__shared__ int conflict;
	if (t == 0) conflict = 0;
	__syncthreads();
	for (int k = t; k < n && P > 1; k += blockDim.x)
	{
		int part = tPart[(unsigned)lev[k] >> LEV_SHIFT];
		int sA = cslot[2 * (size_t)(off + k)], sB = cslot[2 * (size_t)(off + k) + 1];
		if (sA >= 0)
		{
			int old = atomicCAS(&last[sA], -1, part);
			if (old != -1 && old != part) conflict = 1;
		}
		if (sB >= 0)
		{
			int old = atomicCAS(&last[sB], -1, part);
			if (old != -1 && old != part) conflict = 1;
		}
	}
	__syncthreads();
	if (conflict) P = 1;
	int stride = L + 1; /* counters of part p: cnt[p * stride + level] */
	for (int k = t; k < P * stride; k += blockDim.x) cnt[k] = 0;
	__syncthreads();
	for (int k = t; k < n; k += blockDim.x)
	{
		int part = P > 1 ? tPart[(unsigned)lev[k] >> LEV_SHIFT] : 0;
		int lv = lev[k] & LEV_MASK;
		lev[k] = lv | (part << LEV_SHIFT);
		atomicAdd(&cnt[part * stride + lv], 1);
	}
	if (t < B3H_PGS_PARTS) pCount[t] = 0;
	__syncthreads();
	for (int k = t; k < nb; k += blockDim.x) atomicAdd(&pCount[P > 1 ? last[k] : 0], 1);
	__syncthreads();
	if (t == 0)
	{
		/* every part's velocities and level starts must fit the sweep's LDS;
		 * otherwise the cell takes the global walk */
		for (int q = 0; q < P; q++)
		{
			int Lp = 0;
			for (int l = 0; l < L; l++) Lp += cnt[q * stride + l] > 0;
			if (pCount[q] > bodyCap || pgs_part_lds(pCount[q], Lp) > (size_t)pgsLdsBytes) sFit = 0;
		}
		if (!sFit)
		{
			cellLevels[cell] = -1;
			meta[0] = b3hInt4{0, n, 0, -1};
			meta[1] = b3hInt4{0, nb, 0, 0};
		}
	}
	__syncthreads();
	if (!sFit)
	{
		for (int k = t; k < n; k += blockDim.x) posOf[off + k] = -1;
		return;
	}
	if (t == 0)
	{
		int runB = 0;
		for (int q = 0; q < B3H_PGS_PARTS; q++)
		{
			pbFirst[q] = runB;
			pbCursor[q] = runB;
			meta[2 * q + 1] = b3hInt4{runB, pCount[q], 0, 0};
			runB += pCount[q];
		}
		int run = 0, lsOfs = 0;
		for (int q = 0; q < B3H_PGS_PARTS; q++)
		{
			if (q >= P)
			{
				meta[2 * q] = b3hInt4{run, 0, lsOfs, 0};
				continue;
			}
			int first = run, Lp = 0;
			for (int l = 0; l < L; l++)
			{
				int c = cnt[q * stride + l];
				cnt[q * stride + l] = run;
				if (c == 0) continue;
				lsOut[lsOfs + Lp] = run - first;
				Lp++;
				run += c;
			}
			lsOut[lsOfs + Lp] = run - first;
			meta[2 * q] = b3hInt4{first, run - first, lsOfs, Lp};
			lsOfs += Lp + 1;
		}
		cellLevels[cell] = L;
	}
	__syncthreads();
	for (int k = t; k < n; k += blockDim.x)
	{
		int part = (unsigned)lev[k] >> LEV_SHIFT, lv = lev[k] & LEV_MASK;
		int pos = atomicAdd(&cnt[part * stride + lv], 1);
		levelOrder[off + pos] = off + k;
		posOf[off + k] = off + pos;
	}
	/* each body's part-local slot: its place in the part's body list */
	for (int k = t; k < nb; k += blockDim.x)
	{
		int part = P > 1 ? last[k] : 0;
		int pos = atomicAdd(&pbCursor[part], 1);
		int* e = partBodies + 2 * (2 * (size_t)off + pos);
		e[0] = pos - pbFirst[part];
		e[1] = cellBodies[2 * (size_t)off + k];
		last[k] = pos - pbFirst[part];
	}
	__syncthreads();
	for (int k = t; k < n; k += blockDim.x)
	{
		int* cp = cslot + 2 * (size_t)(off + k);
		int sA = cp[0], sB = cp[1];
		cp[0] = sA >= 0 ? last[sA] : -1;
		cp[1] = sB >= 0 ? last[sB] : -1;
	}
}

/* ---------------- PGS sweep, level by level, velocities in LDS ----------------
 * Each thread owns positions t, t+T, ... of the cell's level-ordered list.  The
 * record of a thread's next item is fetched from global memory while its
 * current item is solved and is parked in the thread's LDS slot, so the only
 * wait on global memory comes after the arithmetic.  Level boundaries use an
 * LDS-only barrier (velocities are the only shared state), which does not wait
 * for outstanding global loads. */
__device__ __forceinline__ void lds_barrier()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__device__ __forceinline__ float rel_vel(F4 l0, F4 a0, F4 a1, const Vel& va, const Vel& vb)
{
	F4 l = f4(l0.x, l0.y, l0.z, 0.f);
	return dot3(l, va.lv) + dot3(a0, va.av) + dot3(neg4(l), vb.lv) + dot3(a1, vb.av);
}
/* apply_impulse with P0 = imA*lin, P1 = imB*(-lin) formed as the reference forms them */
__device__ __forceinline__ void apply_pre(F4 lin, float imA, float imB, F4 Q0, F4 Q1, float rl, Vel& a, Vel& b)
{
	F4 linear = f4(lin.x, lin.y, lin.z, 0.f);
	F4 P0 = smul4(imA, linear);
	F4 P1 = smul4(imB, neg4(linear));
	a.lv = add_scaled3(a.lv, P0, rl);
	a.av = add_scaled3(a.av, Q0, rl);
	b.lv = add_scaled3(b.lv, P1, rl);
	b.av = add_scaled3(b.av, Q1, rl);
}

template <int PHASE>
struct Slot
{
	static constexpr int W = PHASE == 1 ? PGS_FSLOT : PGS_NSLOT;
	F4 r[W];
};
/* records and impulses of the constraint at level position p: field k of
 * every record is contiguous over positions (stride apart), so the lanes of a
 * wavefront, which take consecutive positions, read contiguous memory */
struct RecSrc
{
	const F4* recN;
	const F4* recF;
	F4* lamN; /* normal impulses per position */
	F4* lamF; /* friction impulses per position (x, y) */
	int stride;
};
template <int PHASE>
__device__ __forceinline__ void slot_fetch(Slot<PHASE>& s, int p, const RecSrc& R)
{
	if (PHASE == 1)
	{
#pragma unroll
		for (int k = 0; k < PGS_FREC; k++) s.r[k] = R.recF[(size_t)k * R.stride + p];
		s.r[PGS_FREC] = R.lamN[p];
		s.r[PGS_FREC + 1] = R.lamF[p];
	}
	else
	{
#pragma unroll
		for (int k = 0; k < PGS_NREC; k++) s.r[k] = R.recN[(size_t)k * R.stride + p];
		s.r[PGS_NREC] = R.lamN[p];
	}
}

/* float j of a packed record (j is a constant after unrolling) */
__device__ __forceinline__ float rc(const F4* r, int j)
{
	const F4 v = r[j >> 2];
	switch (j & 3)
	{
	case 0: return v.x;
	case 1: return v.y;
	case 2: return v.z;
	default: return v.w;
	}
}
__device__ __forceinline__ F4 rc3(const F4* r, int j) { return f4(rc(r, j), rc(r, j + 1), rc(r, j + 2), 0.f); }
/* TRACK: keep the largest impulse change and impulse (residual tracking) */
template <int PHASE, bool TRACK>
__device__ __forceinline__ void pgs_solve(const F4* __restrict__ Sr, F4* sLv, F4* sAv, int nb, const RecSrc& R,
										  float friction, float& maxD, float& maxL)
{
	struct
	{
		const F4* r;
	} S = {Sr};
	const F4 meta = S.r[PHASE == 1 ? PGS_FREC - 1 : PGS_NREC - 1];
	float imB = meta.x;
	int sA = fbits(meta.y), sB = fbits(meta.z), pos = fbits(meta.w);
	Vel va, vb;
	/* static bodies (slot -1) read the zero slot nb */
	const int rA = sA >= 0 ? sA : nb, rB = sB >= 0 ? sB : nb;
	va.lv = sLv[rA];
	va.av = sAv[rA];
	vb.lv = sLv[rB];
	vb.av = sAv[rB];
	if (PHASE == 1)
	{
		F4 fj = S.r[9];
		if (fj.x == 0) return; /* solveFriction.cl: rows without friction are skipped */
		float imA = fj.w;
		F4 lamN = S.r[PGS_FREC];
		float sum = 0;
		sum += lamN.x;
		sum += lamN.y;
		sum += lamN.z;
		sum += lamN.w;
		float maxR = friction * sum;
		float minR = -maxR;
		float fl[2] = {S.r[PGS_FREC + 1].x, S.r[PGS_FREC + 1].y};
		float fjac[2] = {fj.x, fj.y};
#pragma unroll
		for (int k = 0; k < 2; k++)
		{
			F4 lin = rc3(S.r, 15 * k), a0 = rc3(S.r, 15 * k + 3), a1 = rc3(S.r, 15 * k + 6);
			float rl = rel_vel(lin, a0, a1, va, vb);
			rl *= fjac[k];
			float prev = fl[k];
			float upd = prev;
			upd += rl;
			upd = cl_max(upd, minR);
			upd = cl_min(upd, maxR);
			rl = upd - prev;
			fl[k] = upd;
			if (TRACK)
			{
				if (TRACK)
				{
					maxD = fmaxf(maxD, fabsf(rl));
					maxL = fmaxf(maxL, fabsf(upd));
				}
			}
			apply_pre(lin, imA, imB, rc3(S.r, 15 * k + 9), rc3(S.r, 15 * k + 12), rl, va, vb);
		}
		if (fj.z != 0.f)
		{
			F4 n = f4(rc(S.r, 30), rc(S.r, 31), rc(S.r, 32), rc(S.r, 33));
			float angNA = dot3(n, va.av);
			float angNB = dot3(n, vb.av);
			va.av = sub4(va.av, smul4(angNA * 0.1f, n));
			vb.av = sub4(vb.av, smul4(angNB * 0.1f, n));
		}
		R.lamF[pos] = f4(fl[0], fl[1], 0.f, 0.f);
	}
	else
	{
		F4 lin = S.r[0];
		float imA = lin.w;
		F4 lamv = S.r[PGS_NREC];
		float lam[4] = {lamv.x, lamv.y, lamv.z, lamv.w};
		F4 jv = S.r[13], bv = S.r[14];
		float jac[4] = {jv.x, jv.y, jv.z, jv.w};
		float bb[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
		for (int ic = 0; ic < 4; ic++)
		{
			if (jac[ic] != 0.f)
			{
				F4 a0 = rc3(S.r, 4 + 12 * ic), a1 = rc3(S.r, 7 + 12 * ic);
				float rl = rel_vel(lin, a0, a1, va, vb) + bb[ic];
				rl *= jac[ic];
				float prev = lam[ic];
				float upd = prev;
				upd += rl;
				upd = cl_max(upd, 0.f);
				upd = cl_min(upd, FLT_MAX);
				rl = upd - prev;
				lam[ic] = upd;
				if (TRACK)
				{
					maxD = fmaxf(maxD, fabsf(rl));
					maxL = fmaxf(maxL, fabsf(upd));
				}
				apply_pre(lin, imA, imB, rc3(S.r, 10 + 12 * ic), rc3(S.r, 13 + 12 * ic), rl, va, vb);
			}
		}
		R.lamN[pos] = f4(lam[0], lam[1], lam[2], lam[3]);
	}
	if (sA >= 0)
	{
		sLv[sA] = va.lv;
		sAv[sA] = va.av;
	}
	if (sB >= 0)
	{
		sLv[sB] = vb.lv;
		sAv[sB] = vb.av;
	}
}

/* warm start, in level order, reading the records directly: PART 0 sets and
 * applies the normal impulses before the first normal sweep, PART 1 the
 * friction impulses (clamped by the normal impulses the normal sweeps left)
 * before the first friction sweep.  Each phase thus starts from its own guess
 * and keeps the cold start's fixed point (all normal sweeps run before the
 * friction sweeps). */
template <int PART>
__device__ void pgs_warm(int p, const RecSrc& R, const int* __restrict__ levelOrder, F4* sLv, F4* sAv, int nb,
						 float friction, const float* __restrict__ warm)
{
	const size_t st = (size_t)R.stride;
	const int idx = levelOrder[p]; /* the constraint, for its initial impulses */
	F4 N[PGS_NREC], Fr[PGS_FREC];
	for (int k = 0; k < PGS_NREC; k++) N[k] = R.recN[k * st + p];
	for (int k = 0; k < PGS_FREC; k++) Fr[k] = R.recF[k * st + p];
	F4 meta = N[PGS_NREC - 1];
	float imB = meta.x;
	int sA = fbits(meta.y), sB = fbits(meta.z);
	Vel va, vb;
	/* static bodies (slot -1) read the zero slot nb */
	const int rA = sA >= 0 ? sA : nb, rB = sB >= 0 ? sB : nb;
	va.lv = sLv[rA];
	va.av = sAv[rA];
	vb.lv = sLv[rB];
	vb.av = sAv[rB];
	F4 lin = N[0];
	float imA = lin.w;
	const F4 ln = R.lamN[p];
	float lam[4] = {ln.x, ln.y, ln.z, ln.w};
	if (PART == 0)
	{
		F4 jv = N[13];
		float jac[4] = {jv.x, jv.y, jv.z, jv.w};
		for (int ic = 0; ic < 4; ic++)
		{
			if (jac[ic] == 0.f) continue;
			float rl = cl_max(warm[6 * (size_t)idx + ic], 0.f);
			lam[ic] = rl;
			apply_pre(lin, imA, imB, rc3(N, 10 + 12 * ic), rc3(N, 13 + 12 * ic), rl, va, vb);
		}
		R.lamN[p] = f4(lam[0], lam[1], lam[2], lam[3]);
	}
	F4 fj = Fr[9];
	if (PART == 1 && !(fj.x == 0))
	{
		float sum = 0;
		for (int j = 0; j < 4; j++) sum += lam[j];
		float maxR = friction * sum;
		float fl[2];
		for (int k = 0; k < 2; k++)
		{
			float l = cl_min(cl_max(warm[6 * (size_t)idx + 4 + k], -maxR), maxR);
			fl[k] = l;
			apply_pre(rc3(Fr, 15 * k), imA, imB, rc3(Fr, 15 * k + 9), rc3(Fr, 15 * k + 12), l, va, vb);
		}
		R.lamF[p] = f4(fl[0], fl[1], 0.f, 0.f);
	}
	if (sA >= 0)
	{
		sLv[sA] = va.lv;
		sAv[sA] = va.av;
	}
	if (sB >= 0)
	{
		sLv[sB] = vb.lv;
		sAv[sB] = vb.av;
	}
}

/* one sweep of a part by dependency levels (this thread: positions ls[c] + t,
 * + T, ... of each level c, relative to the part's first position base) */
template <int PHASE, bool TRACK>
__device__ __forceinline__ void pgs_walk(int t, int n, int L, const int* ls, int base, const RecSrc& R, F4* sLv,
										 F4* sAv, int nb, float friction, float& maxD, float& maxL)
{
	const int T = PGS_THREADS;
	/* Three record slots in fixed registers, one per level residue: level
	 * c's first item of this thread (position ls[c] + t) is fetched into
	 * slot c % 3 while level c - 3 is solved, so every record is in flight
	 * for two levels.  The fetch is unconditional (the part's first position
	 * when the thread has no item there), which lets the compiler wait for
	 * exactly the slot it solves.  Further items of a level wider than the
	 * block are fetched on the spot. */
	Slot<PHASE> A, B, C;
	bool hA = false, hB = false, hC = false;
	{
		int p0 = ls[0] + t, p1 = 1 < L ? ls[1] + t : n, p2 = 2 < L ? ls[2] + t : n;
		hA = 0 < L && p0 < ls[1];
		hB = 1 < L && p1 < ls[2];
		hC = 2 < L && p2 < ls[3];
		slot_fetch<PHASE>(A, base + (hA ? p0 : 0), R);
		slot_fetch<PHASE>(B, base + (hB ? p1 : 0), R);
		slot_fetch<PHASE>(C, base + (hC ? p2 : 0), R);
	}
#define PGS_LEVEL(X, hX, c)                                                                   \
	{                                                                                         \
		bool live = (c) < L;                                                                  \
		if (hX) pgs_solve<PHASE, TRACK>(X.r, sLv, sAv, nb, R, friction, maxD, maxL);         \
		for (int q = live ? ls[c] + t + T : n; q < (live ? ls[(c) + 1] : 0); q += T)          \
		{                                                                                     \
			Slot<PHASE> W;                                                                    \
			slot_fetch<PHASE>(W, base + q, R);                                                \
			pgs_solve<PHASE, TRACK>(W.r, sLv, sAv, nb, R, friction, maxD, maxL);             \
		}                                                                                     \
		int pn = (c) + 3 < L ? ls[(c) + 3] + t : n;                                          \
		hX = (c) + 3 < L && pn < ls[(c) + 4];                                                 \
		slot_fetch<PHASE>(X, base + (hX ? pn : 0), R);                                        \
		lds_barrier();                                                                        \
	}
	/* every pass runs all three slots (refills and barriers unconditional,
	 * 3 ceil(L / 3) barriers in all) so the loads in flight are the same on
	 * every path */
	for (int c = 0; c < L; c += 3)
	{
		PGS_LEVEL(A, hA, c)
		PGS_LEVEL(B, hB, c + 1)
		PGS_LEVEL(C, hC, c + 2)
	}
#undef PGS_LEVEL
}

/* global -> LDS copy of n ints, eight per thread per round with every load
 * issued before any store (out-of-range lanes load element 0 and store into
 * the padding entry dst[n]) */
__device__ __forceinline__ void stage_ints(int* dst, const int* __restrict__ src, int n, int t)
{
	for (int k0 = t; k0 < n; k0 += 8 * PGS_THREADS)
	{
		int v[8];
#pragma unroll
		for (int u = 0; u < 8; u++)
		{
			int k = k0 + u * PGS_THREADS;
			v[u] = src[k < n ? k : 0];
		}
#pragma unroll
		for (int u = 0; u < 8; u++)
		{
			int k = k0 + u * PGS_THREADS;
			dst[k < n ? k : n] = v[u];
		}
	}
}

template <int PHASE>
__global__ __launch_bounds__(PGS_THREADS) void k_pgs_level(b3hRigidBodyData* __restrict__ bodies,
														   const b3hInertiaData* __restrict__ inertias,
														   b3hContactConstraint4* __restrict__ cs,
														   const int* __restrict__ cellCounts,
														   const int* __restrict__ cellOffsets,
														   const int* __restrict__ batchSizes,
														   const int* __restrict__ batchIdx,
														   const int* __restrict__ cellBodies,
														   const int* __restrict__ cellBodyCount,
														   const int* __restrict__ cellLevels,
														   const int* __restrict__ levelStart,
														   const int* __restrict__ levelOrder,
														   const b3hInt4* __restrict__ partMeta,
														   const int* __restrict__ partBodies, RecSrc R, int ib,
														   float friction,
														   const float* __restrict__ warm, float* __restrict__ resid,
														   const int* __restrict__ lastClass, int fuseIntegrate,
														   b3hIntegrateParams ip, int ldsBytes)
{
	extern __shared__ F4 smem4[];
	/* blocks [part * cells of the class, ...): part-major, so part 0 of every
	 * cell of the class comes first */
	constexpr int CELLS_PER_CLASS = B3H_N_CELLS / B3H_N_CELL_BATCHES;
	const int part = blockIdx.x / CELLS_PER_CLASS;
	int cell = cell_of(blockIdx.x % CELLS_PER_CLASS, ib);
	int cellN = cellCounts[cell];
	if (cellN == 0) return;
	int start = cellOffsets[cell];
	int end = start + cellN;
	int t = threadIdx.x;
	const int T = PGS_THREADS;
	int L = cellLevels[cell];
	if (L < 0)
	{
		if (part != 0) return;
		/* cell too large for the LDS cache: the reference's global walk on one wave */
		int maxBatch = batchSizes[cell];
		int idx = start + t;
		for (int cur = 0; cur < maxBatch; cur++)
		{
			if (t < 64)
				while (idx < end && batchIdx[idx] == cur)
				{
					solve_one(cs + idx, bodies, inertias, PHASE, friction, warm ? warm + 6 * (size_t)idx : nullptr,
							  resid);
					idx += 64;
				}
			__syncthreads();
		}
		if (fuseIntegrate)
		{
			const int* list = cellBodies + 2 * (size_t)start;
			int nbg = cellBodyCount[cell];
			for (int q = t; q < nbg; q += T)
			{
				int b = list[q];
				if (lastClass[b] != ib) continue;
				b3hRigidBodyData body = bodies[b];
				b3h_integrate_body(body, ip.dt, ip.damping, ip.gravity);
				bodies[b].m_pos = body.m_pos;
				bodies[b].m_quat = body.m_quat;
				bodies[b].m_linVel = body.m_linVel;
				bodies[b].m_angVel = body.m_angVel;
			}
		}
		return;
	}
	const b3hInt4 pm = partMeta[2 * (cell * B3H_PGS_PARTS + part)];
	const int n = pm.y; /* this part's constraints */
	if (n == 0) return;
	L = pm.w;
	const b3hInt4 pb = partMeta[2 * (cell * B3H_PGS_PARTS + part) + 1];
	const int nbp = pb.y; /* this part's bodies: (slot, body) pairs */
	const int* own = partBodies + 2 * (2 * (size_t)start + pb.x);
	const int base = start + pm.x; /* the part's first position */
	/* velocities in part-local slots (own[2 q] == q); the rows carry their
	 * inverse-inertia products, so no inertia is read here */
	const int nb = nbp;
	(void)ldsBytes;
	F4* sLv = smem4;                 /* nb + 2: slot nb zero (static bodies), nb + 1 padding */
	F4* sAv = sLv + nb + 2;          /* nb + 2 */
	int* sLs = (int*)(sAv + nb + 2); /* L + 2: level starts relative to start */
	/* stage the level starts (positions are the part's first position + q;
	 * records are read by position, so the level order itself is not needed) */
	{
		const int* srcL = levelStart + start + B3H_PGS_PARTS * cell + pm.z;
		for (int k = t; k <= L; k += PGS_THREADS) sLs[k] = srcL[k];
	}
	if (t == 3) sLv[nb] = f4(0.f, 0.f, 0.f, 0.f);
	if (t == 4) sAv[nb] = f4(0.f, 0.f, 0.f, 0.f);
	/* gather this part's velocities, four bodies per thread per round with
	 * every load issued before any is stored (lanes past the end read body 0,
	 * which exists, and store into the padding slot) */
	for (int k0 = t; k0 < nbp; k0 += 4 * T)
	{
		int sv[4], bb[4];
		F4 lv[4], av[4];
#pragma unroll
		for (int u = 0; u < 4; u++)
		{
			int k = k0 + u * T;
			sv[u] = k < nbp ? own[2 * k] : nb + 1;
			bb[u] = k < nbp ? own[2 * k + 1] : 0;
		}
#pragma unroll
		for (int u = 0; u < 4; u++)
		{
			lv[u] = bodies[bb[u]].m_linVel;
			av[u] = bodies[bb[u]].m_angVel;
		}
#pragma unroll
		for (int u = 0; u < 4; u++)
		{
			sLv[sv[u]] = lv[u];
			sAv[sv[u]] = av[u];
		}
	}
	const int* ls = sLs;
	float maxD = 0.f, maxL = 0.f;
	__syncthreads();
	if (PHASE >= 2)
	{
		for (int cl = 0; cl < L; cl++)
		{
			for (int k = ls[cl] + t; k < ls[cl + 1]; k += T)
				pgs_warm<PHASE - 2>(base + k, R, levelOrder, sLv, sAv, nb, friction, warm);
			lds_barrier();
		}
	}
	else
	{
		if (resid)
			pgs_walk<PHASE, true>(t, n, L, ls, base, R, sLv, sAv, nb, friction, maxD, maxL);
		else
			pgs_walk<PHASE, false>(t, n, L, ls, base, R, sLv, sAv, nb, friction, maxD, maxL);
	}
	__syncthreads();
	if (!fuseIntegrate)
		for (int q = t; q < nbp; q += T)
		{
			int slot = own[2 * q], b = own[2 * q + 1];
			bodies[b].m_linVel = sLv[slot];
			bodies[b].m_angVel = sAv[slot];
		}
	else
		/* bodies whose last touching class this is have their final velocity:
		 * integrate them here (integrateTransformsKernel).  Two bodies per
		 * thread per round, loads first; the integration is evaluated for
		 * every body and kept only where it applies */
		for (int q0 = t; q0 < nbp; q0 += 2 * T)
		{
			int qq1 = q0 + T;
			int sq0 = own[2 * q0], bq0 = own[2 * q0 + 1];
			int sq1 = qq1 < nbp ? own[2 * qq1] : nb + 1, bq1 = qq1 < nbp ? own[2 * qq1 + 1] : 0;
			int lc0 = lastClass[bq0], lc1 = lastClass[bq1];
			b3hRigidBodyData bd0 = bodies[bq0], bd1 = bodies[bq1];
#pragma unroll
			for (int u = 0; u < 2; u++)
			{
				int sv = u ? sq1 : sq0, b = u ? bq1 : bq0, lc = u ? lc1 : lc0;
				b3hRigidBodyData body = u ? bd1 : bd0;
				body.m_linVel = sLv[sv];
				body.m_angVel = sAv[sv];
				b3h_integrate_body(body, ip.dt, ip.damping, ip.gravity);
				if (sv >= nb) continue;
				if (lc == ib)
				{
					bodies[b].m_pos = body.m_pos;
					bodies[b].m_quat = body.m_quat;
					bodies[b].m_linVel = body.m_linVel;
					bodies[b].m_angVel = body.m_angVel;
				}
				else
				{
					bodies[b].m_linVel = sLv[sv];
					bodies[b].m_angVel = sAv[sv];
				}
			}
		}
	if (resid)
	{
		if (maxD > 0.f) atomic_max_pos(resid, maxD);
		if (maxL > 0.f) atomic_max_pos(resid + 1, maxL);
	}
}

static int s_ldsCapBytes = -1;
static int s_pgsBodyCap = 1 << 30;
static int s_pgsLdsBytes = 32768;

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
	if (hipFuncSetAttribute((const void*)k_batch_cells, hipFuncAttributeMaxDynamicSharedMemorySize, want) != hipSuccess ||
		hipFuncSetAttribute((const void*)k_cell_bodies, hipFuncAttributeMaxDynamicSharedMemorySize, want) != hipSuccess)
		want = 65536 - 1024;
	if (hipFuncSetAttribute((const void*)k_cell_levels, hipFuncAttributeMaxDynamicSharedMemorySize,
							want - PGS_TILE_LDS) != hipSuccess)
		want = 65536 - 1024;
	int pgsBytes = want;
	if (hipFuncSetAttribute((const void*)k_pgs_level<0>, hipFuncAttributeMaxDynamicSharedMemorySize, pgsBytes) !=
			hipSuccess ||
		hipFuncSetAttribute((const void*)k_pgs_level<1>, hipFuncAttributeMaxDynamicSharedMemorySize, pgsBytes) !=
			hipSuccess ||
		hipFuncSetAttribute((const void*)k_pgs_level<2>, hipFuncAttributeMaxDynamicSharedMemorySize, pgsBytes) !=
			hipSuccess ||
		hipFuncSetAttribute((const void*)k_pgs_level<3>, hipFuncAttributeMaxDynamicSharedMemorySize, pgsBytes) !=
			hipSuccess)
		pgsBytes = 32768;
	s_pgsLdsBytes = pgsBytes;
	s_pgsBodyCap = 1 << 30;
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

void b3h_solve_setup(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias, const int* numContacts,
					 int maxContacts, const b3hContact4Data* staging, const int* contactPair,
					 const b3hSolverParams& prm, int staticIdx, b3hSolverBuffers& sb, bool writeSortedContacts,
					 hipStream_t s)
{
	int ldsCapBytes = s_ldsCapBytes > 0 ? s_ldsCapBytes : 32768;
	b3h_fill_i32(sb.cellCounts, 0, B3H_N_CELLS, s);
	int gb = gblocks(maxContacts);
	int kb = gb < B3H_N_CELLS ? gb : B3H_N_CELLS;
	hipLaunchKernelGGL(k_solver_keys, dim3(kb), dim3(256), 0, s, numContacts, staging, contactPair, bodies, staticIdx,
					   1.f / prm.m_batchCellSize, sb.cellOf, sb.cellCounts);
	hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(B3H_N_CELLS), 0, s, sb.cellCounts, sb.cellOffsets, sb.cellCursor);
	hipLaunchKernelGGL(k_cell_scatter, dim3(kb), dim3(256), 0, s, numContacts, sb.cellOf, sb.cellCursor, sb.cellSlots);
	hipLaunchKernelGGL(k_batch_cells, dim3(B3H_N_CELLS), dim3(256), (size_t)ldsCapBytes, s, sb.cellCounts, sb.cellOffsets,
					   sb.cellSlots, staging, contactPair, staticIdx, ldsCapBytes, sb.sortKeys, sb.perm, sb.batchIdx,
					   sb.batchSizes);
	hipLaunchKernelGGL(k_to_constraints, dim3(gb), dim3(256), 0, s, numContacts, sb.perm, sb.batchIdx, staging,
					   contactPair, bodies, inertias, prm, sb.constraints,
					   writeSortedContacts ? sb.sortedContacts : (b3hContact4Data*)nullptr);
	b3h_fill_i32(sb.lastClass, -1, numBodies, s);
	hipLaunchKernelGGL(k_cell_bodies, dim3(B3H_N_CELLS), dim3(256), (size_t)ldsCapBytes, s, sb.cellCounts,
					   sb.cellOffsets, sb.constraints, bodies, numBodies, ldsCapBytes / 4, (unsigned*)sb.sortKeys,
					   sb.cellBodies,
					   sb.cellBodyCount, sb.cslot, sb.lastClass);
	int cap = (sb.pgsBodyCap >= 0 && sb.pgsBodyCap < s_pgsBodyCap) ? sb.pgsBodyCap : s_pgsBodyCap;
	size_t pgsLds = (size_t)s_pgsLdsBytes;
	const int levelsLds = ldsCapBytes - PGS_TILE_LDS; /* beside the kernel's static tile tables */
	hipLaunchKernelGGL(k_cell_levels, dim3(B3H_N_CELLS), dim3(256), (size_t)levelsLds, s, sb.cellCounts,
					   sb.cellOffsets, sb.batchIdx, sb.cellBodyCount, sb.cslot, (const b3hContactConstraint4*)sb.constraints,
					   bodies, 1.f / prm.m_batchCellSize, cap, levelsLds / 4, (int)pgsLds, sb.levelOrder,
					   sb.levelStart, sb.cellLevels, (const int*)sb.cellBodies, sb.partMeta, sb.partBodies, sb.posOf,
					   sb.pgsMaxParts > 0 && sb.pgsMaxParts < B3H_PGS_PARTS ? sb.pgsMaxParts : B3H_PGS_PARTS);
	hipLaunchKernelGGL(k_pgs_prep, dim3(gb), dim3(256), 0, s, numContacts, sb.constraints, sb.cslot, bodies, inertias,
					   (const int*)sb.posOf, sb.maxContacts, sb.recN, sb.recF, sb.lamN, sb.lamF);
}

/* residual tracking (off the hot path): over the rows one sweep of `phase`
 * updates (0 normal, 1 friction), the sums of (impulse - impulse before the
 * sweep)^2 and impulse^2; refreshes the snapshot (phase < 0: snapshot only) */
__global__ __launch_bounds__(256) void k_resid_l2(const int* __restrict__ nPtr,
												  const b3hContactConstraint4* __restrict__ cs, const int* __restrict__ posOf,
												  const F4* __restrict__ lamN, const F4* __restrict__ lamF,
												  float* __restrict__ snap, int phase, double* __restrict__ sums)
{
	__shared__ double red[2][256];
	int n = *nPtr, t = threadIdx.x;
	double d2 = 0.0, l2 = 0.0;
	for (int k = blockIdx.x * blockDim.x + t; k < n; k += gridDim.x * blockDim.x)
	{
		const b3hContactConstraint4& c = cs[k];
		/* impulses of level-walked constraints live at their positions during
		 * the sweeps; those of global-walk cells in the constraints */
		const int p = posOf[k];
		float ln[4], lf[2];
		if (p >= 0)
		{
			const F4 a = lamN[p], b = lamF[p];
			ln[0] = a.x, ln[1] = a.y, ln[2] = a.z, ln[3] = a.w;
			lf[0] = b.x, lf[1] = b.y;
		}
		else
		{
			for (int i = 0; i < 4; i++) ln[i] = c.m_appliedRambdaDt[i];
			for (int m = 0; m < 2; m++) lf[m] = c.m_fAppliedRambdaDt[m];
		}
		float* sn = snap + 6 * (size_t)k;
		if (phase != 1)
			for (int i = 0; i < 4; i++)
			{
				float l = ln[i];
				if (phase == 0 && c.m_jacCoeffInv[i] != 0.f)
				{
					double d = (double)l - sn[i];
					d2 += d * d;
					l2 += (double)l * l;
				}
				sn[i] = l;
			}
		if (phase != 0)
			for (int m = 0; m < 2; m++)
			{
				float l = lf[m];
				if (phase == 1)
				{
					double d = (double)l - sn[4 + m];
					d2 += d * d;
					l2 += (double)l * l;
				}
				sn[4 + m] = l;
			}
	}
	red[0][t] = d2;
	red[1][t] = l2;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1)
	{
		if (t < o)
		{
			red[0][t] += red[0][t + o];
			red[1][t] += red[1][t + o];
		}
		__syncthreads();
	}
	if (t == 0 && phase >= 0)
	{
		atomicAdd(sums, red[0][0]);
		atomicAdd(sums + 1, red[1][0]);
	}
}

void b3h_solve_sweeps(b3hRigidBodyData* bodies, const b3hInertiaData* inertias, const b3hSolverParams& prm,
					  b3hSolverBuffers& sb, bool trackResidual, hipStream_t s, hipEvent_t evNormalDone,
					  const b3hIntegrateParams* integrate)
{
	int K = prm.m_numIterations;
	size_t pgsLds = (size_t)s_pgsLdsBytes;
	if (trackResidual) b3h_fill_i32((int*)sb.resid, 0, 4 * K, s);
	const dim3 grid(B3H_N_CELLS / B3H_N_CELL_BATCHES * B3H_PGS_PARTS), block(PGS_THREADS);
	b3hIntegrateParams ip = integrate ? *integrate : b3hIntegrateParams{0.f, 0.f, f4(0, 0, 0, 0)};
#define B3H_PGS_ARGS(ib, resid_, warm_, fuse_)                                                                         \
	bodies, inertias, sb.constraints, sb.cellCounts, sb.cellOffsets, sb.batchSizes, sb.batchIdx, sb.cellBodies, \
		sb.cellBodyCount, sb.cellLevels, sb.levelStart, sb.levelOrder, (const b3hInt4*)sb.partMeta,               \
		(const int*)sb.partBodies, RecSrc{(const F4*)sb.recN, (const F4*)sb.recF, (F4*)sb.lamN, (F4*)sb.lamF,      \
										 sb.maxContacts},                                                      \
		ib,                                                                                                    \
		prm.m_frictionCoeff, warm_, resid_, (const int*)sb.lastClass, fuse_, ip, (int)pgsLds
	if (sb.warmImpulses)
		for (int ib = 0; ib < B3H_N_CELL_BATCHES; ib++)
			hipLaunchKernelGGL(k_pgs_level<2>, grid, block, pgsLds, s,
							   B3H_PGS_ARGS(ib, (float*)nullptr, (const float*)sb.warmImpulses, 0));
	bool l2 = trackResidual && sb.residSnap && sb.residL2 && sb.numContacts;
	int rb = gblocks(sb.maxContacts);
	if (rb > 1024) rb = 1024;
	if (l2)
	{
		b3h_fill_i32((int*)sb.residL2, 0, 8 * K, s);
		hipLaunchKernelGGL(k_resid_l2, dim3(rb), dim3(256), 0, s, sb.numContacts, (const b3hContactConstraint4*)sb.constraints,
						   (const int*)sb.posOf, (const F4*)sb.lamN, (const F4*)sb.lamF, sb.residSnap, -1,
						   (double*)nullptr);
	}
	for (int phase = 0; phase < 2; phase++)
	{
		if (phase == 1)
		{
			if (evNormalDone) hipEventRecord(evNormalDone, s);
			if (sb.warmImpulses)
			{
				/* the friction part of the warm start, after the normal sweeps */
				for (int ib = 0; ib < B3H_N_CELL_BATCHES; ib++)
					hipLaunchKernelGGL(k_pgs_level<3>, grid, block, pgsLds, s,
									   B3H_PGS_ARGS(ib, (float*)nullptr, (const float*)sb.warmImpulses, 0));
				if (l2)
					hipLaunchKernelGGL(k_resid_l2, dim3(rb), dim3(256), 0, s, sb.numContacts,
									   (const b3hContactConstraint4*)sb.constraints, (const int*)sb.posOf,
									   (const F4*)sb.lamN, (const F4*)sb.lamF, sb.residSnap, -1, (double*)nullptr);
			}
		}
		for (int it = 0; it < K; it++)
		{
			for (int ib = 0; ib < B3H_N_CELL_BATCHES; ib++)
			{
				float* r = trackResidual ? sb.resid + 2 * (phase * K + it) : (float*)nullptr;
				int fuse = (integrate && phase == 1 && it == K - 1) ? 1 : 0;
				if (phase == 0)
					hipLaunchKernelGGL(k_pgs_level<0>, grid, block, pgsLds, s,
									   B3H_PGS_ARGS(ib, r, (const float*)nullptr, 0));
				else
					hipLaunchKernelGGL(k_pgs_level<1>, grid, block, pgsLds, s,
									   B3H_PGS_ARGS(ib, r, (const float*)nullptr, fuse));
			}
			if (l2)
				hipLaunchKernelGGL(k_resid_l2, dim3(rb), dim3(256), 0, s, sb.numContacts,
								   (const b3hContactConstraint4*)sb.constraints, (const int*)sb.posOf,
								   (const F4*)sb.lamN, (const F4*)sb.lamF, sb.residSnap, phase,
								   sb.residL2 + 2 * (phase * K + it));
		}
	}
	/* the final impulses of level-walked constraints back into the constraints */
	hipLaunchKernelGGL(k_lambda_writeback, dim3(rb), dim3(256), 0, s, sb.numContacts, (const int*)sb.posOf,
					   (const F4*)sb.lamN, (const F4*)sb.lamF, sb.constraints);
#undef B3H_PGS_ARGS
}

void b3h_solve_contacts(int numBodies, b3hRigidBodyData* bodies, const b3hInertiaData* inertias, const int* numContacts,
						int maxContacts, const b3hContact4Data* staging, const int* contactPair,
						const b3hSolverParams& prm, int staticIdx, b3hSolverBuffers& sb, bool writeSortedContacts,
						bool trackResidual, hipStream_t s, hipEvent_t evSetupDone, hipEvent_t evNormalDone,
						const b3hIntegrateParams* integrate)
{
	b3h_solve_setup(numBodies, bodies, inertias, numContacts, maxContacts, staging, contactPair, prm, staticIdx, sb,
					writeSortedContacts, s);
	if (evSetupDone) hipEventRecord(evSetupDone, s);
	b3h_solve_sweeps(bodies, inertias, prm, sb, trackResidual, s, evNormalDone, integrate);
}
