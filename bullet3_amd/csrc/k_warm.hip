/*
 * k_warm.hip -- initial impulses for the PGS solve (warm start).
 *
 * The reference solver starts every step from zero impulses
 * (b3ConvertConstraint4.h:73,107,132 zero m_appliedRambdaDt and
 * m_fAppliedRambdaDt).  Two alternatives change only the initial iterate, so
 * the solve's fixed point is the same:
 *   previous frame: each constraint takes the final impulses of the previous
 *     step's constraint on the same body pair; a normal row takes the impulse
 *     of the nearest previous contact point (within matchTol), friction rows
 *     the previous world-space friction impulse projected on the new tangents;
 *   learned: a network maps 24 per-point features (contact geometry, relative
 *     velocity, bias, mass ratios, the matched previous impulses) to a
 *     correction of those previous-frame impulses, in units of the row's
 *     effective mass times |g| dt.
 * History: after the solve every constraint's final impulses, contact points
 * and friction vector are stored, keyed by (bodyA, bodyB) in an open-addressing
 * hash table that the next step probes.  Keys are unique per step, so lookups
 * do not depend on insertion order.
 *
 * Feature row of contact point i of a constraint (u = |g| dt, s_i = |m_i| u
 * with m_i the row's effective mass, sf_k the same for friction row k, n the
 * normal, t1/t2 the friction tangents, dv the velocity of A's point relative to
 * B's, np the number of points, lp/fp the matched previous impulses):
 *   0-2 n | 3 dv.n/u | 4 dv.t1/u | 5 dv.t2/u | 6 b_i/u | 7 depth_i/0.05 |
 *   8 imA/(imA+imB) | 9 imB/(imA+imB) | 10 |m_i|(imA+imB) | 11-12 |mf_k|(imA+imB) |
 *   13 lp_i/s_i | 14-15 fp_k/(np sf_k) | 16 pair matched | 17 point matched |
 *   18 np/4 | 19 sum(lp)/(np s_i) | 20-21 (p_i - center).t_k/0.5 |
 *   22 (wA - wB).n 0.5/u | 23 friction coefficient
 * Targets (logged after the solve, same units): (sum_j lambda_j - sum_j lp_j)
 * / (np s_i) -- the manifold's total load change, shared by its points, since
 * the split of the load over the points is not determined -- and
 * (f_k - fp_k)/(np sf_k); the initial impulses are lp_i + s_i y0 and
 * fp_k + sf_k sum_i y_k.
 */
#include "b3h_internal.h"
#include "b3h_mlp.h"

#define WARM_EMPTY 0xffffffffffffffffull

__device__ __forceinline__ unsigned long long pair_key(unsigned a, unsigned b)
{
	return ((unsigned long long)a << 32) | (unsigned long long)b;
}
/* MurmurHash3's 64-bit finaliser */
__device__ __forceinline__ unsigned mix64(unsigned long long k)
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return (unsigned)k;
}
__device__ __forceinline__ int num_points(const b3hContactConstraint4& c)
{
	int n = 0;
	for (int i = 0; i < 4; i++) n += c.m_jacCoeffInv[i] != 0.f;
	return n;
}

__global__ __launch_bounds__(256) void k_warm_match(const int* __restrict__ nPtr,
													const b3hContactConstraint4* __restrict__ cs,
													const b3hRigidBodyData* __restrict__ bodies,
													const b3hWarmHist* __restrict__ hist,
													const unsigned long long* __restrict__ htKeys,
													const int* __restrict__ htVals, int htMask,
													const int* __restrict__ histValid, b3hWarmParams wp,
													float* __restrict__ prev, float* __restrict__ warm,
													float* __restrict__ X)
{
	int n = *nPtr;
	int valid = *histValid;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
	{
		const b3hContactConstraint4& c = cs[k];
		int np = num_points(c);
		int h = -1;
		if (valid)
		{
			unsigned long long key = pair_key(c.m_bodyA, c.m_bodyB);
			for (unsigned slot = mix64(key) & (unsigned)htMask;; slot = (slot + 1) & (unsigned)htMask)
			{
				unsigned long long kk = htKeys[slot];
				if (kk == key)
				{
					h = htVals[slot];
					break;
				}
				if (kk == WARM_EMPTY) break;
			}
		}
		F4 nrm = xyz0(c.m_linear);
		F4 t1, t2;
		plane_space(nrm, &t1, &t2);
		float lp[4] = {0.f, 0.f, 0.f, 0.f}, pm[4] = {0.f, 0.f, 0.f, 0.f}, fp[2] = {0.f, 0.f};
		if (h >= 0)
		{
			const b3hWarmHist& H = hist[h];
			int pn = (int)H.fric.w;
			float tol2 = wp.matchTol * wp.matchTol;
			for (int i = 0; i < np; i++)
			{
				F4 p = c.m_worldPos[i];
				int best = -1;
				float bestD = tol2;
				for (int j = 0; j < pn; j++)
				{
					F4 d = sub4(xyz0(p), xyz0(H.pos[j]));
					float dd = dot3(d, d);
					if (dd < bestD)
					{
						bestD = dd;
						best = j;
					}
				}
				if (best >= 0)
				{
					lp[i] = H.pos[best].w;
					pm[i] = 1.f;
				}
			}
			fp[0] = dot3(H.fric, t1);
			fp[1] = dot3(H.fric, t2);
		}
		float* pr = prev + 8 * (size_t)k;
		for (int i = 0; i < 4; i++) pr[i] = lp[i];
		pr[4] = fp[0];
		pr[5] = fp[1];
		pr[6] = h >= 0 ? 1.f : 0.f;
		pr[7] = (float)np;
		if (warm)
		{
			float* w = warm + 6 * (size_t)k;
			for (int i = 0; i < 4; i++) w[i] = lp[i];
			w[4] = fp[0];
			w[5] = fp[1];
		}
		if (!X) continue;
		const b3hRigidBodyData& A = bodies[c.m_bodyA];
		const b3hRigidBodyData& B = bodies[c.m_bodyB];
		float u = wp.velUnit;
		float imS = A.m_invMass + B.m_invMass;
		float sumLp = (lp[0] + lp[1]) + (lp[2] + lp[3]);
		float fnp = np > 0 ? (float)np : 1.f;
		float sf0 = fabsf(c.m_fJacCoeffInv[0]) * u, sf1 = fabsf(c.m_fJacCoeffInv[1]) * u;
		F4 dw = sub4(A.m_angVel, B.m_angVel);
		for (int i = 0; i < 4; i++)
		{
			float* x = X + (size_t)(4 * k + i) * B3H_WARM_FEATURES;
			if (i >= np)
			{
				for (int f = 0; f < B3H_WARM_FEATURES; f++) x[f] = 0.f;
				continue;
			}
			F4 p = xyz0(c.m_worldPos[i]);
			F4 rA = sub4(p, xyz0(A.m_pos)), rB = sub4(p, xyz0(B.m_pos));
			F4 vA = add4(xyz0(A.m_linVel), cross3(A.m_angVel, rA));
			F4 vB = add4(xyz0(B.m_linVel), cross3(B.m_angVel, rB));
			F4 dv = sub4(vA, vB);
			float si = fabsf(c.m_jacCoeffInv[i]) * u;
			F4 off = sub4(p, xyz0(c.m_center));
			x[0] = nrm.x;
			x[1] = nrm.y;
			x[2] = nrm.z;
			x[3] = dot3(dv, nrm) / u;
			x[4] = dot3(dv, t1) / u;
			x[5] = dot3(dv, t2) / u;
			x[6] = c.m_b[i] / u;
			x[7] = c.m_worldPos[i].w / 0.05f;
			x[8] = A.m_invMass / imS;
			x[9] = B.m_invMass / imS;
			x[10] = fabsf(c.m_jacCoeffInv[i]) * imS;
			x[11] = fabsf(c.m_fJacCoeffInv[0]) * imS;
			x[12] = fabsf(c.m_fJacCoeffInv[1]) * imS;
			x[13] = lp[i] / si;
			x[14] = fp[0] / (fnp * sf0);
			x[15] = fp[1] / (fnp * sf1);
			x[16] = h >= 0 ? 1.f : 0.f;
			x[17] = pm[i];
			x[18] = fnp * 0.25f;
			x[19] = sumLp / (fnp * si);
			x[20] = dot3(off, t1) / 0.5f;
			x[21] = dot3(off, t2) / 0.5f;
			x[22] = dot3(dw, nrm) * 0.5f / u;
			x[23] = c.m_linear.w;
		}
	}
}

__global__ __launch_bounds__(256) void k_warm_from_net(const int* __restrict__ nPtr,
													   const b3hContactConstraint4* __restrict__ cs,
													   const float* __restrict__ prev, const float* __restrict__ Y,
													   float velUnit, float* __restrict__ warm)
{
	int n = *nPtr;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
	{
		const b3hContactConstraint4& c = cs[k];
		const float* pr = prev + 8 * (size_t)k;
		int np = (int)pr[7];
		float* w = warm + 6 * (size_t)k;
		float f0 = pr[4], f1 = pr[5];
		float sf0 = fabsf(c.m_fJacCoeffInv[0]) * velUnit, sf1 = fabsf(c.m_fJacCoeffInv[1]) * velUnit;
		float v[4];
		bool finite = isfinite(f0) && isfinite(f1);
		for (int i = 0; i < 4; i++)
		{
			if (i >= np)
			{
				v[i] = 0.f;
				continue;
			}
			const float* y = Y + (size_t)(4 * k + i) * B3H_WARM_OUTPUTS;
			float si = fabsf(c.m_jacCoeffInv[i]) * velUnit;
			v[i] = cl_max(pr[i] + si * y[0], 0.f);
			f0 += sf0 * y[1];
			f1 += sf1 * y[2];
			finite = finite && isfinite(v[i]);
		}
		finite = finite && isfinite(f0) && isfinite(f1);
		/* a non-finite network output falls back to the previous-frame guess
		 * for the whole constraint (the warm pass then clamps the friction
		 * rows to the friction bound of the normal impulses, as the sweeps do) */
		for (int i = 0; i < 4; i++) w[i] = finite ? v[i] : pr[i];
		w[4] = finite ? f0 : pr[4];
		w[5] = finite ? f1 : pr[5];
	}
}

/* targets of the logged rows: the solve's final impulses relative to the
 * previous-frame guess; rowCount = points per constraint */
__global__ __launch_bounds__(256) void k_warm_targets(const int* __restrict__ nPtr,
													  const b3hContactConstraint4* __restrict__ cs,
													  const float* __restrict__ prev, float velUnit,
													  float* __restrict__ T, int* __restrict__ rowCount)
{
	int n = *nPtr;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
	{
		const b3hContactConstraint4& c = cs[k];
		const float* pr = prev + 8 * (size_t)k;
		int np = (int)pr[7];
		float fnp = np > 0 ? (float)np : 1.f;
		float sf0 = fabsf(c.m_fJacCoeffInv[0]) * velUnit, sf1 = fabsf(c.m_fJacCoeffInv[1]) * velUnit;
		float t1 = (c.m_fAppliedRambdaDt[0] - pr[4]) / (fnp * sf0);
		float t2 = (c.m_fAppliedRambdaDt[1] - pr[5]) / (fnp * sf1);
		/* the normal target is the manifold's total: how the load spreads over
		 * its points is not determined by the contact problem (PGS leaves it to
		 * the sweep order and the initial iterate) */
		float tot = 0.f, totPrev = 0.f;
		for (int i = 0; i < np; i++)
		{
			tot += c.m_appliedRambdaDt[i];
			totPrev += pr[i];
		}
		for (int i = 0; i < 4; i++)
		{
			float* t = T + (size_t)(4 * k + i) * B3H_WARM_OUTPUTS;
			if (i < np)
			{
				float si = fabsf(c.m_jacCoeffInv[i]) * velUnit;
				t[0] = (tot - totPrev) / (fnp * si);
				t[1] = t1;
				t[2] = t2;
			}
			else
				t[0] = t[1] = t[2] = 0.f;
		}
		rowCount[k] = np;
	}
}

/* append the valid rows (features, targets) to the log, in constraint order */
__global__ __launch_bounds__(256) void k_warm_log(const int* __restrict__ nPtr, const int* __restrict__ rowScan,
												  const float* __restrict__ X, const float* __restrict__ T,
												  float* __restrict__ logX, float* __restrict__ logT, int logCap,
												  const int* __restrict__ logCount)
{
	int n = *nPtr;
	int base0 = *logCount;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
	{
		int np = rowScan[k + 1] - rowScan[k];
		for (int i = 0; i < np; i++)
		{
			long long dst = (long long)base0 + rowScan[k] + i;
			if (dst >= logCap) break;
			const float* x = X + (size_t)(4 * k + i) * B3H_WARM_FEATURES;
			float* lx = logX + (size_t)dst * B3H_WARM_FEATURES;
			for (int f = 0; f < B3H_WARM_FEATURES; f++) lx[f] = x[f];
			const float* t = T + (size_t)(4 * k + i) * B3H_WARM_OUTPUTS;
			float* lt = logT + (size_t)dst * B3H_WARM_OUTPUTS;
			for (int f = 0; f < B3H_WARM_OUTPUTS; f++) lt[f] = t[f];
		}
	}
}
__global__ void k_warm_log_advance(const int* __restrict__ nPtr, const int* __restrict__ rowScan, int logCap,
								   int* __restrict__ logCount)
{
	long long c = (long long)*logCount + rowScan[*nPtr];
	*logCount = c < logCap ? (int)c : logCap;
}

__device__ __forceinline__ void ht_insert(unsigned long long key, int k, unsigned long long* __restrict__ htKeys,
										  int* __restrict__ htVals, int htMask)
{
	for (unsigned slot = mix64(key) & (unsigned)htMask;; slot = (slot + 1) & (unsigned)htMask)
	{
		unsigned long long old = atomicCAS(&htKeys[slot], WARM_EMPTY, key);
		if (old == WARM_EMPTY || old == key)
		{
			htVals[slot] = k;
			break;
		}
	}
}

/* history for the next step */
__global__ __launch_bounds__(256) void k_warm_store(const int* __restrict__ nPtr,
													const b3hContactConstraint4* __restrict__ cs,
													b3hWarmHist* __restrict__ hist,
													unsigned long long* __restrict__ histKeys,
													unsigned long long* __restrict__ htKeys, int* __restrict__ htVals,
													int htMask)
{
	int n = *nPtr;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
	{
		const b3hContactConstraint4& c = cs[k];
		int np = num_points(c);
		F4 nrm = xyz0(c.m_linear);
		F4 t1, t2;
		plane_space(nrm, &t1, &t2);
		b3hWarmHist H;
		for (int i = 0; i < 4; i++)
		{
			F4 p = c.m_worldPos[i];
			H.pos[i] = f4(p.x, p.y, p.z, i < np ? c.m_appliedRambdaDt[i] : 0.f);
		}
		F4 fr = add4(mul4(t1, c.m_fAppliedRambdaDt[0]), mul4(t2, c.m_fAppliedRambdaDt[1]));
		H.fric = f4(fr.x, fr.y, fr.z, (float)np);
		hist[k] = H;
		unsigned long long key = pair_key(c.m_bodyA, c.m_bodyB);
		histKeys[k] = key;
		ht_insert(key, k, htKeys, htVals, htMask);
	}
}
/* the entry count doubles as the validity flag of the history */
__global__ void k_warm_valid(const int* __restrict__ nPtr, int* __restrict__ histValid) { *histValid = *nPtr; }
__global__ void k_warm_set_count(int n, int* __restrict__ histValid) { *histValid = n; }
__global__ __launch_bounds__(256) void k_warm_reinsert(int n, const unsigned long long* __restrict__ histKeys,
													   unsigned long long* __restrict__ htKeys,
													   int* __restrict__ htVals, int htMask)
{
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
		ht_insert(histKeys[k], k, htKeys, htVals, htMask);
}

static inline int wblocks(int maxN)
{
	int b = (maxN + 255) / 256;
	if (b > 4096) b = 4096;
	return b < 1 ? 1 : b;
}

void b3h_warm_prepare(const int* numContacts, int maxContacts, const b3hContactConstraint4* cs,
					  const b3hRigidBodyData* bodies, b3hWarmBuffers& wb, const b3hWarmParams& wp, const b3hMlp* net,
					  bool features, hipStream_t s)
{
	int gb = wblocks(maxContacts);
	bool learned = wp.mode == 2 && net;
	float* X = (learned || features) ? wb.X : nullptr;
	hipLaunchKernelGGL(k_warm_match, dim3(gb), dim3(256), 0, s, numContacts, cs, bodies, (const b3hWarmHist*)wb.hist,
					   (const unsigned long long*)wb.htKeys, (const int*)wb.htVals, wb.htCap - 1,
					   (const int*)wb.histValid, wp, wb.prev, learned ? (float*)nullptr : wb.warm, X);
	if (learned)
	{
		b3h_mlp_run_forward_dev(*net, numContacts, 4, 4 * maxContacts, wb.X, B3H_WARM_FEATURES, wb.Y,
								B3H_WARM_OUTPUTS, s);
		hipLaunchKernelGGL(k_warm_from_net, dim3(gb), dim3(256), 0, s, numContacts, cs, (const float*)wb.prev,
						   (const float*)wb.Y, wp.velUnit, wb.warm);
	}
}

void b3h_warm_finish(const int* numContacts, int maxContacts, const b3hContactConstraint4* cs, b3hWarmBuffers& wb,
					 const b3hWarmParams& wp, bool logRows, hipStream_t s)
{
	int gb = wblocks(maxContacts);
	if (logRows && wb.logX && wb.logT && wb.logCount)
	{
		hipLaunchKernelGGL(k_warm_targets, dim3(gb), dim3(256), 0, s, numContacts, cs, (const float*)wb.prev,
						   wp.velUnit, wb.T, wb.rowCount);
		b3h_exclusive_scan(wb.rowCount, wb.rowScan, numContacts, maxContacts, wb.scanTmp, s);
		hipLaunchKernelGGL(k_warm_log, dim3(gb), dim3(256), 0, s, numContacts, (const int*)wb.rowScan,
						   (const float*)wb.X, (const float*)wb.T, wb.logX, wb.logT, wb.logCap,
						   (const int*)wb.logCount);
		hipLaunchKernelGGL(k_warm_log_advance, dim3(1), dim3(1), 0, s, numContacts, (const int*)wb.rowScan, wb.logCap,
						   wb.logCount);
	}
	b3h_fill_i32((int*)wb.htKeys, -1, 2 * wb.htCap, s);
	hipLaunchKernelGGL(k_warm_store, dim3(gb), dim3(256), 0, s, numContacts, cs, wb.hist, wb.histKeys, wb.htKeys,
					   wb.htVals, wb.htCap - 1);
	hipLaunchKernelGGL(k_warm_valid, dim3(1), dim3(1), 0, s, numContacts, wb.histValid);
}

void b3h_warm_rebuild(int n, b3hWarmBuffers& wb, hipStream_t s)
{
	b3h_fill_i32((int*)wb.htKeys, -1, 2 * wb.htCap, s);
	if (n > 0)
		hipLaunchKernelGGL(k_warm_reinsert, dim3(wblocks(n)), dim3(256), 0, s, n,
						   (const unsigned long long*)wb.histKeys, wb.htKeys, wb.htVals, wb.htCap - 1);
	hipLaunchKernelGGL(k_warm_set_count, dim3(1), dim3(1), 0, s, n, wb.histValid);
}
