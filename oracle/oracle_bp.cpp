/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * World AABBs, integration, uniform-grid / SAP / brute-force broadphase and the
 * stable key sort, restated on the CPU.
 */
#include "oracle.h"
#include <vector>
#include <algorithm>
#include <stdlib.h>

/* b3ComputeWorldAabb + b3TransformAabb2 (margin 0):
 * src/Bullet3Collision/NarrowPhaseCollision/shared/b3UpdateAabbs.h:8-33,
 * src/Bullet3Collision/BroadPhaseCollision/shared/b3Aabb.h:24-43,
 * rotation matrix as src/Bullet3Common/shared/b3Mat3x3.h:37-60.
 * m_signedMaxIndices[3] is taken from the body itself (the reference indexes
 * body[bodyId] there, out of range; the lane is never read downstream). */
void o_quat_to_abs_rows(F4 q, F4 rows[3])
{
	F4 q2 = f4(q.x * q.x, q.y * q.y, q.z * q.z, 0.f);
	F4 r0, r1, r2;
	r0.x = 1 - 2 * q2.y - 2 * q2.z;
	r0.y = 2 * q.x * q.y - 2 * q.w * q.z;
	r0.z = 2 * q.x * q.z + 2 * q.w * q.y;
	r1.x = 2 * q.x * q.y + 2 * q.w * q.z;
	r1.y = 1 - 2 * q2.x - 2 * q2.z;
	r1.z = 2 * q.y * q.z - 2 * q.w * q.x;
	r2.x = 2 * q.x * q.z - 2 * q.w * q.y;
	r2.y = 2 * q.y * q.z + 2 * q.w * q.x;
	r2.z = 1 - 2 * q2.x - 2 * q2.y;
	rows[0] = f4(fabsf(r0.x), fabsf(r0.y), fabsf(r0.z), 0.f);
	rows[1] = f4(fabsf(r1.x), fabsf(r1.y), fabsf(r1.z), 0.f);
	rows[2] = f4(fabsf(r2.x), fabsf(r2.y), fabsf(r2.z), 0.f);
}

extern "C" void b3o_compute_world_aabbs(int n, const b3hRigidBodyData* bodies, const b3hCollidable* cols,
										  const b3hAabb* localAabbs, b3hAabb* worldAabbs)
{
#pragma omp parallel for num_threads(b3o_threads()) schedule(static)
	for (int i = 0; i < n; i++)
	{
		const b3hRigidBodyData& b = bodies[i];
		int ci = b.m_collidableIdx;
		if (cols[ci].m_shapeIndex < 0) continue;
		b3hAabb la = localAabbs[ci];
		F4 half = smul4(0.5f, sub4(la.m_max, la.m_min));
		half = add4(half, f4(0.f, 0.f, 0.f, 0.f));
		F4 center = smul4(0.5f, add4(la.m_max, la.m_min));
		F4 rows[3];
		o_quat_to_abs_rows(b.m_quat, rows);
		F4 c = xform(center, b.m_pos, b.m_quat);
		F4 ext = f4(dot3(half, rows[0]), dot3(half, rows[1]), dot3(half, rows[2]), 0.f);
		b3hAabb w;
		w.m_min = sub4(c, ext);
		w.m_max = add4(c, ext);
		w.m_min.w = i2f(i);
		w.m_max.w = i2f(b.m_invMass == 0.f ? 0 : 1);
		worldAabbs[i] = w;
	}
}

/* integrateSingleTransform: src/Bullet3Dynamics/shared/b3IntegrateTransforms.h:5-55 */
extern "C" void b3o_integrate(int n, b3hRigidBodyData* bodies, float dt, float angDamp, const float* grav3)
{
	F4 g = f4(grav3[0], grav3[1], grav3[2], 0.f);
#pragma omp parallel for num_threads(b3o_threads()) schedule(static)
	for (int i = 0; i < n; i++)
	{
		b3hRigidBodyData& b = bodies[i];
		if (b.m_invMass == 0.f) continue;
		const float thr = 0.25f * 3.14159254f;
		b.m_angVel.x *= angDamp;
		b.m_angVel.y *= angDamp;
		b.m_angVel.z *= angDamp;
		F4 av = b.m_angVel;
		float ang = sqrtf(dot3(av, av));
		if (ang * dt > thr) ang = thr / dt;
		F4 axis;
		if (ang < 0.001f)
			axis = mul4(av, 0.5f * dt - (dt * dt * dt) * 0.020833333333f * ang * ang);
		else
			axis = mul4(av, sin_small(0.5f * ang * dt) / ang);
		F4 dorn = f4(axis.x, axis.y, axis.z, cos_small(ang * dt * 0.5f));
		F4 q = qmul(dorn, b.m_quat);
		float len = sqrtf(dot4(q, q));
		if (len > 0.f)
			q = mul4(q, 1.f / len);
		else
			q = f4(0, 0, 0, 1);
		b.m_quat = q;
		b.m_pos = add4(b.m_pos, mul4(b.m_linVel, dt));
		b.m_linVel = add4(b.m_linVel, mul4(g, dt));
	}
}

/* stable sort of (key,value) by unsigned key == b3RadixSort32CL::executeHost
 * (src/Bullet3OpenCL/ParallelPrimitives/b3RadixSort32CL.cpp:89-164) */
extern "C" void b3o_sort_keys(int n, b3hSortData* kv)
{
	std::stable_sort(kv, kv + n, [](const b3hSortData& a, const b3hSortData& b) { return a.m_key < b.m_key; });
}

/* inclusive overlap test: gridBroadphase.cl:89-94, sap.cl:37-44 */
static inline bool overlap(const b3hAabb& a, const b3hAabb& b)
{
	return (a.m_min.x <= b.m_max.x) && (b.m_min.x <= a.m_max.x) &&
		   (a.m_min.y <= b.m_max.y) && (b.m_min.y <= a.m_max.y) &&
		   (a.m_min.z <= b.m_max.z) && (b.m_min.z <= a.m_max.z);
}
static inline int handle(const b3hAabb& a) { return f2i(a.m_min.w); }

/* large x small pairs: computePairsKernelTwoArrays, sap.cl:66-105 (pair
 * normalised min,max).  Emitted small-major, large-minor. */
static void large_small(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
						std::vector<b3hInt4>& out)
{
	for (int i = 0; i < nSmall; i++)
		for (int j = 0; j < nLarge; j++)
		{
			const b3hAabb& s = all[smallMap[i]];
			const b3hAabb& l = all[largeMap[j]];
			if (overlap(l, s))
			{
				int x = handle(l), y = handle(s);
				if (x > y) { int t = x; x = y; y = t; }
				b3hInt4 p = {x, y, -1, -1};
				out.push_back(p);
			}
		}
}

/* Uniform grid: kCalcHashAABB / kFindCellStart / kFindOverlappingPairs,
 * src/Bullet3OpenCL/BroadphaseCollision/kernels/gridBroadphase.cl:3-211,
 * host flow b3GpuGridBroadphase.cpp:117-259.  Pair order: large x small pairs
 * first, then per sorted small AABB the 27 neighbour cells (z,y,x) in bucket
 * order. */
extern "C" int b3o_grid_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
							   const float* cellSize3, const int* dims3, int maxBodiesPerCell,
							   b3hInt4* pairsOut, int maxPairs, int* totalFound)
{
	std::vector<b3hInt4> out;
	large_small(all, smallMap, nSmall, largeMap, nLarge, out);
	float inv[3] = {1.f / cellSize3[0], 1.f / cellSize3[1], 1.f / cellSize3[2]};
	int dx = dims3[0], dy = dims3[1], dz = dims3[2];
	std::vector<b3hSortData> hash(nSmall);
	auto gridpos = [&](const b3hAabb& a, int g[3]) {
		float px = (a.m_min.x + a.m_max.x) * 0.5f;
		float py = (a.m_min.y + a.m_max.y) * 0.5f;
		float pz = (a.m_min.z + a.m_max.z) * 0.5f;
		g[0] = (int)floorf(px * inv[0]) & (dx - 1);
		g[1] = (int)floorf(py * inv[1]) & (dy - 1);
		g[2] = (int)floorf(pz * inv[2]) & (dz - 1);
	};
	auto poshash = [&](int x, int y, int z) {
		x &= dx - 1; y &= dy - 1; z &= dz - 1;
		return z * dy * dx + y * dx + x;
	};
	for (int i = 0; i < nSmall; i++)
	{
		int g[3];
		gridpos(all[smallMap[i]], g);
		hash[i].m_key = (unsigned)poshash(g[0], g[1], g[2]);
		hash[i].m_value = (unsigned)i;
	}
	b3o_sort_keys(nSmall, hash.data());
	std::vector<int> cellStart((size_t)dx * dy * dz, -1);
	for (int i = 0; i < nSmall; i++)
		if (i == 0 || hash[i].m_key != hash[i - 1].m_key) cellStart[hash[i].m_key] = i;
	for (int s = 0; s < nSmall; s++)
	{
		int ui = (int)hash[s].m_value;
		const b3hAabb& a = all[smallMap[ui]];
		int g[3];
		gridpos(a, g);
		int ha = handle(a);
		for (int z = -1; z <= 1; z++)
			for (int y = -1; y <= 1; y++)
				for (int x = -1; x <= 1; x++)
				{
					int h = poshash(g[0] + x, g[1] + y, g[2] + z);
					int st = cellStart[h];
					if (st == -1) continue;
					int end = st + maxBodiesPerCell;
					if (end > nSmall) end = nSmall;
					for (int k = st; k < end; k++)
					{
						if ((int)hash[k].m_key != h) break;
						int uj = (int)hash[k].m_value;
						if (uj == ui) continue;
						const b3hAabb& b = all[smallMap[uj]];
						if (!overlap(a, b)) continue;
						int hb = handle(b);
						if (ha < hb)
						{
							b3hInt4 p = {ha, hb, -1, -1};
							out.push_back(p);
						}
					}
				}
	}
	int total = (int)out.size();
	if (totalFound) *totalFound = total;
	int n = total < maxPairs ? total : maxPairs;
	for (int i = 0; i < n; i++) pairsOut[i] = out[i];
	return n;
}

/* order-preserving float -> uint (sap.cl:324-330) */
static inline unsigned float_flip(float f)
{
	unsigned u = f2u(f);
	unsigned mask = (unsigned)(-(int)(u >> 31)) | 0x80000000u;
	return u ^ mask;
}

/* variance axis: prepareSumVarianceKernel + float4 scans (sap.cl:377-389,
 * b3GpuSapBroadphase.cpp:1018-1052).  The sums are accumulated in double so
 * that the axis does not depend on the summation order. */
extern "C" int b3o_sap_axis(const b3hAabb* all, const int* smallMap, int nSmall)
{
	if (nSmall == 0) return 0;
	double s[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
	for (int i = 0; i < nSmall; i++)
	{
		const b3hAabb& a = all[smallMap[i]];
		float c[3] = {(a.m_max.x + a.m_min.x) * 0.5f, (a.m_max.y + a.m_min.y) * 0.5f, (a.m_max.z + a.m_min.z) * 0.5f};
		for (int k = 0; k < 3; k++)
		{
			s[k] += (double)c[k];
			s2[k] += (double)(c[k] * c[k]);
		}
	}
	double v[3];
	for (int k = 0; k < 3; k++) v[k] = s2[k] - (s[k] * s[k]) / (double)nSmall;
	int axis = 0;
	if (v[1] > v[0]) axis = 1;
	if (v[2] > v[axis]) axis = 2;
	return axis;
}

/* 1-axis sweep and prune: flipFloatKernel, radix sort, scatterKernel and
 * computePairsKernelLocalSharedMemory (sap.cl:231-373; host
 * b3GpuSapBroadphase.cpp:997-1231).  Pairs are (sorted-earlier, later), not
 * normalised, as sap.cl:291-292.  axisOverride >= 0 forces the axis. */
extern "C" int b3o_sap_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
							  int axisOverride, b3hInt4* pairsOut, int maxPairs, int* totalFound, int* axisOut)
{
	std::vector<b3hInt4> out;
	large_small(all, smallMap, nSmall, largeMap, nLarge, out);
	int axis = axisOverride >= 0 ? axisOverride : b3o_sap_axis(all, smallMap, nSmall);
	if (axisOut) *axisOut = axis;
	std::vector<b3hSortData> kv(nSmall);
	for (int i = 0; i < nSmall; i++)
	{
		const b3hAabb& a = all[smallMap[i]];
		float m = axis == 0 ? a.m_min.x : (axis == 1 ? a.m_min.y : a.m_min.z);
		kv[i].m_key = float_flip(m);
		kv[i].m_value = (unsigned)i;
	}
	b3o_sort_keys(nSmall, kv.data());
	std::vector<b3hAabb> sorted(nSmall);
	for (int i = 0; i < nSmall; i++) sorted[i] = all[smallMap[kv[i].m_value]];
	for (int i = 0; i < nSmall; i++)
	{
		const b3hAabb& a = sorted[i];
		float amax = axis == 0 ? a.m_max.x : (axis == 1 ? a.m_max.y : a.m_max.z);
		for (int j = i + 1; j < nSmall; j++)
		{
			const b3hAabb& b = sorted[j];
			float bmin = axis == 0 ? b.m_min.x : (axis == 1 ? b.m_min.y : b.m_min.z);
			if (amax < bmin) break;
			if (overlap(a, b))
			{
				b3hInt4 p = {handle(a), handle(b), -1, -1};
				out.push_back(p);
			}
		}
	}
	int total = (int)out.size();
	if (totalFound) *totalFound = total;
	int n = total < maxPairs ? total : maxPairs;
	for (int i = 0; i < n; i++) pairsOut[i] = out[i];
	return n;
}

/* brute force over all small/small and small/large combinations, pairs
 * normalised (x<y) and sorted; the reference's own host check
 * b3GpuSapBroadphase::calculateOverlappingPairsHost (b3GpuSapBroadphase.cpp:862-981). */
extern "C" int b3o_brute_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
								b3hInt4* pairsOut, int maxPairs)
{
	std::vector<b3hInt4> out;
	for (int i = 0; i < nSmall; i++)
		for (int j = i + 1; j < nSmall; j++)
		{
			const b3hAabb& a = all[smallMap[i]];
			const b3hAabb& b = all[smallMap[j]];
			if (overlap(a, b))
			{
				int x = handle(a), y = handle(b);
				if (x > y) { int t = x; x = y; y = t; }
				b3hInt4 p = {x, y, -1, -1};
				out.push_back(p);
			}
		}
	large_small(all, smallMap, nSmall, largeMap, nLarge, out);
	for (auto& p : out)
		if (p.x > p.y) { int t = p.x; p.x = p.y; p.y = t; }
	std::sort(out.begin(), out.end(), [](const b3hInt4& a, const b3hInt4& b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
	int n = (int)out.size() < maxPairs ? (int)out.size() : maxPairs;
	for (int i = 0; i < n; i++) pairsOut[i] = out[i];
	return (int)out.size();
}
