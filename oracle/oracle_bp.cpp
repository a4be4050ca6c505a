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

/* ---- parallel linear BVH broadphase (b3GpuParallelLinearBvhBroadphase::
 * calculateOverlappingPairs, b3GpuParallelLinearBvhBroadphase.cpp:49-57 ->
 * b3GpuParallelLinearBvh::build / calculateOverlappingPairs,
 * b3GpuParallelLinearBvh.cpp:120-355, kernels parallelLinearBvh.cl), restated
 * kernel by kernel as the reference runs them: the tree's parents come from the
 * reference's binary searches, its boxes from the per-depth passes.  The
 * reference appends pairs with an atomic counter, so its list order is not
 * defined; the oracle writes the device's fixed order: leaves in Morton order,
 * each leaf's partners in traversal order (right child, then left; of two
 * overlapping internal children the right is entered and the left kept on the
 * stack), then the large x small pairs small-major as (large, small). */
#define PLBVH_INVALID_PREFIX 128
#define PLBVH_ROOT_MARKER -1
static inline bool plbvh_is_leaf(int idx) { return (idx >> 31) == 0; }
static inline int plbvh_strip(int idx) { return idx & 0x7fffffff; }
static inline int plbvh_mark(bool leaf, int idx) { return leaf ? idx : (int)((unsigned)idx | 0x80000000u); }

/* interleaveBits, parallelLinearBvh.cl:43-72 */
static unsigned plbvh_interleave(unsigned x)
{
	x &= 0x000003FF;
	x = (x ^ (x << 16)) & 0xFF0000FF;
	x = (x ^ (x << 8)) & 0x0300F00F;
	x = (x ^ (x << 4)) & 0x030C30C3;
	x = (x ^ (x << 2)) & 0x09249249;
	return x;
}
/* one axis of assignMortonCodesAndAabbIndicies, parallelLinearBvh.cl:105-140;
 * (int)NaN (every centre on one plane) taken as 0 */
static int plbvh_discrete(float rel, float cell)
{
	float g = rel / cell;
	if (g != g) return 0 + 512;
	float f = (g >= 0.0f) ? g : floorf(g);
	int d;
	if (f >= 2147483647.f)
		d = 511;
	else if (f <= -2147483648.f)
		d = -512;
	else
		d = (int)f;
	d = d < -512 ? -512 : (d > 511 ? 511 : d);
	return d + 512;
}
static int plbvh_clz64(unsigned long long x)
{
	if (!x) return 64;
	return __builtin_clzll(x);
}

/* The tree over nLeaves >= 2 leaves (m_* arrays of b3GpuParallelLinearBvh.h). */
struct OPlbvh
{
	std::vector<b3hAabb> leafAabbs;   /* small AABBs, small order */
	std::vector<b3hSortData> morton;  /* sorted (code, small index) */
	std::vector<int> childX, childY;  /* internal nodes */
	std::vector<int> parent;          /* internal nodes (-1 root) */
	std::vector<b3hAabb> nodeAabbs;
	std::vector<int> rangeLo, rangeHi;
	int root;
};

static void plbvh_build(const b3hAabb* all, const int* smallMap, int n, OPlbvh& t)
{
	t.leafAabbs.resize(n);
	for (int i = 0; i < n; i++) t.leafAabbs[i] = all[smallMap[i]]; /* separateAabbs, :80-87 */
	/* findAllNodesMergedAabb, :89-111, as the host loop runs it (:177-196) */
	std::vector<b3hAabb> m(t.leafAabbs);
	for (int need = n; need >= 2; need = need / 2 + need % 2)
	{
		int remain = need / 2 + need % 2, merged = need - remain;
		for (int i = 0; i < merged; i++)
		{
			b3hAabb& a = m[i];
			const b3hAabb& b = m[i + remain];
			a.m_min = f4(fminf(a.m_min.x, b.m_min.x), fminf(a.m_min.y, b.m_min.y), fminf(a.m_min.z, b.m_min.z), 0.f);
			a.m_max = f4(fmaxf(a.m_max.x, b.m_max.x), fmaxf(a.m_max.y, b.m_max.y), fmaxf(a.m_max.z, b.m_max.z), 0.f);
		}
	}
	const b3hAabb mg = m[0];
	/* assignMortonCodesAndAabbIndicies, :113-146 */
	float gc[3] = {(mg.m_min.x + mg.m_max.x) * 0.5f, (mg.m_min.y + mg.m_max.y) * 0.5f, (mg.m_min.z + mg.m_max.z) * 0.5f};
	float cs[3] = {(mg.m_max.x - mg.m_min.x) / 1024.f, (mg.m_max.y - mg.m_min.y) / 1024.f,
				   (mg.m_max.z - mg.m_min.z) / 1024.f};
	t.morton.resize(n);
	for (int i = 0; i < n; i++)
	{
		const b3hAabb& a = t.leafAabbs[i];
		float c[3] = {(a.m_min.x + a.m_max.x) * 0.5f, (a.m_min.y + a.m_max.y) * 0.5f, (a.m_min.z + a.m_max.z) * 0.5f};
		unsigned d[3];
		for (int k = 0; k < 3; k++) d[k] = (unsigned)plbvh_discrete(c[k] - gc[k], cs[k]);
		t.morton[i].m_key = plbvh_interleave(d[0]) << 0 | plbvh_interleave(d[1]) << 1 | plbvh_interleave(d[2]) << 2;
		t.morton[i].m_value = (unsigned)i;
	}
	b3o_sort_keys(n, t.morton.data()); /* m_radixSorter.execute, stable (:206-209) */

	const int ni = n - 1;
	/* computeAdjacentPairCommonPrefix, :440-470 */
	std::vector<unsigned long long> prefix(ni);
	std::vector<int> plen(ni);
	for (int i = 0; i < ni; i++)
	{
		unsigned long long a = ((unsigned long long)t.morton[i].m_key << 32) | (unsigned)i;
		unsigned long long b = ((unsigned long long)t.morton[i + 1].m_key << 32) | (unsigned)(i + 1);
		int len = plbvh_clz64(a ^ b);
		unsigned long long mask = len == 0 ? 0ull : (~0ull) << (64 - len);
		prefix[i] = (a & b) & mask;
		plen[i] = len;
	}
	auto shared = [&](unsigned long long pa, int la, unsigned long long pb, int lb) {
		int c = plbvh_clz64(pa ^ pb);
		int l = la < lb ? la : lb;
		return c < l ? c : l;
	};
	t.childX.assign(ni, 0);
	t.childY.assign(ni, 0);
	t.parent.assign(ni, 0);
	/* buildBinaryRadixTreeLeafNodes, :473-505 */
	for (int leaf = 0; leaf < n; leaf++)
	{
		int ls = leaf - 1, rs = leaf;
		int lp = ls >= 0 ? plen[ls] : PLBVH_INVALID_PREFIX;
		int rp = rs < ni ? plen[rs] : PLBVH_INVALID_PREFIX;
		bool leftHigher = lp > rp;
		if (lp == PLBVH_INVALID_PREFIX) leftHigher = false;
		if (rp == PLBVH_INVALID_PREFIX) leftHigher = true;
		int p = leftHigher ? ls : rs;
		(leftHigher ? t.childY : t.childX)[p] = plbvh_mark(true, leaf);
	}
	/* buildBinaryRadixTreeInternalNodes (binary searches), :507-640 */
	t.root = -1;
	for (int node = 0; node < ni; node++)
	{
		unsigned long long np_ = prefix[node];
		int nl = plen[node];
		int leftIndex = -1;
		{
			int lower = 0, upper = node - 1;
			while (lower <= upper)
			{
				int mid = (lower + upper) / 2;
				if (shared(np_, nl, prefix[mid], plen[mid]) < nl)
				{
					int right = mid + 1;
					if (right < node)
					{
						if (shared(np_, nl, prefix[right], plen[right]) < nl)
						{
							lower = right;
							leftIndex = right;
						}
						else
						{
							leftIndex = mid;
							break;
						}
					}
					else
					{
						leftIndex = mid;
						break;
					}
				}
				else
					upper = mid - 1;
			}
		}
		int rightIndex = -1;
		{
			int lower = node + 1, upper = ni - 1;
			while (lower <= upper)
			{
				int mid = (lower + upper) / 2;
				if (shared(np_, nl, prefix[mid], plen[mid]) < nl)
				{
					int left = mid - 1;
					if (left > node)
					{
						if (shared(np_, nl, prefix[left], plen[left]) < nl)
						{
							upper = left;
							rightIndex = left;
						}
						else
						{
							rightIndex = mid;
							break;
						}
					}
					else
					{
						rightIndex = mid;
						break;
					}
				}
				else
					lower = mid + 1;
			}
		}
		int lp = leftIndex != -1 ? plen[leftIndex] : PLBVH_INVALID_PREFIX;
		int rp = rightIndex != -1 ? plen[rightIndex] : PLBVH_INVALID_PREFIX;
		bool leftHigher = lp > rp;
		if (lp == PLBVH_INVALID_PREFIX)
			leftHigher = false;
		else if (rp == PLBVH_INVALID_PREFIX)
			leftHigher = true;
		int p = leftHigher ? leftIndex : rightIndex;
		bool isRoot = leftIndex == -1 && rightIndex == -1;
		t.parent[node] = isRoot ? PLBVH_ROOT_MARKER : p;
		if (!isRoot)
			(leftHigher ? t.childY : t.childX)[p] = plbvh_mark(false, node);
		else
			t.root = plbvh_mark(false, node);
	}
	/* findDistanceFromRoot + buildBinaryRadixTreeAabbsRecursive, :643-720, per depth from the deepest */
	std::vector<int> dist(ni);
	int maxDist = 0;
	for (int node = 0; node < ni; node++)
	{
		int d = 0, p = t.parent[node];
		while (p != PLBVH_ROOT_MARKER)
		{
			p = t.parent[p];
			d++;
		}
		dist[node] = d;
		if (d > maxDist) maxDist = d;
	}
	t.nodeAabbs.assign(ni, b3hAabb());
	auto child_aabb = [&](int c) -> b3hAabb {
		return plbvh_is_leaf(c) ? t.leafAabbs[t.morton[c].m_value] : t.nodeAabbs[plbvh_strip(c)];
	};
	for (int d = maxDist; d >= 0; d--)
		for (int node = 0; node < ni; node++)
			if (dist[node] == d)
			{
				b3hAabb a = child_aabb(t.childX[node]), b = child_aabb(t.childY[node]);
				b3hAabb o;
				o.m_min = f4(fminf(a.m_min.x, b.m_min.x), fminf(a.m_min.y, b.m_min.y), fminf(a.m_min.z, b.m_min.z), 0.f);
				o.m_max = f4(fmaxf(a.m_max.x, b.m_max.x), fmaxf(a.m_max.y, b.m_max.y), fmaxf(a.m_max.z, b.m_max.z), 0.f);
				t.nodeAabbs[node] = o;
			}
	/* findLeafIndexRanges, :723-767 */
	t.rangeLo.resize(ni);
	t.rangeHi.resize(ni);
	for (int node = 0; node < ni; node++)
	{
		int lo = t.childX[node];
		while (!plbvh_is_leaf(lo)) lo = t.childX[plbvh_strip(lo)];
		int hi = t.childY[node];
		while (!plbvh_is_leaf(hi)) hi = t.childY[plbvh_strip(hi)];
		t.rangeLo[node] = lo;
		t.rangeHi[node] = hi;
	}
}

extern "C" int b3o_plbvh_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
								b3hInt4* pairsOut, int maxPairs, int* totalFound, int* rootOut)
{
	std::vector<b3hInt4> out;
	if (rootOut) *rootOut = nSmall == 1 ? 0 : -1;
	if (nSmall >= 2)
	{
		OPlbvh t;
		plbvh_build(all, smallMap, nSmall, t);
		if (rootOut) *rootOut = t.root;
		/* plbvhCalculateOverlappingPairs, :185-260: query leaf q pairs with
		 * leaves after it, (query, other) */
		for (int q = 0; q < nSmall - 1; q++)
		{
			const b3hAabb& a = t.leafAabbs[t.morton[q].m_value];
			std::vector<int> stack;
			int v = plbvh_strip(t.root);
			while (true)
			{
				int next = -1;
				int ch[2] = {t.childY[v], t.childX[v]};
				for (int k = 0; k < 2; k++)
				{
					int c = ch[k];
					if (plbvh_is_leaf(c))
					{
						if (c <= q) continue;
						const b3hAabb& b = t.leafAabbs[t.morton[c].m_value];
						if (overlap(a, b))
						{
							b3hInt4 p = {handle(a), handle(b), -1, -1};
							out.push_back(p);
						}
					}
					else
					{
						int idx = plbvh_strip(c);
						if (t.rangeHi[idx] <= q) continue;
						if (overlap(a, t.nodeAabbs[idx]))
						{
							if (next < 0)
								next = idx;
							else
								stack.push_back(idx);
						}
					}
				}
				if (next < 0)
				{
					if (stack.empty()) break;
					next = stack.back();
					stack.pop_back();
				}
				v = next;
			}
		}
	}
	/* plbvhLargeAabbAabbTest, :395-420: (large, small), unnormalised */
	for (int i = 0; i < nSmall; i++)
		for (int j = 0; j < nLarge; j++)
		{
			const b3hAabb& s = all[smallMap[i]];
			const b3hAabb& l = all[largeMap[j]];
			if (overlap(s, l))
			{
				b3hInt4 p = {handle(l), handle(s), -1, -1};
				out.push_back(p);
			}
		}
	int total = (int)out.size();
	if (totalFound) *totalFound = total;
	int n = total < maxPairs ? total : maxPairs;
	for (int i = 0; i < n; i++) pairsOut[i] = out[i];
	return n;
}

/* ---- ray casts (b3GpuRaycast::castRays, b3GpuRaycast.cpp:245-373) ----
 * candidates: every body whose broadphase AABB the segment crosses
 * (plbvhRayTraverse / plbvhLargeAabbRayTest, parallelLinearBvh.cl:262-393,
 * rayIntersectsAabb :290-322), visited in the device's order (the BVH walk,
 * right child first, then the large proxies); each candidate tested by
 * rayCastPairsKernel's narrow tests (rayCastKernels.cl:155-240, 330-393),
 * keeping the nearest hit. */
static bool o_ray_box(F4 o, float len, F4 nd, const b3hAabb& a)
{
	float t0x = ((nd.x < 0.f ? a.m_max.x : a.m_min.x) - o.x) / nd.x;
	float t0y = ((nd.y < 0.f ? a.m_max.y : a.m_min.y) - o.y) / nd.y;
	float t0z = ((nd.z < 0.f ? a.m_max.z : a.m_min.z) - o.z) / nd.z;
	float t1x = ((nd.x < 0.f ? a.m_min.x : a.m_max.x) - o.x) / nd.x;
	float t1y = ((nd.y < 0.f ? a.m_min.y : a.m_max.y) - o.y) / nd.y;
	float t1z = ((nd.z < 0.f ? a.m_min.z : a.m_max.z) - o.z) / nd.z;
	float tmin = fmaxf(t0z, fmaxf(t0y, fmaxf(t0x, 0.f)));
	float tmax = fminf(t1z, fminf(t1y, fminf(t1x, len)));
	return tmin <= tmax;
}
static bool o_ray_convex(F4 from, F4 to, const b3hConvexPolyhedronData& cv, const b3hGpuFace* faces, float* hitFraction,
						 F4* hitNormal)
{
	bool result = true;
	float exitFraction = *hitFraction, enterFraction = -0.3f;
	F4 cur = f4(0.f, 0.f, 0.f, 0.f);
	for (int i = 0; i < cv.m_numFaces && result; i++)
	{
		F4 pl = faces[cv.m_faceOffset + i].m_plane;
		float fd = dot3(from, pl) + pl.w;
		float td = dot3(to, pl) + pl.w;
		if (fd < 0.f)
		{
			if (td >= 0.f)
			{
				float fr = fd / (fd - td);
				if (exitFraction > fr) exitFraction = fr;
			}
		}
		else
		{
			if (td < 0.f)
			{
				float fr = fd / (fd - td);
				if (enterFraction <= fr)
				{
					enterFraction = fr;
					cur = f4(pl.x, pl.y, pl.z, 0.f);
				}
			}
			else
				result = false;
		}
		if (exitFraction <= enterFraction) result = false;
	}
	if (enterFraction < 0.f) result = false;
	if (result)
	{
		*hitFraction = enterFraction;
		*hitNormal = cur;
	}
	return result;
}
struct ORay
{
	F4 from, to, normal, point;
	float frac;
	int body, ignore;
};
static void o_ray_body(ORay& st, int b, const b3hRigidBodyData* bodies, const b3hCollidable* cols,
					   const b3hConvexPolyhedronData* convexes, const b3hGpuFace* faces)
{
	if (b == st.ignore) return;
	const b3hRigidBodyData& bd = bodies[b];
	const b3hCollidable& col = cols[bd.m_collidableIdx];
	if (col.m_shapeType == 3) /* SHAPE_CONVEX_HULL */
	{
		F4 invOrn = qinv(bd.m_quat);
		F4 invPos = qrot(invOrn, neg4(bd.m_pos));
		F4 fl = xyz0(add4(qrot(invOrn, st.from), invPos));
		F4 tl = xyz0(add4(qrot(invOrn, st.to), invPos));
		const b3hConvexPolyhedronData& cv = convexes[col.m_shapeIndex];
		if (cv.m_numFaces && o_ray_convex(fl, tl, cv, faces, &st.frac, &st.normal))
		{
			st.body = b;
			float s = 1.f - st.frac;
			st.point = f4(s * st.from.x + st.frac * st.to.x, s * st.from.y + st.frac * st.to.y,
						  s * st.from.z + st.frac * st.to.z, 0.f);
		}
	}
	if (col.m_shapeType == 7) /* SHAPE_SPHERE: sphere_intersect */
	{
		F4 rs = xyz0(sub4(st.from, bd.m_pos));
		F4 d = xyz0(sub4(st.to, st.from));
		float A = dot3(d, d), B = dot3(rs, d), C = dot3(rs, rs) - col.m_radius * col.m_radius;
		float D = B * B - A * C;
		if (D > 0.f)
		{
			float t = (-B - sqrtf(D)) / A;
			if (t >= 0.f && t < st.frac)
			{
				st.frac = t;
				st.body = b;
				float s = 1.f - st.frac;
				st.point = f4(s * st.from.x + st.frac * st.to.x, s * st.from.y + st.frac * st.to.y,
							  s * st.from.z + st.frac * st.to.z, 0.f);
				st.normal = sub4(st.point, bd.m_pos);
			}
		}
	}
}

extern "C" int b3o_cast_rays(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
							  const b3hRigidBodyData* bodies, const b3hCollidable* cols,
							  const b3hConvexPolyhedronData* convexes, const b3hGpuFace* faces, const b3hRayInfo* rays,
							  b3hRayHit* hits, int numRays)
{
	OPlbvh t;
	if (nSmall >= 2) plbvh_build(all, smallMap, nSmall, t);
#pragma omp parallel for num_threads(b3o_threads()) schedule(dynamic, 64)
	for (int i = 0; i < numRays; i++)
	{
		b3hRayHit h = hits[i];
		ORay st;
		st.from = rays[i].m_from;
		st.to = rays[i].m_to;
		st.frac = 1.f;
		st.body = -1;
		st.ignore = h.m_hitResult2;
		st.normal = f4(0.f, 0.f, 0.f, 0.f);
		st.point = f4(0.f, 0.f, 0.f, 0.f);
		F4 d = xyz0(sub4(st.to, st.from));
		F4 nd = nrm3(d);
		float len = sqrtf(dot3(d, d));
		if (nSmall == 1)
		{
			const b3hAabb& a = all[smallMap[0]];
			if (o_ray_box(st.from, len, nd, a)) o_ray_body(st, handle(a), bodies, cols, convexes, faces);
		}
		else if (nSmall >= 2)
		{
			std::vector<int> stack;
			int v = plbvh_strip(t.root);
			while (true)
			{
				int next = -1;
				int ch[2] = {t.childY[v], t.childX[v]};
				for (int k = 0; k < 2; k++)
				{
					int c = ch[k];
					if (plbvh_is_leaf(c))
					{
						const b3hAabb& a = t.leafAabbs[t.morton[c].m_value];
						if (o_ray_box(st.from, len, nd, a)) o_ray_body(st, handle(a), bodies, cols, convexes, faces);
					}
					else
					{
						int idx = plbvh_strip(c);
						if (o_ray_box(st.from, len, nd, t.nodeAabbs[idx]))
						{
							if (next < 0)
								next = idx;
							else
								stack.push_back(idx);
						}
					}
				}
				if (next < 0)
				{
					if (stack.empty()) break;
					next = stack.back();
					stack.pop_back();
				}
				v = next;
			}
		}
		for (int j = 0; j < nLarge; j++)
		{
			const b3hAabb& a = all[largeMap[j]];
			if (o_ray_box(st.from, len, nd, a)) o_ray_body(st, handle(a), bodies, cols, convexes, faces);
		}
		h.m_hitFraction = 1.f;
		if (st.body >= 0)
		{
			h.m_hitFraction = st.frac;
			h.m_hitPoint = st.point;
			float k = 1.f / sqrtf(dot4(st.normal, st.normal));
			h.m_hitNormal = f4(st.normal.x * k, st.normal.y * k, st.normal.z * k, st.normal.w * k);
			h.m_hitBody = st.body;
		}
		hits[i] = h;
	}
	return 0;
}
