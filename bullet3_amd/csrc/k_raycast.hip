/*
 * k_raycast.hip -- ray casts against the rigid bodies (b3GpuRaycast::castRays,
 * src/Bullet3OpenCL/Raycast/b3GpuRaycast.cpp:245-373, kernels
 * Raycast/kernels/rayCastKernels.cl and BroadphaseCollision/kernels/
 * parallelLinearBvh.cl:262-393).
 *
 * The reference builds a linear BVH over the broadphase's AABBs, writes every
 * (ray, body) whose AABB the ray segment crosses into a pair buffer with an
 * atomic counter (at most 16 per ray), sorts the pairs by ray, finds each ray's
 * range with atomics and then tests each ray against its candidate bodies.
 * Here one thread per ray walks the same BVH (stack in LDS) and tests each
 * candidate body as the walk meets it, keeping the nearest hit: no pair
 * buffer, no sort, no atomics, and no cap on the candidates of a ray.  The
 * candidates are the reference's (the same AABB test on the same tree, large
 * proxies after the tree); they are tested in traversal order, so of two
 * bodies hit at exactly the same fraction the first in that order wins (the
 * reference's order is the atomic append order).
 *
 * The AABBs are the broadphase's as the last step left them (computed before
 * that step's integration), as in the reference.
 */
#include "b3h_internal.h"

#define RAY_TPB 128
#define RAY_STACK 64

/* rayIntersectsAabb, parallelLinearBvh.cl:290-322: slab test of the segment
 * [0, len] along the unit direction nd */
__device__ __forceinline__ bool ray_box(F4 o, float len, F4 nd, float lx, float ly, float lz, float hx, float hy,
										float hz)
{
	const float t0x = ((nd.x < 0.f ? hx : lx) - o.x) / nd.x;
	const float t0y = ((nd.y < 0.f ? hy : ly) - o.y) / nd.y;
	const float t0z = ((nd.z < 0.f ? hz : lz) - o.z) / nd.z;
	const float t1x = ((nd.x < 0.f ? lx : hx) - o.x) / nd.x;
	const float t1y = ((nd.y < 0.f ? ly : hy) - o.y) / nd.y;
	const float t1z = ((nd.z < 0.f ? lz : hz) - o.z) / nd.z;
	const float tmin = fmaxf(t0z, fmaxf(t0y, fmaxf(t0x, 0.f)));
	const float tmax = fminf(t1z, fminf(t1y, fminf(t1x, len)));
	return tmin <= tmax;
}

/* rayConvex, rayCastKernels.cl:155-215: clip the local segment against the
 * hull's face planes; enter fraction from -0.3 */
__device__ __forceinline__ bool ray_convex(F4 from, F4 to, int numFaces, int faceOffset,
										   const b3hGpuFace* __restrict__ faces, float* hitFraction, F4* hitNormal)
{
	bool result = true;
	float exitFraction = *hitFraction;
	float enterFraction = -0.3f;
	F4 cur = f4(0.f, 0.f, 0.f, 0.f);
	for (int i = 0; i < numFaces && result; i++)
	{
		const F4 pl = faces[faceOffset + i].m_plane;
		const float fd = dot3(from, pl) + pl.w;
		const float td = dot3(to, pl) + pl.w;
		if (fd < 0.f)
		{
			if (td >= 0.f)
			{
				const float fr = fd / (fd - td);
				if (exitFraction > fr) exitFraction = fr;
			}
		}
		else
		{
			if (td < 0.f)
			{
				const float fr = fd / (fd - td);
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

/* sphere_intersect, rayCastKernels.cl:218-240 */
__device__ __forceinline__ bool ray_sphere(F4 c, float r, F4 from, F4 to, float* hitFraction)
{
	const F4 rs = xyz0(sub4(from, c));
	const F4 d = xyz0(sub4(to, from));
	const float A = dot3(d, d), B = dot3(rs, d), C = dot3(rs, rs) - r * r;
	const float D = B * B - A * C;
	if (D > 0.f)
	{
		const float t = (-B - sqrtf(D)) / A;
		if (t >= 0.f && t < *hitFraction)
		{
			*hitFraction = t;
			return true;
		}
	}
	return false;
}

struct RayState
{
	F4 from, to;
	float frac;
	int body;
	int ignore;
	F4 normal, point;
};

/* the narrow test of rayCastPairsKernel (rayCastKernels.cl:330-393) for one candidate */
__device__ __forceinline__ void ray_body(RayState& st, int b, const b3hRigidBodyData* __restrict__ bodies,
										 const b3hCollidable* __restrict__ cols,
										 const b3hConvexPolyhedronData* __restrict__ convexes,
										 const b3hGpuFace* __restrict__ faces)
{
	if (b == st.ignore) return;
	const b3hRigidBodyData& bd = bodies[b];
	const b3hCollidable col = cols[bd.m_collidableIdx];
	const F4 pos = bd.m_pos, orn = bd.m_quat;
	if (col.m_shapeType == B3H_SHAPE_CONVEX_HULL)
	{
		const F4 invOrn = qinv(orn);
		const F4 invPos = qrot(invOrn, neg4(pos));
		const F4 fl = xyz0(add4(qrot(invOrn, st.from), invPos));
		const F4 tl = xyz0(add4(qrot(invOrn, st.to), invPos));
		const b3hConvexPolyhedronData& cv = convexes[col.m_shapeIndex];
		const int nf = cv.m_numFaces, fo = cv.m_faceOffset;
		if (nf && ray_convex(fl, tl, nf, fo, faces, &st.frac, &st.normal))
		{
			st.body = b;
			const float s = 1.f - st.frac;
			st.point = f4(s * st.from.x + st.frac * st.to.x, s * st.from.y + st.frac * st.to.y,
						  s * st.from.z + st.frac * st.to.z, 0.f);
		}
	}
	if (col.m_shapeType == B3H_SHAPE_SPHERE)
	{
		if (ray_sphere(pos, col.m_radius, st.from, st.to, &st.frac))
		{
			st.body = b;
			const float s = 1.f - st.frac;
			st.point = f4(s * st.from.x + st.frac * st.to.x, s * st.from.y + st.frac * st.to.y,
						  s * st.from.z + st.frac * st.to.z, 0.f);
			st.normal = sub4(st.point, pos);
		}
	}
}

__global__ __launch_bounds__(RAY_TPB) void k_cast_rays(int numRays, const b3hRayInfo* __restrict__ rays,
													   b3hRayHit* __restrict__ hits, int nSmall,
													   const b3hAabb* __restrict__ leaves,
													   const int2* __restrict__ child,
													   const b3hFloat4* __restrict__ box,
													   const int* __restrict__ meta, const b3hAabb* __restrict__ all,
													   const int* __restrict__ largeMap, int nLarge,
													   const b3hRigidBodyData* __restrict__ bodies,
													   const b3hCollidable* __restrict__ cols,
													   const b3hConvexPolyhedronData* __restrict__ convexes,
													   const b3hGpuFace* __restrict__ faces, int* __restrict__ overflow)
{
	__shared__ int stkAll[RAY_STACK * RAY_TPB];
	int* stk = stkAll + threadIdx.x;
	const int i = blockIdx.x * RAY_TPB + threadIdx.x;
	if (i >= numRays) return;
	b3hRayHit h = hits[i];
	RayState st;
	st.from = rays[i].m_from;
	st.to = rays[i].m_to;
	st.frac = 1.f;
	st.body = -1;
	st.ignore = h.m_hitResult2;
	st.normal = f4(0.f, 0.f, 0.f, 0.f);
	st.point = f4(0.f, 0.f, 0.f, 0.f);
	const F4 d = xyz0(sub4(st.to, st.from));
	const F4 nd = nrm3(d);
	const float len = sqrtf(dot3(d, d));
	const F4 o = st.from;
	if (nSmall == 1)
	{
		const b3hAabb a = leaves[0];
		if (ray_box(o, len, nd, a.m_min.x, a.m_min.y, a.m_min.z, a.m_max.x, a.m_max.y, a.m_max.z))
			ray_body(st, f2i(a.m_min.w), bodies, cols, convexes, faces);
	}
	else if (nSmall > 1)
	{
		int v = meta[0] & 0x7fffffff, sp = 0;
		while (true)
		{
			const int2 ch = child[v];
			int next = -1;
#pragma unroll
			for (int k = 0; k < 2; k++)
			{
				const int cc = k == 0 ? ch.y : ch.x;
				if (cc >= 0)
				{
					const b3hAabb a = leaves[cc];
					if (ray_box(o, len, nd, a.m_min.x, a.m_min.y, a.m_min.z, a.m_max.x, a.m_max.y, a.m_max.z))
						ray_body(st, f2i(a.m_min.w), bodies, cols, convexes, faces);
				}
				else
				{
					const int idx = cc & 0x7fffffff;
					const b3hFloat4 lo = box[2 * idx], hi = box[2 * idx + 1];
					if (ray_box(o, len, nd, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z))
					{
						if (next < 0)
							next = idx;
						else if (sp < RAY_STACK)
							stk[(sp++) * RAY_TPB] = idx;
						else
							*overflow = 1;
					}
				}
			}
			if (next < 0)
			{
				if (sp == 0) break;
				next = stk[(--sp) * RAY_TPB];
			}
			v = next;
		}
	}
	for (int j = 0; j < nLarge; j++)
	{
		const b3hAabb a = all[largeMap[j]];
		if (ray_box(o, len, nd, a.m_min.x, a.m_min.y, a.m_min.z, a.m_max.x, a.m_max.y, a.m_max.z))
			ray_body(st, f2i(a.m_min.w), bodies, cols, convexes, faces);
	}
	h.m_hitFraction = 1.f;
	if (st.body >= 0)
	{
		h.m_hitFraction = st.frac;
		h.m_hitPoint = st.point;
		const float l2 = dot4(st.normal, st.normal);
		const float k = 1.f / sqrtf(l2);
		h.m_hitNormal = f4(st.normal.x * k, st.normal.y * k, st.normal.z * k, st.normal.w * k);
		h.m_hitBody = st.body;
	}
	hits[i] = h;
}

void b3h_cast_rays(int numRays, const b3hRayInfo* rays, b3hRayHit* hits, const b3hAabb* all, const int* smallMap,
				   int nSmall, const int* largeMap, int nLarge, const b3hRigidBodyData* bodies,
				   const b3hShapeTables& sh, b3hBroadphaseScratch& sc, hipStream_t s)
{
	if (numRays <= 0) return;
	b3h_fill_i32(sc.bvhMeta, 0, 2, s);
	if (nSmall > 0) b3h_plbvh_build(all, smallMap, nSmall, sc, s);
	hipLaunchKernelGGL(k_cast_rays, dim3((numRays + RAY_TPB - 1) / RAY_TPB), dim3(RAY_TPB), 0, s, numRays, rays, hits,
					   nSmall, (const b3hAabb*)sc.sortedAabbs, (const int2*)sc.bvhChild, (const b3hFloat4*)sc.bvhBox,
					   (const int*)sc.bvhMeta, all, largeMap, nLarge, bodies, sh.collidables, sh.convexes, sh.faces,
					   sc.bvhMeta + 1);
}
