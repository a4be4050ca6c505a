/*
 * k_narrowphase.hip -- convex-convex contact generation, fused into one kernel
 * per pair followed by an order-preserving compaction.
 *
 * Per pair (reference kernels, all in src/Bullet3OpenCL/NarrowphaseCollision/kernels/):
 *   1. MPR penetration + 1-point contact      mpr.cl:14-90 -> b3MprPenetration.h:816-886
 *   2. SAT over face normals of A then B      sat.cl:597-651, 1417-1509
 *   3. exact edge-edge SAT if |EA|*|EB| <= 162 sat.cl:719-798, 1512-1607
 *      else the 162-direction approximation   mpr.cl:180-311
 *   4. clip B's reference face against A's     satClipHullContacts.cl:235-382, 488-541
 *      incident-face side planes, keep depth <= 0.02, reduce to <= 4 points
 *      (extractManifoldSequential :720-810); point 0 stays the MPR point
 *      when MPR found one (clipHullHullKernel :890-1000).
 * The reference runs these as 7 launches with a host sync after each; here one
 * launch writes a pair-indexed contact record and a flag, and a prefix sum over
 * the flags assigns contact slots in pair order (deterministic; the reference
 * uses an atomic counter).  The hull tables are small and instanced, so they
 * stay resident in L2/L1; the kernel is bound by the divergent per-pair
 * arithmetic rather than by HBM.
 */
#include "b3h_internal.h"
#include "unit_sphere_162.h"

#define MPR_TOL 1E-6f
#define MPR_MAX_ITER 1000
#define LARGE_FLOAT 1e18f
#define CLIP_CAP 64

__constant__ float c_unitSphere[162][3];

struct NpCtx
{
	const b3hRigidBodyData* bodies;
	b3hShapeTables sh;
	__device__ const b3hConvexPolyhedronData& hullOf(int body) const
	{
		return sh.convexes[sh.collidables[bodies[body].m_collidableIdx].m_shapeIndex];
	}
};

/* ---------------- MPR ---------------- */
struct Sup
{
	F4 v, v1, v2;
};
struct Simplex
{
	Sup ps[4];
	int last;
};

__device__ __forceinline__ int mpr_is_zero(float v) { return fabsf(v) < FLT_EPSILON; }
__device__ __forceinline__ int mpr_eq(float a_, float b_)
{
	float ab = fabsf(a_ - b_);
	if (fabsf(ab) < FLT_EPSILON) return 1;
	float a = fabsf(a_), b = fabsf(b_);
	if (b > a) return ab < FLT_EPSILON * b;
	return ab < FLT_EPSILON * a;
}
__device__ __forceinline__ int mpr_vec_eq(F4 a, F4 b) { return mpr_eq(a.x, b.x) && mpr_eq(a.y, b.y) && mpr_eq(a.z, b.z); }
__device__ __forceinline__ void mpr_normalize(F4* d)
{
	float k = 1.f / sqrtf(dot3(*d, *d));
	*d = mul4(*d, k);
}
__device__ __forceinline__ void vadd3(F4* v, F4 w)
{
	v->x += w.x;
	v->y += w.y;
	v->z += w.z;
}

__device__ int max_dot(F4 dir, const F4* v, int n)
{
	float best = -LARGE_FLOAT;
	int idx = -1;
	for (int i = 0; i < n; i++)
	{
		float d = dot3(v[i], dir);
		if (d > best)
		{
			best = d;
			idx = i;
		}
	}
	return idx < 0 ? 0 : idx;
}
__device__ F4 convex_support(const NpCtx& C, int body, F4 dirIn)
{
	const b3hRigidBodyData& b = C.bodies[body];
	F4 q = b.m_quat;
	F4 dir = f4(dirIn.x, dirIn.y, dirIn.z, 0.f);
	F4 localDir = qrot(qinv(q), dir);
	const b3hConvexPolyhedronData& h = C.hullOf(body);
	const F4* vs = C.sh.vertices + h.m_vertexOffset;
	F4 p = vs[max_dot(localDir, vs, h.m_numVertices)];
	return xform(p, b.m_pos, q);
}
__device__ void mpr_support(const NpCtx& C, int a, int b, F4 dir, Sup* s)
{
	s->v1 = convex_support(C, a, dir);
	s->v2 = convex_support(C, b, mul4(dir, -1.f));
	s->v = sub4(s->v1, s->v2);
}
__device__ void portal_dir(const Simplex* p, F4* dir)
{
	F4 v2v1 = sub4(p->ps[2].v, p->ps[1].v);
	F4 v3v1 = sub4(p->ps[3].v, p->ps[1].v);
	*dir = cross3(v2v1, v3v1);
	mpr_normalize(dir);
}
__device__ int portal_reach_tolerance(const Simplex* p, const Sup* v4, F4 dir)
{
	float dv1 = dot3(p->ps[1].v, dir);
	float dv2 = dot3(p->ps[2].v, dir);
	float dv3 = dot3(p->ps[3].v, dir);
	float dv4 = dot3(v4->v, dir);
	float d1 = dv4 - dv1, d2 = dv4 - dv2, d3 = dv4 - dv3;
	d1 = d1 < d2 ? d1 : d2;
	d1 = d1 < d3 ? d1 : d3;
	return mpr_eq(d1, MPR_TOL) || d1 < MPR_TOL;
}
__device__ void expand_portal(Simplex* p, const Sup* v4)
{
	F4 v4v0 = cross3(v4->v, p->ps[0].v);
	float d = dot3(p->ps[1].v, v4v0);
	if (d > 0.f)
	{
		d = dot3(p->ps[2].v, v4v0);
		if (d > 0.f)
			p->ps[1] = *v4;
		else
			p->ps[3] = *v4;
	}
	else
	{
		d = dot3(p->ps[3].v, v4v0);
		if (d > 0.f)
			p->ps[2] = *v4;
		else
			p->ps[1] = *v4;
	}
}
__device__ int discover_portal(const NpCtx& C, int a, int b, Simplex* p)
{
	F4 dir, va, vb;
	float d;
	p->ps[0].v1 = C.bodies[a].m_pos;
	p->ps[0].v2 = C.bodies[b].m_pos;
	p->ps[0].v = sub4(p->ps[0].v1, p->ps[0].v2);
	p->last = 0;
	F4 zero = f4(0, 0, 0, 0);
	if (mpr_vec_eq(p->ps[0].v, zero))
	{
		va = f4(FLT_EPSILON * 10.f, 0.f, 0.f, 0.f);
		vadd3(&p->ps[0].v, va);
	}
	dir = mul4(p->ps[0].v, -1.f);
	mpr_normalize(&dir);
	mpr_support(C, a, b, dir, &p->ps[1]);
	p->last = 1;
	d = dot3(p->ps[1].v, dir);
	if (mpr_is_zero(d) || d < 0.f) return -1;
	dir = cross3(p->ps[0].v, p->ps[1].v);
	if (mpr_is_zero(dot3(dir, dir)))
	{
		if (mpr_vec_eq(p->ps[1].v, zero)) return 1;
		return 2;
	}
	mpr_normalize(&dir);
	mpr_support(C, a, b, dir, &p->ps[2]);
	d = dot3(p->ps[2].v, dir);
	if (mpr_is_zero(d) || d < 0.f) return -1;
	p->last = 2;
	va = sub4(p->ps[1].v, p->ps[0].v);
	vb = sub4(p->ps[2].v, p->ps[0].v);
	dir = cross3(va, vb);
	mpr_normalize(&dir);
	d = dot3(dir, p->ps[0].v);
	if (d > 0.f)
	{
		Sup t = p->ps[1];
		p->ps[1] = p->ps[2];
		p->ps[2] = t;
		dir = mul4(dir, -1.f);
	}
	while (p->last + 1 < 4)
	{
		mpr_support(C, a, b, dir, &p->ps[3]);
		d = dot3(p->ps[3].v, dir);
		if (mpr_is_zero(d) || d < 0.f) return -1;
		int cont = 0;
		va = cross3(p->ps[1].v, p->ps[3].v);
		d = dot3(va, p->ps[0].v);
		if (d < 0.f && !mpr_is_zero(d))
		{
			p->ps[2] = p->ps[3];
			cont = 1;
		}
		if (!cont)
		{
			va = cross3(p->ps[3].v, p->ps[2].v);
			d = dot3(va, p->ps[0].v);
			if (d < 0.f && !mpr_is_zero(d))
			{
				p->ps[1] = p->ps[3];
				cont = 1;
			}
		}
		if (cont)
		{
			va = sub4(p->ps[1].v, p->ps[0].v);
			vb = sub4(p->ps[2].v, p->ps[0].v);
			dir = cross3(va, vb);
			mpr_normalize(&dir);
		}
		else
			p->last = 3;
	}
	return 0;
}
__device__ int refine_portal(const NpCtx& C, int a, int b, Simplex* p)
{
	F4 dir;
	Sup v4;
	for (int i = 0; i < MPR_MAX_ITER; i++)
	{
		portal_dir(p, &dir);
		float d = dot3(dir, p->ps[1].v);
		if (mpr_is_zero(d) || d > 0.f) return 0;
		mpr_support(C, a, b, dir, &v4);
		float dv4 = dot3(v4.v, dir);
		int canEnc = mpr_is_zero(dv4) || dv4 > 0.f;
		if (!canEnc || portal_reach_tolerance(p, &v4, dir)) return -1;
		expand_portal(p, &v4);
	}
	return -1;
}
__device__ void find_pos(const Simplex* p, F4* pos)
{
	F4 dir, vec, p1, p2;
	float b[4], sum, inv;
	portal_dir(p, &dir);
	vec = cross3(p->ps[1].v, p->ps[2].v);
	b[0] = dot3(vec, p->ps[3].v);
	vec = cross3(p->ps[3].v, p->ps[2].v);
	b[1] = dot3(vec, p->ps[0].v);
	vec = cross3(p->ps[0].v, p->ps[1].v);
	b[2] = dot3(vec, p->ps[3].v);
	vec = cross3(p->ps[2].v, p->ps[1].v);
	b[3] = dot3(vec, p->ps[0].v);
	sum = b[0] + b[1] + b[2] + b[3];
	if (mpr_is_zero(sum) || sum < 0.f)
	{
		b[0] = 0.f;
		vec = cross3(p->ps[2].v, p->ps[3].v);
		b[1] = dot3(vec, dir);
		vec = cross3(p->ps[3].v, p->ps[1].v);
		b[2] = dot3(vec, dir);
		vec = cross3(p->ps[1].v, p->ps[2].v);
		b[3] = dot3(vec, dir);
		sum = b[1] + b[2] + b[3];
	}
	inv = 1.f / sum;
	p1 = f4(0, 0, 0, 0);
	p2 = f4(0, 0, 0, 0);
	for (int i = 0; i < 4; i++)
	{
		vec = mul4(p->ps[i].v1, b[i]);
		vadd3(&p1, vec);
		vec = mul4(p->ps[i].v2, b[i]);
		vadd3(&p2, vec);
	}
	p1 = mul4(p1, inv);
	p2 = mul4(p2, inv);
	*pos = p1;
	vadd3(pos, p2);
	*pos = mul4(*pos, 0.5f);
}
__device__ float point_segment_dist2(F4 P, F4 x0, F4 b, F4* witness)
{
	float dist, t;
	F4 d = sub4(b, x0);
	F4 a = sub4(x0, P);
	t = -1.f * dot3(a, d);
	t /= dot3(d, d);
	if (t < 0.f || mpr_is_zero(t))
	{
		F4 ab = sub4(x0, P);
		dist = dot3(ab, ab);
		*witness = x0;
	}
	else if (t > 1.f || mpr_eq(t, 1.f))
	{
		F4 ab = sub4(b, P);
		dist = dot3(ab, ab);
		*witness = b;
	}
	else
	{
		*witness = mul4(d, t);
		vadd3(witness, x0);
		F4 ab = sub4(*witness, P);
		dist = dot3(ab, ab);
	}
	return dist;
}
__device__ float point_tri_dist2(F4 P, F4 x0, F4 B, F4 Cc, F4* witness)
{
	F4 d1 = sub4(B, x0), d2 = sub4(Cc, x0), a = sub4(x0, P);
	float v = dot3(d1, d1), w = dot3(d2, d2);
	float p = dot3(a, d1), q = dot3(a, d2), r = dot3(d1, d2);
	float s = (q * r - w * p) / (w * v - r * r);
	float t = (-s * r - q) / w;
	float dist, dist2;
	if ((mpr_is_zero(s) || s > 0.f) && (mpr_eq(s, 1.f) || s < 1.f) && (mpr_is_zero(t) || t > 0.f) &&
		(mpr_eq(t, 1.f) || t < 1.f) && (mpr_eq(t + s, 1.f) || t + s < 1.f))
	{
		d1 = mul4(d1, s);
		d2 = mul4(d2, t);
		*witness = x0;
		vadd3(witness, d1);
		vadd3(witness, d2);
		F4 ab = sub4(*witness, P);
		dist = dot3(ab, ab);
	}
	else
	{
		F4 w2;
		dist = point_segment_dist2(P, x0, B, witness);
		dist2 = point_segment_dist2(P, x0, Cc, &w2);
		if (dist2 < dist)
		{
			dist = dist2;
			*witness = w2;
		}
		dist2 = point_segment_dist2(P, B, Cc, &w2);
		if (dist2 < dist)
		{
			dist = dist2;
			*witness = w2;
		}
	}
	return dist;
}
__device__ void find_penetr(const NpCtx& C, int a, int b, Simplex* p, float* depth, F4* pdir, F4* pos)
{
	F4 dir;
	Sup v4;
	unsigned iterations = 1u;
	F4 zero = f4(0, 0, 0, 0);
	for (int i = 0; i < MPR_MAX_ITER; i++)
	{
		portal_dir(p, &dir);
		mpr_support(C, a, b, dir, &v4);
		if (portal_reach_tolerance(p, &v4, dir) || iterations == MPR_MAX_ITER)
		{
			*depth = point_tri_dist2(zero, p->ps[1].v, p->ps[2].v, p->ps[3].v, pdir);
			*depth = sqrtf(*depth);
			if (mpr_is_zero(pdir->x) && mpr_is_zero(pdir->y) && mpr_is_zero(pdir->z)) *pdir = dir;
			mpr_normalize(pdir);
			find_pos(p, pos);
			return;
		}
		expand_portal(p, &v4);
		iterations++;
	}
}
__device__ int mpr_penetration(const NpCtx& C, int a, int b, float* depth, F4* dir, F4* pos)
{
	Simplex portal;
	int res = discover_portal(C, a, b, &portal);
	if (res == 0)
	{
		if (refine_portal(C, a, b, &portal) < 0) return -1;
		find_penetr(C, a, b, &portal, depth, dir, pos);
		return 0;
	}
	if (res == 1)
	{
		*depth = 0.f;
		*dir = f4(0, 0, 0, 0);
		*pos = portal.ps[1].v1;
		vadd3(pos, portal.ps[1].v2);
		*pos = mul4(*pos, 0.5f);
		return 0;
	}
	if (res == 2)
	{
		*pos = portal.ps[1].v1;
		vadd3(pos, portal.ps[1].v2);
		*pos = mul4(*pos, 0.5f);
		*dir = portal.ps[1].v;
		*depth = sqrtf(dot3(*dir, *dir));
		mpr_normalize(dir);
		return 0;
	}
	return -1;
}

/* ---------------- SAT ---------------- */
__device__ void project(const NpCtx& C, const b3hConvexPolyhedronData& h, F4 pos, F4 orn, F4 dir, float* mn, float* mx)
{
	float lo = FLT_MAX, hi = -FLT_MAX;
	F4 localDir = qinvrot(orn, dir);
	float offset = dot4(pos, dir);
	const F4* vs = C.sh.vertices + h.m_vertexOffset;
	for (int i = 0; i < h.m_numVertices; i++)
	{
		float dp = dot4(vs[i], localDir);
		if (dp < lo) lo = dp;
		if (dp > hi) hi = dp;
	}
	if (lo > hi)
	{
		float t = lo;
		lo = hi;
		hi = t;
	}
	*mn = lo + offset;
	*mx = hi + offset;
}
__device__ bool test_sep_axis(const NpCtx& C, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb, F4 posA,
							  F4 ornA, F4 posB, F4 ornB, F4 axis, float* depth)
{
	float mn0, mx0, mn1, mx1;
	project(C, ha, posA, ornA, axis, &mn0, &mx0);
	project(C, hb, posB, ornB, axis, &mn1, &mx1);
	if (mx0 < mn1 || mx1 < mn0) return false;
	float d0 = mx0 - mn1, d1 = mx1 - mn0;
	*depth = d0 < d1 ? d0 : d1;
	return true;
}
__device__ bool sat_faces(const NpCtx& C, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb, F4 posA,
						  F4 ornA, F4 posB, F4 ornB, F4 dc, F4* sep, float* dmin)
{
	for (int i = 0; i < ha.m_numFaces; i++)
	{
		F4 n = C.sh.faces[ha.m_faceOffset + i].m_plane;
		F4 nw = qrot(ornA, n);
		if (dot3(dc, nw) < 0) nw = mul4(nw, -1.f);
		float d;
		if (!test_sep_axis(C, ha, hb, posA, ornA, posB, ornB, nw, &d)) return false;
		if (d < *dmin)
		{
			*dmin = d;
			*sep = nw;
		}
	}
	if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
	return true;
}
__device__ __forceinline__ bool almost_zero(F4 v) { return !(fabsf(v.x) > 1e-6f || fabsf(v.y) > 1e-6f || fabsf(v.z) > 1e-6f); }
__device__ bool sat_edges(const NpCtx& C, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb, F4 posA,
						  F4 ornA, F4 posB, F4 ornB, F4 dc, F4* sep, float* dmin)
{
	for (int e0 = 0; e0 < ha.m_numUniqueEdges; e0++)
	{
		F4 ew0 = qrot(ornA, C.sh.uniqueEdges[ha.m_uniqueEdgesOffset + e0]);
		for (int e1 = 0; e1 < hb.m_numUniqueEdges; e1++)
		{
			F4 ew1 = qrot(ornB, C.sh.uniqueEdges[hb.m_uniqueEdgesOffset + e1]);
			F4 c = cross3(ew0, ew1);
			if (almost_zero(c)) continue;
			c = nrm3(c);
			if (dot3(dc, c) < 0) c = mul4(c, -1.f);
			float d;
			if (!test_sep_axis(C, ha, hb, posA, ornA, posB, ornB, c, &d)) return false;
			if (d < *dmin)
			{
				*dmin = d;
				*sep = c;
			}
		}
	}
	if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
	return true;
}
__device__ bool sat_sphere(const NpCtx& C, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb, F4 posA,
						   F4 ornA, F4 posB, F4 ornB, F4 dc, F4* sep, float* dmin)
{
	for (int i = 0; i < B3_UNIT_SPHERE_162_COUNT; i++)
	{
		F4 c = f4(c_unitSphere[i][0], c_unitSphere[i][1], c_unitSphere[i][2], 0.f);
		if (dot3(dc, c) > 0) c = mul4(c, -1.f);
		float d;
		if (!test_sep_axis(C, ha, hb, posA, ornA, posB, ornB, c, &d)) return false;
		if (d < *dmin)
		{
			*dmin = d;
			*sep = c;
		}
	}
	if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
	return true;
}

/* ---------------- clipping ---------------- */
__device__ __forceinline__ F4 lerp3(F4 a, F4 b, float t)
{
	return f4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, 0.f);
}
__device__ int clip_face(const F4* in, int nIn, F4 n, float eq, F4* out)
{
	int nOut = 0;
	if (nIn < 2) return 0;
	F4 first = in[nIn - 1];
	float ds = dot3(n, first) + eq;
	for (int ve = 0; ve < nIn; ve++)
	{
		F4 end = in[ve];
		float de = dot3(n, end) + eq;
		if (ds < 0)
		{
			if (de < 0)
			{
				if (nOut < CLIP_CAP) out[nOut++] = end;
			}
			else
			{
				if (nOut < CLIP_CAP) out[nOut++] = lerp3(first, end, ds * 1.f / (ds - de));
			}
		}
		else
		{
			if (de < 0)
			{
				if (nOut < CLIP_CAP) out[nOut++] = lerp3(first, end, ds * 1.f / (ds - de));
				if (nOut < CLIP_CAP) out[nOut++] = end;
			}
		}
		first = end;
		ds = de;
	}
	return nOut;
}
__device__ int clip_hull_hull(const NpCtx& C, F4 sep, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb,
							  F4 posA, F4 ornA, F4 posB, F4 ornB, F4* bufA, F4* bufB, F4* out)
{
	const float minDist = -1e30f, maxDist = 0.02f;
	int closestB = -1;
	float dmax = -FLT_MAX;
	for (int f = 0; f < hb.m_numFaces; f++)
	{
		F4 wn = qrot(ornB, xyz0(C.sh.faces[hb.m_faceOffset + f].m_plane));
		float d = dot3(wn, sep);
		if (d > dmax)
		{
			dmax = d;
			closestB = f;
		}
	}
	if (closestB < 0) return 0;
	int nB = 0;
	{
		const b3hGpuFace& poly = C.sh.faces[hb.m_faceOffset + closestB];
		for (int e = 0; e < poly.m_numIndices && nB < CLIP_CAP; e++)
		{
			F4 v = C.sh.vertices[hb.m_vertexOffset + C.sh.indices[poly.m_indexOffset + e]];
			bufA[nB++] = xform(v, posB, ornB);
		}
	}
	F4* pin = bufA;
	F4* pout = bufB;
	int nIn = nB;
	int closestA = -1;
	float dmin = FLT_MAX;
	for (int f = 0; f < ha.m_numFaces; f++)
	{
		F4 wn = qrot(ornA, xyz0(C.sh.faces[ha.m_faceOffset + f].m_plane));
		float d = dot3(wn, sep);
		if (d < dmin)
		{
			dmin = d;
			closestA = f;
		}
	}
	if (closestA < 0) return 0;
	const b3hGpuFace polyA = C.sh.faces[ha.m_faceOffset + closestA];
	int nvA = polyA.m_numIndices;
	const F4* va = C.sh.vertices + ha.m_vertexOffset;
	const int* ia = C.sh.indices + polyA.m_indexOffset;
	F4 wpn = qrot(ornA, xyz0(polyA.m_plane));
	for (int e0 = 0; e0 < nvA; e0++)
	{
		F4 a = va[ia[e0]];
		F4 b = va[ia[(e0 + 1) % nvA]];
		F4 we0 = qrot(ornA, sub4(a, b));
		F4 pn = neg4(cross3(we0, wpn));
		F4 wa = xform(a, posA, ornA);
		float peq = -dot3(wa, pn);
		int nOut = clip_face(pin, nIn, pn, peq, pout);
		F4* t = pout;
		pout = pin;
		pin = t;
		nIn = nOut;
	}
	int nc = 0;
	float peq = polyA.m_plane.w - dot3(wpn, posA);
	for (int i = 0; i < nIn; i++)
	{
		float depth = dot3(wpn, pin[i]) + peq;
		if (depth <= minDist) depth = minDist;
		if (depth <= maxDist)
		{
			F4 p = pin[i];
			out[nc++] = f4(p.x, p.y, p.z, depth);
		}
	}
	return nc;
}
__device__ int extract_manifold(const F4* p, int n, F4 nearNormal, int* idx)
{
	if (n == 0) return 0;
	n = n < 64 ? n : 64;
	F4 sum = p[0];
	for (int j = 1; j < n; j++) sum = add4(sum, p[j]);
	float fn = (float)n;
	F4 center = f4(sum.x / fn, sum.y / fn, sum.z / fn, sum.w / fn);
	if (n < 4)
	{
		for (int i = 0; i < n; i++) idx[i] = i;
		return n;
	}
	F4 aVector = sub4(p[0], center);
	F4 u = cross3(nearNormal, aVector);
	F4 v = cross3(nearNormal, u);
	u = nrm3(u);
	v = nrm3(v);
	int best0 = 0, best1 = 0, best2 = 0, best3 = 0;
	for (int ie = 0; ie < n; ie++)
	{
		F4 r = sub4(p[ie], center);
		unsigned lo = 0xffu & (unsigned)ie;
		int a0 = (int)((f2u(dot3(u, r)) & 0xffffff00u) | lo);
		int a1 = (int)((f2u(dot3(neg4(u), r)) & 0xffffff00u) | lo);
		int a2 = (int)((f2u(dot3(v, r)) & 0xffffff00u) | lo);
		int a3 = (int)((f2u(dot3(neg4(v), r)) & 0xffffff00u) | lo);
		if (ie == 0 || a0 > best0) best0 = a0;
		if (ie == 0 || a1 > best1) best1 = a1;
		if (ie == 0 || a2 > best2) best2 = a2;
		if (ie == 0 || a3 > best3) best3 = a3;
	}
	idx[0] = best0 & 0xff;
	idx[1] = best1 & 0xff;
	idx[2] = best2 & 0xff;
	idx[3] = best3 & 0xff;
	return 4;
}

/* ---------------- the fused per-pair kernel ---------------- */
__global__ __launch_bounds__(64) void k_narrowphase(const b3hInt4* __restrict__ pairs, const int* __restrict__ numPairsPtr,
													const b3hRigidBodyData* __restrict__ bodies, b3hShapeTables sh,
													b3hContact4Data* __restrict__ staging, int* __restrict__ flags,
													int* __restrict__ sepStateOut, b3hFloat4* __restrict__ sepOut)
{
	int numPairs = *numPairsPtr;
	NpCtx C;
	C.bodies = bodies;
	C.sh = sh;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numPairs; i += gridDim.x * blockDim.x)
	{
		b3hInt4 pr = pairs[i];
		int a = pr.x, b = pr.y;
		const b3hRigidBodyData& A = bodies[a];
		const b3hRigidBodyData& B = bodies[b];
		bool hull = !(A.m_invMass == 0 && B.m_invMass == 0) &&
					sh.collidables[A.m_collidableIdx].m_shapeType == B3H_SHAPE_CONVEX_HULL &&
					sh.collidables[B.m_collidableIdx].m_shapeType == B3H_SHAPE_CONVEX_HULL;
		int state = 0;
		F4 sep = f4(0, 0, 0, 0);
		bool mprHit = false;
		float mprDepth = 0.f;
		F4 mprDir = f4(0, 0, 0, 0), mprPos = f4(0, 0, 0, 0);
		if (hull) mprHit = mpr_penetration(C, a, b, &mprDepth, &mprDir, &mprPos) == 0;
		if (mprHit) sep = neg4(mprDir);
		F4 posA = A.m_pos, posB = B.m_pos;
		F4 ornA = A.m_quat, ornB = B.m_quat;
		if (hull)
		{
			const b3hConvexPolyhedronData& ha = C.hullOf(a);
			const b3hConvexPolyhedronData& hb = C.hullOf(b);
			F4 pA = posA, pB = posB;
			pA.w = 0.f;
			pB.w = 0.f;
			F4 c0 = xform(ha.m_localCenter, pA, ornA);
			F4 c1 = xform(hb.m_localCenter, pB, ornB);
			F4 dc = sub4(c0, c1);
			float dmin = FLT_MAX;
			F4 sepN = sep;
			bool sepA = sat_faces(C, ha, hb, pA, ornA, pB, ornB, dc, &sepN, &dmin);
			float dminSaved = FLT_MAX;
			if (sepA)
			{
				bool sepB = sat_faces(C, hb, ha, pB, ornB, pA, ornA, dc, &sepN, &dmin);
				if (sepB)
				{
					state = 1;
					dminSaved = dmin;
					sep = sepN;
				}
				else
					state = 4;
			}
			if (state != 0)
			{
				float dm = dminSaved;
				F4 s2 = sep;
				bool ok;
				if (ha.m_numUniqueEdges * hb.m_numUniqueEdges <= B3_UNIT_SPHERE_162_COUNT)
					ok = sat_edges(C, ha, hb, pA, ornA, pB, ornB, dc, &s2, &dm);
				else
					ok = sat_sphere(C, ha, hb, pA, ornA, pB, ornB, dc, &s2, &dm);
				if (!ok)
					state = 0;
				else
				{
					state = 1;
					sep = s2;
				}
			}
		}
		if (sepStateOut) sepStateOut[i] = state;
		if (sepOut) sepOut[i] = sep;
		int nReduced = 0;
		int idx[4] = {-1, -1, -1, -1};
		F4 pts[CLIP_CAP];
		if (hull && state == 1)
		{
			F4 bufA[CLIP_CAP], bufB[CLIP_CAP];
			int np = clip_hull_hull(C, sep, C.hullOf(a), C.hullOf(b), posA, ornA, posB, ornB, bufA, bufB, pts);
			if (np > 0) nReduced = extract_manifold(pts, np, neg4(sep), idx);
		}
		int flag = (mprHit || nReduced > 0) ? 1 : 0;
		flags[i] = flag;
		if (!flag) continue;
		b3hContact4Data c;
		for (int k = 0; k < 4; k++) c.m_worldPosB[k] = f4(0, 0, 0, 0);
		c.m_restituitionCoeffCmp = (unsigned short)(0.f * 0xffff);
		c.m_frictionCoeffCmp = (unsigned short)(0.7f * 0xffff);
		c.m_batchIdx = i;
		c.m_bodyAPtrAndSignBit = A.m_invMass == 0 ? -a : a;
		c.m_bodyBPtrAndSignBit = B.m_invMass == 0 ? -b : b;
		c.m_childIndexA = -1;
		c.m_childIndexB = -1;
		c.m_unused1 = 0;
		c.m_unused2 = 0;
		c.m_worldNormalOnB = f4(0, 0, 0, 0);
		if (mprHit)
		{
			c.m_worldNormalOnB = neg4(mprDir);
			F4 p = mprPos;
			p.w = -mprDepth;
			c.m_worldPosB[0] = p;
			c.m_worldNormalOnB.w = 1.f;
		}
		if (nReduced > 0)
		{
			c.m_worldNormalOnB = sep;
			for (int k = 0; k < nReduced; k++)
				if (k > 0 || !mprHit) c.m_worldPosB[k] = pts[idx[k]];
			c.m_worldNormalOnB.w = (float)nReduced;
		}
		staging[i] = c;
	}
}

/* slots in pair order: pairs[i].z = slot or -1, contactPair[slot] = i */
__global__ __launch_bounds__(256) void k_np_compact(b3hInt4* __restrict__ pairs, const int* __restrict__ numPairsPtr,
													const int* __restrict__ flags, const int* __restrict__ scan,
													int* __restrict__ contactPair, int maxContacts)
{
	int numPairs = *numPairsPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numPairs; i += gridDim.x * blockDim.x)
	{
		int slot = flags[i] ? scan[i] : -1;
		if (slot >= maxContacts) slot = -1;
		pairs[i].z = slot;
		if (slot >= 0) contactPair[slot] = i;
	}
}
__global__ void k_np_count(const int* __restrict__ numPairsPtr, const int* __restrict__ scan, int maxContacts,
						   int* __restrict__ numContacts, int* __restrict__ totalContacts)
{
	int t = scan[*numPairsPtr];
	*totalContacts = t;
	*numContacts = t < maxContacts ? t : maxContacts;
}

int b3h_np_module_init()
{
	static bool done = false;
	if (done) return B3H_OK;
	hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_unitSphere), b3UnitSphere162, sizeof(b3UnitSphere162));
	if (e != hipSuccess)
	{
		b3h_set_error("unit-sphere table upload: %s", hipGetErrorString(e));
		return B3H_ERR_HIP;
	}
	done = true;
	return B3H_OK;
}

void b3h_narrowphase(b3hInt4* pairs, const int* numPairs, int maxPairs, const b3hRigidBodyData* bodies,
					 const b3hShapeTables& sh, b3hContact4Data* staging, int* flags, int* flagScan, unsigned* scanTmp,
					 int* contactPair, int maxContacts, int* numContacts, int* sepState, b3hFloat4* sepNormal,
					 int pairsPerGroup, hipStream_t s)
{
	(void)pairsPerGroup;
	int blocks = (maxPairs + 63) / 64;
	if (blocks > 65536) blocks = 65536;
	if (blocks < 1) blocks = 1;
	hipLaunchKernelGGL(k_narrowphase, dim3(blocks), dim3(64), 0, s, pairs, numPairs, bodies, sh, staging, flags, sepState,
					   sepNormal);
	b3h_exclusive_scan(flags, flagScan, numPairs, maxPairs, scanTmp, s);
	int cblocks = (maxPairs + 255) / 256;
	if (cblocks > 16384) cblocks = 16384;
	if (cblocks < 1) cblocks = 1;
	hipLaunchKernelGGL(k_np_compact, dim3(cblocks), dim3(256), 0, s, pairs, numPairs, flags, flagScan, contactPair,
					   maxContacts);
	hipLaunchKernelGGL(k_np_count, dim3(1), dim3(1), 0, s, numPairs, flagScan, maxContacts, numContacts, numContacts + 1);
}

__global__ __launch_bounds__(256) void k_gather_contacts(const int* __restrict__ contactPair, const int* __restrict__ nPtr,
														 const b3hContact4Data* __restrict__ staging,
														 b3hContact4Data* __restrict__ out)
{
	int n = *nPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = staging[contactPair[i]];
}
void b3h_gather_contacts(const int* contactPair, const int* numContacts, int maxContacts, const b3hContact4Data* staging,
						 b3hContact4Data* out, hipStream_t s)
{
	int blocks = (maxContacts + 255) / 256;
	if (blocks > 16384) blocks = 16384;
	if (blocks < 1) blocks = 1;
	hipLaunchKernelGGL(k_gather_contacts, dim3(blocks), dim3(256), 0, s, contactPair, numContacts, staging, out);
}
