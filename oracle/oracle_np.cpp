/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * Convex-convex narrowphase of the Bullet3 GPU pipeline restated on the CPU:
 * MPR contact (mpr.cl:14-90 / b3MprPenetration.h), SAT over face normals
 * (sat.cl:597-651,1417-1509), exact edge-edge SAT (sat.cl:719-798,1512-1607),
 * 162-direction approximation (mpr.cl:180-311), face clipping and manifold
 * reduction (satClipHullContacts.cl:235-382,488-541,720-810,890-1000).
 *
 * Contact slots are assigned in pair order (a pair that yields a contact in
 * either the MPR or the clipping stage takes the next slot); the reference
 * assigns them with atomics in arrival order, which is not deterministic.
 */
#include "oracle.h"
#include "unit_sphere_162.h"
#include <vector>
#include <algorithm>
#include <cstring>

#define MPR_TOL 1E-6f
#define MPR_MAX_ITER 1000
#define LARGE_FLOAT 1e18f

struct OShapes
{
	const b3hRigidBodyData* bodies;
	const b3hCollidable* cols;
	const b3hConvexPolyhedronData* convexes;
	const b3hFloat4* verts;
	const b3hFloat4* edges;
	const b3hGpuFace* faces;
	const int* indices;
	const b3hConvexPolyhedronData& hullOf(int body) const { return convexes[cols[bodies[body].m_collidableIdx].m_shapeIndex]; }
};

/* ---------------- MPR (b3MprPenetration.h) ---------------- */
struct Sup { F4 v, v1, v2; };
struct Simplex { Sup ps[4]; int last; };

static inline int mpr_is_zero(float v) { return fabsf(v) < FLT_EPSILON; }
static inline int mpr_eq(float a_, float b_)
{
	float ab = fabsf(a_ - b_);
	if (fabsf(ab) < FLT_EPSILON) return 1;
	float a = fabsf(a_), b = fabsf(b_);
	if (b > a) return ab < FLT_EPSILON * b;
	return ab < FLT_EPSILON * a;
}
static inline int mpr_vec_eq(F4 a, F4 b) { return mpr_eq(a.x, b.x) && mpr_eq(a.y, b.y) && mpr_eq(a.z, b.z); }
static inline void mpr_normalize(F4* d)
{
	float k = 1.f / sqrtf(dot3(*d, *d));
	*d = mul4(*d, k);
}
/* b3MaxDot, first index of the maximum (b3Float4.h:66-87) */
static int max_dot(F4 dir, const F4* v, int n)
{
	float best = -LARGE_FLOAT;
	int idx = -1;
	for (int i = 0; i < n; i++)
	{
		float d = dot3(v[i], dir);
		if (d > best) { best = d; idx = i; }
	}
	if (idx < 0) idx = 0;
	return idx;
}
/* b3MprConvexSupport (b3MprPenetration.h:139-179) */
static F4 convex_support(const OShapes& S, int body, F4 dirIn)
{
	const b3hRigidBodyData& b = S.bodies[body];
	F4 dir = f4(dirIn.x, dirIn.y, dirIn.z, 0.f);
	F4 localDir = qrot(qinv(b.m_quat), dir);
	const b3hConvexPolyhedronData& h = S.hullOf(body);
	F4 p = S.verts[h.m_vertexOffset + max_dot(localDir, S.verts + h.m_vertexOffset, h.m_numVertices)];
	return xform(p, b.m_pos, b.m_quat);
}
static void mpr_support(const OShapes& S, int a, int b, F4 dir, Sup* s)
{
	s->v1 = convex_support(S, a, dir);
	s->v2 = convex_support(S, b, mul4(dir, -1.f));
	s->v = sub4(s->v1, s->v2);
}
static inline void vadd3(F4* v, F4 w) { v->x += w.x; v->y += w.y; v->z += w.z; }

static void portal_dir(const Simplex* p, F4* dir)
{
	F4 v2v1 = sub4(p->ps[2].v, p->ps[1].v);
	F4 v3v1 = sub4(p->ps[3].v, p->ps[1].v);
	*dir = cross3(v2v1, v3v1);
	mpr_normalize(dir);
}
static int portal_encapsules_origin(const Simplex* p, F4 dir)
{
	float d = dot3(dir, p->ps[1].v);
	return mpr_is_zero(d) || d > 0.f;
}
static int portal_reach_tolerance(const Simplex* p, const Sup* v4, F4 dir)
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
static int portal_can_encapsule_origin(const Sup* v4, F4 dir)
{
	float d = dot3(v4->v, dir);
	return mpr_is_zero(d) || d > 0.f;
}
static void expand_portal(Simplex* p, const Sup* v4)
{
	F4 v4v0 = cross3(v4->v, p->ps[0].v);
	float d = dot3(p->ps[1].v, v4v0);
	if (d > 0.f)
	{
		d = dot3(p->ps[2].v, v4v0);
		if (d > 0.f) p->ps[1] = *v4;
		else p->ps[3] = *v4;
	}
	else
	{
		d = dot3(p->ps[3].v, v4v0);
		if (d > 0.f) p->ps[2] = *v4;
		else p->ps[1] = *v4;
	}
}

/* b3DiscoverPortal (b3MprPenetration.h:346-479) */
static int discover_portal(const OShapes& S, int a, int b, Simplex* p)
{
	F4 dir, va, vb;
	float d;
	p->ps[0].v1 = S.bodies[a].m_pos;
	p->ps[0].v2 = S.bodies[b].m_pos;
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
	mpr_support(S, a, b, dir, &p->ps[1]);
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
	mpr_support(S, a, b, dir, &p->ps[2]);
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
		Sup t = p->ps[1]; p->ps[1] = p->ps[2]; p->ps[2] = t;
		dir = mul4(dir, -1.f);
	}
	while (p->last + 1 < 4)
	{
		mpr_support(S, a, b, dir, &p->ps[3]);
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

/* b3RefinePortal (b3MprPenetration.h:481-517) */
static int refine_portal(const OShapes& S, int a, int b, Simplex* p)
{
	F4 dir;
	Sup v4;
	for (int i = 0; i < MPR_MAX_ITER; i++)
	{
		portal_dir(p, &dir);
		if (portal_encapsules_origin(p, dir)) return 0;
		mpr_support(S, a, b, dir, &v4);
		if (!portal_can_encapsule_origin(&v4, dir) || portal_reach_tolerance(p, &v4, dir)) return -1;
		expand_portal(p, &v4);
	}
	return -1;
}

/* b3FindPos (b3MprPenetration.h:519-580) */
static void find_pos(const Simplex* p, F4* pos)
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

static float point_segment_dist2(F4 P, F4 x0, F4 b, F4* witness)
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

/* b3MprVec3PointTriDist2 with witness (b3MprPenetration.h:648-729) */
static float point_tri_dist2(F4 P, F4 x0, F4 B, F4 C, F4* witness)
{
	F4 d1 = sub4(B, x0), d2 = sub4(C, x0), a = sub4(x0, P);
	float u = dot3(a, a), v = dot3(d1, d1), w = dot3(d2, d2);
	float p = dot3(a, d1), q = dot3(a, d2), r = dot3(d1, d2);
	(void)u;
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
		dist2 = point_segment_dist2(P, x0, C, &w2);
		if (dist2 < dist) { dist = dist2; *witness = w2; }
		dist2 = point_segment_dist2(P, B, C, &w2);
		if (dist2 < dist) { dist = dist2; *witness = w2; }
	}
	return dist;
}

/* b3FindPenetr (b3MprPenetration.h:731-777) */
static void find_penetr(const OShapes& S, int a, int b, Simplex* p, float* depth, F4* pdir, F4* pos)
{
	F4 dir;
	Sup v4;
	unsigned long iterations = 1UL;
	F4 zero = f4(0, 0, 0, 0);
	for (int i = 0; i < MPR_MAX_ITER; i++)
	{
		portal_dir(p, &dir);
		mpr_support(S, a, b, dir, &v4);
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

/* b3MprPenetration (b3MprPenetration.h:816-886): 0 = contact */
static int mpr_penetration(const OShapes& S, int a, int b, float* depth, F4* dir, F4* pos)
{
	Simplex portal;
	int res = discover_portal(S, a, b, &portal);
	switch (res)
	{
		case 0:
			if (refine_portal(S, a, b, &portal) < 0) return -1;
			find_penetr(S, a, b, &portal, depth, dir, pos);
			return 0;
		case 1:
			*depth = 0.f;
			*dir = f4(0, 0, 0, 0);
			*pos = portal.ps[1].v1;
			vadd3(pos, portal.ps[1].v2);
			*pos = mul4(*pos, 0.5f);
			return 0;
		case 2:
			*pos = portal.ps[1].v1;
			vadd3(pos, portal.ps[1].v2);
			*pos = mul4(*pos, 0.5f);
			*dir = portal.ps[1].v;
			*depth = sqrtf(dot3(*dir, *dir));
			mpr_normalize(dir);
			return 0;
		default:
			return -1;
	}
}

static void write_contact_header(b3hContact4Data* c, const OShapes& S, const b3hInt4& pr, int pairIndex, float friction)
{
	c->m_restituitionCoeffCmp = (unsigned short)(0.f * 0xffff);
	c->m_frictionCoeffCmp = (unsigned short)(friction * 0xffff);
	c->m_batchIdx = pairIndex;
	int bodyA = pr.x, bodyB = pr.y;
	c->m_bodyAPtrAndSignBit = S.bodies[bodyA].m_invMass == 0 ? -bodyA : bodyA;
	c->m_bodyBPtrAndSignBit = S.bodies[bodyB].m_invMass == 0 ? -bodyB : bodyB;
	c->m_childIndexA = -1;
	c->m_childIndexB = -1;
	c->m_unused1 = 0;
	c->m_unused2 = 0;
}

static bool pair_is_hull_hull(const OShapes& S, int a, int b)
{
	if (S.bodies[a].m_invMass == 0 && S.bodies[b].m_invMass == 0) return false;
	if (S.cols[S.bodies[a].m_collidableIdx].m_shapeType != B3H_SHAPE_CONVEX_HULL) return false;
	if (S.cols[S.bodies[b].m_collidableIdx].m_shapeType != B3H_SHAPE_CONVEX_HULL) return false;
	return true;
}

/* ---------------- SAT (sat.cl) ---------------- */
/* project (sat.cl:327-350) */
static void project(const OShapes& S, const b3hConvexPolyhedronData& h, F4 pos, F4 orn, F4 dir, float* mn, float* mx)
{
	*mn = FLT_MAX;
	*mx = -FLT_MAX;
	F4 localDir = qinvrot(orn, dir);
	float offset = dot4(pos, dir);
	for (int i = 0; i < h.m_numVertices; i++)
	{
		float dp = dot4(S.verts[h.m_vertexOffset + i], localDir);
		if (dp < *mn) *mn = dp;
		if (dp > *mx) *mx = dp;
	}
	if (*mn > *mx) { float t = *mn; *mn = *mx; *mx = t; }
	*mn += offset;
	*mx += offset;
}
/* TestSepAxis (sat.cl:577-594): false = separated */
static bool test_sep_axis(const OShapes& S, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb,
						  F4 posA, F4 ornA, F4 posB, F4 ornB, F4 axis, float* depth)
{
	float mn0, mx0, mn1, mx1;
	project(S, ha, posA, ornA, axis, &mn0, &mx0);
	project(S, hb, posB, ornB, axis, &mn1, &mx1);
	if (mx0 < mn1 || mx1 < mn0) return false;
	float d0 = mx0 - mn1, d1 = mx1 - mn0;
	*depth = d0 < d1 ? d0 : d1;
	return true;
}
/* findSeparatingAxis over the faces of hullA (sat.cl:597-651) */
static bool sat_faces(const OShapes& S, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb,
					  F4 posA, F4 ornA, F4 posB, F4 ornB, F4 dc, F4* sep, float* dmin)
{
	posA.w = 0.f;
	posB.w = 0.f;
	for (int i = 0; i < ha.m_numFaces; i++)
	{
		F4 n = S.faces[ha.m_faceOffset + i].m_plane;
		F4 nw = qrot(ornA, n);
		if (dot3(dc, nw) < 0) nw = mul4(nw, -1.f);
		float d;
		if (!test_sep_axis(S, ha, hb, posA, ornA, posB, ornB, nw, &d)) return false;
		if (d < *dmin) { *dmin = d; *sep = nw; }
	}
	if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
	return true;
}
static inline bool almost_zero(F4 v) { return !(fabsf(v.x) > 1e-6f || fabsf(v.y) > 1e-6f || fabsf(v.z) > 1e-6f); }
/* findSeparatingAxisEdgeEdge (sat.cl:719-798) */
static bool sat_edges(const OShapes& S, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb,
					  F4 posA, F4 ornA, F4 posB, F4 ornB, F4 dc, F4* sep, float* dmin)
{
	posA.w = 0.f;
	posB.w = 0.f;
	for (int e0 = 0; e0 < ha.m_numUniqueEdges; e0++)
	{
		F4 ew0 = qrot(ornA, S.edges[ha.m_uniqueEdgesOffset + e0]);
		for (int e1 = 0; e1 < hb.m_numUniqueEdges; e1++)
		{
			F4 ew1 = qrot(ornB, S.edges[hb.m_uniqueEdgesOffset + e1]);
			F4 c = cross3(ew0, ew1);
			if (almost_zero(c)) continue;
			c = nrm3(c);
			if (dot3(dc, c) < 0) c = mul4(c, -1.f);
			float d;
			if (!test_sep_axis(S, ha, hb, posA, ornA, posB, ornB, c, &d)) return false;
			if (d < *dmin) { *dmin = d; *sep = c; }
		}
	}
	if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
	return true;
}
/* findSeparatingAxisUnitSphere (mpr.cl:180-240) */
static bool sat_sphere(const OShapes& S, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb,
					   F4 posA, F4 ornA, F4 posB, F4 ornB, F4 dc, F4* sep, float* dmin)
{
	posA.w = 0.f;
	posB.w = 0.f;
	for (int i = 0; i < B3_UNIT_SPHERE_162_COUNT; i++)
	{
		F4 c = f4(b3UnitSphere162[i][0], b3UnitSphere162[i][1], b3UnitSphere162[i][2], 0.f);
		if (dot3(dc, c) > 0) c = mul4(c, -1.f);
		float d;
		if (!test_sep_axis(S, ha, hb, posA, ornA, posB, ornB, c, &d)) return false;
		if (d < *dmin) { *dmin = d; *sep = c; }
	}
	if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
	return true;
}

/* ---------------- clipping (satClipHullContacts.cl) ---------------- */
#define CLIP_CAP 64
static F4 lerp3(F4 a, F4 b, float t)
{
	return f4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, 0.f);
}
/* clipFace (satClipHullContacts.cl:235-282); output capped at CLIP_CAP */
static int clip_face(const F4* in, int nIn, F4 n, float eq, F4* out)
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

/* clipHullAgainstHull + clipFaceAgainstHull (satClipHullContacts.cl:285-382,488-541) */
static int clip_hull_hull(const OShapes& S, F4 sep, const b3hConvexPolyhedronData& ha, const b3hConvexPolyhedronData& hb,
						  F4 posA, F4 ornA, F4 posB, F4 ornB, F4* contactsOut)
{
	const float minDist = -1e30f, maxDist = 0.02f;
	F4 bufA[CLIP_CAP], bufB[CLIP_CAP];
	int closestB = -1;
	float dmax = -FLT_MAX;
	for (int f = 0; f < hb.m_numFaces; f++)
	{
		F4 nrm = xyz0(S.faces[hb.m_faceOffset + f].m_plane);
		F4 wn = qrot(ornB, nrm);
		float d = dot3(wn, sep);
		if (d > dmax) { dmax = d; closestB = f; }
	}
	if (closestB < 0) return 0;
	int nB = 0;
	{
		const b3hGpuFace& poly = S.faces[hb.m_faceOffset + closestB];
		for (int e = 0; e < poly.m_numIndices && nB < CLIP_CAP; e++)
		{
			F4 v = S.verts[hb.m_vertexOffset + S.indices[poly.m_indexOffset + e]];
			bufA[nB++] = xform(v, posB, ornB);
		}
	}
	/* clipFaceAgainstHull */
	F4* pin = bufA;
	F4* pout = bufB;
	int nIn = nB;
	int closestA = -1;
	float dmin = FLT_MAX;
	for (int f = 0; f < ha.m_numFaces; f++)
	{
		F4 nrm = xyz0(S.faces[ha.m_faceOffset + f].m_plane);
		F4 wn = qrot(ornA, nrm);
		float d = dot3(wn, sep);
		if (d < dmin) { dmin = d; closestA = f; }
	}
	if (closestA < 0) return 0;
	const b3hGpuFace& polyA = S.faces[ha.m_faceOffset + closestA];
	int nvA = polyA.m_numIndices;
	for (int e0 = 0; e0 < nvA; e0++)
	{
		F4 a = S.verts[ha.m_vertexOffset + S.indices[polyA.m_indexOffset + e0]];
		F4 b = S.verts[ha.m_vertexOffset + S.indices[polyA.m_indexOffset + ((e0 + 1) % nvA)]];
		F4 edge0 = sub4(a, b);
		F4 we0 = qrot(ornA, edge0);
		F4 pnA = xyz0(polyA.m_plane);
		F4 wpn = qrot(ornA, pnA);
		F4 pn = neg4(cross3(we0, wpn));
		F4 wa = xform(a, posA, ornA);
		float peq = -dot3(wa, pn);
		int nOut = clip_face(pin, nIn, pn, peq, pout);
		F4* t = pout; pout = pin; pin = t;
		nIn = nOut;
	}
	int nc = 0;
	{
		F4 lpn = xyz0(polyA.m_plane);
		float leq = polyA.m_plane.w;
		F4 pn = qrot(ornA, lpn);
		float peq = leq - dot3(pn, posA);
		for (int i = 0; i < nIn; i++)
		{
			float depth = dot3(pn, pin[i]) + peq;
			if (depth <= minDist) depth = minDist;
			if (depth <= maxDist)
			{
				F4 p = pin[i];
				contactsOut[nc++] = f4(p.x, p.y, p.z, depth);
			}
		}
	}
	return nc;
}

/* extractManifoldSequential (satClipHullContacts.cl:720-810) */
static int extract_manifold(const F4* p, int n, F4 nearNormal, int* idx)
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
	int best[4];
	for (int ie = 0; ie < n; ie++)
	{
		F4 r = sub4(p[ie], center);
		float f[4] = {dot3(u, r), dot3(neg4(u), r), dot3(v, r), dot3(neg4(v), r)};
		for (int k = 0; k < 4; k++)
		{
			int a = (int)((f2u(f[k]) & 0xffffff00u) | (0xffu & (unsigned)ie));
			if (ie == 0 || a > best[k]) best[k] = a;
		}
	}
	for (int k = 0; k < 4; k++) idx[k] = best[k] & 0xff;
	return 4;
}

/* MPR stage alone (mprPenetrationKernel, mpr.cl:14-90).  Returns the number of
 * contacts appended; used for the golden-vector check. */
extern "C" int b3o_mpr_contacts(b3hInt4* pairs, int nPairs, const b3hRigidBodyData* bodies,
								 const b3hCollidable* cols, const b3hConvexPolyhedronData* convexes,
								 const b3hFloat4* vertices, b3hContact4Data* contacts, int capacity)
{
	OShapes S = {bodies, cols, convexes, vertices, 0, 0, 0};
	int nc = 0;
	for (int i = 0; i < nPairs; i++)
	{
		int a = pairs[i].x, b = pairs[i].y;
		if (!pair_is_hull_hull(S, a, b)) continue;
		float depth;
		F4 dir, pos;
		if (mpr_penetration(S, a, b, &depth, &dir, &pos) == 0)
		{
			int slot = nc++;
			if (slot < capacity)
			{
				pairs[i].z = slot;
				b3hContact4Data* c = contacts + slot;
				memset(c, 0, sizeof(*c));
				c->m_worldNormalOnB = neg4(dir);
				write_contact_header(c, S, pairs[i], i, 0.7f);
				pos.w = -depth;
				c->m_worldPosB[0] = pos;
				c->m_worldNormalOnB.w = 1.f;
			}
		}
	}
	return nc;
}

/* ---------------- primitive pairs (primitiveContactsKernel,
 * primitiveContacts.cl:793-977) ---------------- */
static F4 p_transform(F4 p, F4 t, F4 q) { return add4(qrot(q, p), t); }                 /* :230-234 */
static void p_tr_inverse(F4 t, F4 q, F4* to, F4* qo) { *qo = qinv(q); *to = qrot(*qo, neg4(t)); } /* :236-241 */
static void p_tr_mul(F4 tA, F4 qA, F4 tB, F4 qB, F4* to, F4* qo)                       /* :243-251 */
{
	*qo = qmul(qA, qB);
	*to = p_transform(tB, tA, qA);
}
static void p_header(b3hContact4Data* c, const OShapes& S, int pairIndex, int bA, int bB)
{
	b3hInt4 pr = {bA, bB, 0, 0};
	memset(c, 0, sizeof(*c));
	write_contact_header(c, S, pr, pairIndex, 0.7f);
}
/* extractManifoldSequential, :489-582 */
static int p_extract_manifold(const F4* p, int n, F4 nearNormal, int* idx)
{
	if (n > 64) n = 64;
	F4 center = f4(0, 0, 0, 0);
	for (int i = 0; i < n; i++) center = add4(center, p[i]);
	float fn = (float)n;
	center = f4(center.x / fn, center.y / fn, center.z / fn, center.w / fn);
	F4 aVector = sub4(p[0], center);
	F4 u = cross3(nearNormal, aVector);
	F4 v = cross3(nearNormal, u);
	u = nrm3(u);
	v = nrm3(v);
	float minW = FLT_MAX;
	int minIndex = -1;
	float maxDots[4] = {FLT_MIN, FLT_MIN, FLT_MIN, FLT_MIN};
	for (int ie = 0; ie < n; ie++)
	{
		if (p[ie].w < minW)
		{
			minW = p[ie].w;
			minIndex = ie;
		}
		F4 r = sub4(p[ie], center);
		float f = dot3(u, r);
		if (f < maxDots[0]) { maxDots[0] = f; idx[0] = ie; }
		f = dot3(neg4(u), r);
		if (f < maxDots[1]) { maxDots[1] = f; idx[1] = ie; }
		f = dot3(v, r);
		if (f < maxDots[2]) { maxDots[2] = f; idx[2] = ie; }
		f = dot3(neg4(v), r);
		if (f < maxDots[3]) { maxDots[3] = f; idx[3] = ie; }
	}
	if (idx[0] != minIndex && idx[1] != minIndex && idx[2] != minIndex && idx[3] != minIndex) idx[0] = minIndex;
	return 4;
}
/* the hull side of a primitive contact: its collidable and pose (the body's
 * own, or a compound child's: processCompoundPairsPrimitivesKernel, :980-1090) */
struct HullPose
{
	int col;
	F4 m_pos, m_quat;
};
static HullPose body_hull(const OShapes& S, int b)
{
	return HullPose{S.bodies[b].m_collidableIdx, S.bodies[b].m_pos, S.bodies[b].m_quat};
}
/* computeContactPlaneConvex, :584-727 (contact bodies: plane, hull) */
static bool p_plane_convex(const OShapes& S, int pairIndex, int bA, int bB, b3hContact4Data* c, const HullPose& B)
{
	const b3hRigidBodyData& A = S.bodies[bA];
	const b3hConvexPolyhedronData& hullB = S.convexes[S.cols[B.col].m_shapeIndex];
	F4 planeEq = S.faces[S.cols[A.m_collidableIdx].m_shapeIndex].m_plane;
	F4 planeNormal = f4(planeEq.x, planeEq.y, planeEq.z, 0.f);
	F4 planeNormalWorld = qrot(A.m_quat, planeNormal);
	float planeConstant = planeEq.w;
	F4 invPosA, invOrnA, invPosB, invOrnB, pic, qic;
	p_tr_inverse(A.m_pos, A.m_quat, &invPosA, &invOrnA);
	p_tr_inverse(B.m_pos, B.m_quat, &invPosB, &invOrnB);
	p_tr_mul(invPosB, invOrnB, A.m_pos, A.m_quat, &pic, &qic);
	F4 nInConvex = qrot(qic, neg4(planeNormal));
	float maxDot = -1e30f;
	F4 pts[64];
	int numPoints = 0;
	int idx[4] = {0, 1, 2, 3};
	for (int i = 0; i < hullB.m_numVertices; i++)
	{
		F4 vtx = S.verts[hullB.m_vertexOffset + i];
		float curDot = dot4(vtx, nInConvex);
		if (curDot > maxDot)
		{
			maxDot = curDot;
			if (numPoints == 64) numPoints--;
		}
		if (numPoints < 64)
		{
			F4 vtxWorld = p_transform(vtx, B.m_pos, B.m_quat);
			F4 vtxInPlane = p_transform(vtxWorld, invPosA, invOrnA);
			float dist = dot4(planeNormal, vtxInPlane) - planeConstant;
			if (dist < 0.f)
			{
				vtxWorld.w = dist;
				pts[numPoints++] = vtxWorld;
			}
		}
	}
	int nr = numPoints;
	if (numPoints > 4) nr = p_extract_manifold(pts, numPoints, nInConvex, idx);
	if (nr <= 0) return false;
	p_header(c, S, pairIndex, bA, bB);
	c->m_worldNormalOnB = neg4(planeNormalWorld);
	switch (nr)
	{
		case 4: c->m_worldPosB[3] = pts[idx[3]]; /* fall through */
		case 3: c->m_worldPosB[2] = pts[idx[2]]; /* fall through */
		case 2: c->m_worldPosB[1] = pts[idx[1]]; /* fall through */
		case 1: c->m_worldPosB[0] = pts[idx[0]];
	}
	c->m_worldNormalOnB.w = (float)nr;
	return true;
}
/* computeContactPlaneSphere, :729-791 (contact bodies: plane, sphere) */
static bool p_plane_sphere(const OShapes& S, int pairIndex, int bA, int bB, b3hContact4Data* c)
{
	const b3hRigidBodyData& A = S.bodies[bA];
	const b3hRigidBodyData& B = S.bodies[bB];
	F4 planeEq = S.faces[S.cols[A.m_collidableIdx].m_shapeIndex].m_plane;
	float radius = S.cols[B.m_collidableIdx].m_radius;
	F4 n1 = f4(planeEq.x, planeEq.y, planeEq.z, 0.f);
	float planeConstant = planeEq.w;
	F4 invPosA, invOrnA, cpPos, cpOrn, invPosB, invOrnB, pcPos, pcOrn;
	p_tr_inverse(A.m_pos, A.m_quat, &invPosA, &invOrnA);
	p_tr_mul(invPosA, invOrnA, B.m_pos, B.m_quat, &cpPos, &cpOrn);
	p_tr_inverse(B.m_pos, B.m_quat, &invPosB, &invOrnB);
	p_tr_mul(invPosB, invOrnB, A.m_pos, A.m_quat, &pcPos, &pcOrn);
	F4 vtx1 = mul4(qrot(pcOrn, neg4(n1)), radius);
	F4 vtxInPlane1 = p_transform(vtx1, cpPos, cpOrn);
	float distance = dot3(n1, vtxInPlane1) - planeConstant;
	if (!(distance < 0.f)) return false;
	F4 projected = sub4(vtxInPlane1, smul4(distance, n1));
	F4 world = p_transform(projected, A.m_pos, A.m_quat);
	F4 normalOnSurfaceB1 = qrot(A.m_quat, n1);
	F4 pOnB1 = add4(world, mul4(normalOnSurfaceB1, distance));
	pOnB1.w = distance;
	p_header(c, S, pairIndex, bA, bB);
	c->m_worldNormalOnB = neg4(normalOnSurfaceB1);
	c->m_worldPosB[0] = pOnB1;
	c->m_worldNormalOnB.w = 1.f;
	return true;
}
/* IsPointInPolygon, :280-337 (out.w on an edge's interior: 0; the reference leaves it unset) */
static bool p_point_in_polygon(F4 p, const b3hGpuFace& face, const F4* base, const int* indices, F4* out)
{
	F4 plane = f4(face.m_plane.x, face.m_plane.y, face.m_plane.z, 0.f);
	if (face.m_numIndices < 2) return false;
	F4 b = base[indices[face.m_indexOffset + face.m_numIndices - 1]];
	for (int i = 0; i < face.m_numIndices; i++)
	{
		F4 a = b;
		b = base[indices[face.m_indexOffset + i]];
		F4 ab = sub4(b, a), ap = sub4(p, a);
		F4 v = cross3(ab, plane);
		if (dot4(ap, v) > 0.f)
		{
			float ab_m2 = dot4(ab, ab);
			float rt = ab_m2 != 0.f ? dot4(ab, ap) / ab_m2 : 0.f;
			if (rt <= 0.f)
				*out = a;
			else if (rt >= 1.f)
				*out = b;
			else
			{
				float sv = 1.f - rt;
				*out = f4(sv * a.x + rt * b.x, sv * a.y + rt * b.y, sv * a.z + rt * b.z, 0.f);
			}
			return false;
		}
	}
	return true;
}
/* computeContactSphereConvex, :339-487 (contact bodies: sphere, hull) */
static bool p_sphere_convex(const OShapes& S, int pairIndex, int bA, int bB, b3hContact4Data* c, const HullPose& H)
{
	const b3hRigidBodyData& Sp = S.bodies[bA];
	float radius = S.cols[Sp.m_collidableIdx].m_radius;
	F4 pos = H.m_pos, quat = H.m_quat;
	F4 invPos, invOrn;
	p_tr_inverse(pos, quat, &invPos, &invOrn);
	F4 spherePos = p_transform(Sp.m_pos, invPos, invOrn);
	const b3hConvexPolyhedronData& hull = S.convexes[S.cols[H.col].m_shapeIndex];
	F4 closestPnt = f4(0, 0, 0, 0), hitNormalWorld = f4(0, 0, 0, 0);
	float minDist = -1000000.f;
	bool bCollide = true;
	for (int f = 0; f < hull.m_numFaces; f++)
	{
		const b3hGpuFace& face = S.faces[hull.m_faceOffset + f];
		F4 planeEqn = face.m_plane;
		F4 n = f4(planeEqn.x, planeEqn.y, planeEqn.z, 0.f);
		/* signedDistanceFromPointToPlane, :270-278 */
		float dist = dot3(n, spherePos) + planeEqn.w;
		F4 pntReturn = sub4(spherePos, smul4(dist, n));
		if (dist > radius)
		{
			bCollide = false;
			break;
		}
		if (dist > 0)
		{
			F4 out = f4(0, 0, 0, 0);
			if (p_point_in_polygon(spherePos, face, S.verts + hull.m_vertexOffset, S.indices, &out))
			{
				if (dist > minDist)
				{
					minDist = dist;
					closestPnt = pntReturn;
					hitNormalWorld = planeEqn;
				}
			}
			else
			{
				F4 tmp = sub4(spherePos, out);
				float l2 = dot4(tmp, tmp);
				if (l2 < radius * radius)
				{
					dist = sqrtf(l2);
					if (dist > minDist)
					{
						minDist = dist;
						closestPnt = out;
						hitNormalWorld = f4(tmp.x / dist, tmp.y / dist, tmp.z / dist, tmp.w / dist);
					}
				}
				else
				{
					bCollide = false;
					break;
				}
			}
		}
		else
		{
			if (dist > minDist)
			{
				minDist = dist;
				closestPnt = pntReturn;
				hitNormalWorld.x = planeEqn.x;
				hitNormalWorld.y = planeEqn.y;
				hitNormalWorld.z = planeEqn.z;
			}
		}
	}
	if (!(bCollide && minDist > -10000)) return false;
	F4 normalOnSurfaceB1 = qrot(quat, neg4(hitNormalWorld));
	F4 pOnB1 = p_transform(closestPnt, pos, quat);
	float actualDepth = minDist - radius;
	if (!(actualDepth <= 0.f)) return false;
	pOnB1.w = actualDepth;
	p_header(c, S, pairIndex, bA, bB);
	c->m_worldNormalOnB = neg4(normalOnSurfaceB1);
	c->m_worldPosB[0] = pOnB1;
	c->m_worldNormalOnB.w = 1.f;
	return true;
}
/* sphere-sphere, primitiveContactsKernel :905-951 */
static bool p_sphere_sphere(const OShapes& S, int pairIndex, int bA, int bB, b3hContact4Data* c)
{
	float radiusA = S.cols[S.bodies[bA].m_collidableIdx].m_radius;
	float radiusB = S.cols[S.bodies[bB].m_collidableIdx].m_radius;
	F4 posA = S.bodies[bA].m_pos, posB = S.bodies[bB].m_pos;
	F4 diff = sub4(posA, posB);
	float len = sqrtf(dot4(diff, diff));
	if (!(len <= radiusA + radiusB)) return false;
	float dist = len - (radiusA + radiusB);
	F4 normalOnSurfaceB = f4(1.f, 0.f, 0.f, 0.f);
	if ((double)len > 0.00001) normalOnSurfaceB = f4(diff.x / len, diff.y / len, diff.z / len, diff.w / len);
	F4 contactPosB = add4(posB, mul4(normalOnSurfaceB, radiusB));
	contactPosB.w = dist;
	p_header(c, S, pairIndex, bA, bB);
	c->m_worldNormalOnB = normalOnSurfaceB;
	c->m_worldPosB[0] = contactPosB;
	c->m_worldNormalOnB.w = 1.f;
	return true;
}
/* the primitive kernel's dispatch by shape types; false for hull x hull */
static bool p_primitive_pair(const OShapes& S, const b3hInt4& pr, int i, b3hContact4Data* c, bool* handled)
{
	int ta = S.cols[S.bodies[pr.x].m_collidableIdx].m_shapeType;
	int tb = S.cols[S.bodies[pr.y].m_collidableIdx].m_shapeType;
	*handled = !(ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_CONVEX_HULL);
	if (!*handled) return false;
	if (ta == B3H_SHAPE_PLANE && tb == B3H_SHAPE_CONVEX_HULL) return p_plane_convex(S, i, pr.x, pr.y, c, body_hull(S, pr.y));
	if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_PLANE) return p_plane_convex(S, i, pr.y, pr.x, c, body_hull(S, pr.x));
	if (ta == B3H_SHAPE_PLANE && tb == B3H_SHAPE_SPHERE) return p_plane_sphere(S, i, pr.x, pr.y, c);
	if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_PLANE) return p_plane_sphere(S, i, pr.y, pr.x, c);
	if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_CONVEX_HULL) return p_sphere_convex(S, i, pr.x, pr.y, c, body_hull(S, pr.y));
	if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_SPHERE) return p_sphere_convex(S, i, pr.y, pr.x, c, body_hull(S, pr.x));
	if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_SPHERE) return p_sphere_sphere(S, i, pr.x, pr.y, c);
	return false;
}

/* One pair of the convex-convex narrowphase: MPR, SAT over faces and edges,
 * clipping and reduction.  Writes the pair's manifold into *c (slot-free: the
 * header names the pair) and returns whether the pair makes a contact;
 * state / sep receive the final SAT state and normal. */
static bool narrow_pair(const OShapes& S, const b3hInt4& pr, int i, b3hContact4Data* c, int* stateOut, F4* sepOut)
{
	const b3hRigidBodyData* bodies = S.bodies;
	int a = pr.x, b = pr.y;
	int state = 0;
	F4 sep = f4(0, 0, 0, 0);
	{
		bool handled;
		bool hit = p_primitive_pair(S, pr, i, c, &handled);
		if (handled)
		{
			*stateOut = 0;
			*sepOut = sep;
			return hit;
		}
	}
	bool hull = pair_is_hull_hull(S, a, b);
	/* MPR */
	bool mprHit = false;
	float mprDepth = 0.f;
	F4 mprDir = f4(0, 0, 0, 0), mprPos = f4(0, 0, 0, 0);
	if (hull) mprHit = mpr_penetration(S, a, b, &mprDepth, &mprDir, &mprPos) == 0;
	/* the separating-normal slot starts with MPR's -dir when MPR found a
	 * penetration (b3MprPenetration.h:857); it only survives in the
	 * degenerate "state 4" path below */
	if (mprHit) sep = neg4(mprDir);
	if (hull)
	{
		const b3hRigidBodyData& A = bodies[a];
		const b3hRigidBodyData& B = bodies[b];
		const b3hConvexPolyhedronData& ha = S.hullOf(a);
		const b3hConvexPolyhedronData& hb = S.hullOf(b);
		F4 posA = A.m_pos; posA.w = 0.f;
		F4 posB = B.m_pos; posB.w = 0.f;
		F4 c0 = xform(ha.m_localCenter, posA, A.m_quat);
		F4 c1 = xform(hb.m_localCenter, posB, B.m_quat);
		F4 dc = sub4(c0, c1);
		/* findSeparatingAxisVertexFaceKernel (sat.cl:1417-1509) */
		float dmin = FLT_MAX;
		F4 sepN = sep;
		bool sepA = sat_faces(S, ha, hb, posA, A.m_quat, posB, B.m_quat, dc, &sepN, &dmin);
		float dminSaved = FLT_MAX;
		if (!sepA)
			state = 0;
		else
		{
			bool sepB = sat_faces(S, hb, ha, posB, B.m_quat, posA, A.m_quat, dc, &sepN, &dmin);
			if (sepB)
			{
				state = 1;
				dminSaved = dmin;
				sep = sepN;
			}
			else
				state = 4; /* sat.cl:1490-1506: flag left at 4, normal slot untouched */
		}
		int nEE = ha.m_numUniqueEdges * hb.m_numUniqueEdges;
		if (state != 0)
		{
			float dm = dminSaved;
			F4 s2 = sep;
			bool ok;
			if (nEE <= B3_UNIT_SPHERE_162_COUNT)
				ok = sat_edges(S, ha, hb, posA, A.m_quat, posB, B.m_quat, dc, &s2, &dm);
			else
				ok = sat_sphere(S, ha, hb, posA, A.m_quat, posB, B.m_quat, dc, &s2, &dm);
			if (!ok)
				state = 0;
			else
			{
				state = 1;
				sep = s2;
			}
		}
	}
	*stateOut = state;
	*sepOut = sep;

	/* clipHullHullKernel (satClipHullContacts.cl:890-1000) */
	int nReduced = 0;
	int idx[4] = {-1, -1, -1, -1};
	F4 pts[CLIP_CAP];
	if (hull && state == 1)
	{
		const b3hRigidBodyData& A = bodies[a];
		const b3hRigidBodyData& B = bodies[b];
		int np = clip_hull_hull(S, sep, S.hullOf(a), S.hullOf(b), A.m_pos, A.m_quat, B.m_pos, B.m_quat, pts);
		if (np > 0) nReduced = extract_manifold(pts, np, neg4(sep), idx);
	}
	if (!mprHit && nReduced == 0) return false;
	memset(c, 0, sizeof(*c));
	write_contact_header(c, S, pr, i, 0.7f);
	if (mprHit)
	{
		c->m_worldNormalOnB = neg4(mprDir);
		F4 p = mprPos;
		p.w = -mprDepth;
		c->m_worldPosB[0] = p;
		c->m_worldNormalOnB.w = 1.f;
	}
	if (nReduced > 0)
	{
		c->m_worldNormalOnB = sep;
		for (int k = 0; k < nReduced; k++)
			if (k > 0 || !mprHit) c->m_worldPosB[k] = pts[idx[k]];
		c->m_worldNormalOnB.w = (float)nReduced;
	}
	return true;
}

/* Full convex-convex narrowphase of one step: pairs in order, contacts take
 * slots in pair order.  With more than one oracle thread the pairs are
 * evaluated in parallel into a per-pair buffer and compacted in pair order
 * afterwards, so the output is the same bytes.  hasSepOut/sepOut (optional)
 * receive the final SAT state and normal of each pair. */
/* ---------------- compounds ---------------- */
/* b3TransformAabb2 (shared/b3Aabb.h:24-43), margin 0 */
static void o_transform_aabb2(F4 lmin, F4 lmax, F4 pos, F4 q, F4* mn, F4* mx)
{
	F4 half = smul4(0.5f, sub4(lmax, lmin));
	half = add4(half, f4(0.f, 0.f, 0.f, 0.f));
	F4 center = smul4(0.5f, add4(lmax, lmin));
	F4 q2 = f4(q.x * q.x, q.y * q.y, q.z * q.z, 0.f);
	F4 r0, r1, r2;
	r0 = f4(1 - 2 * q2.y - 2 * q2.z, 2 * q.x * q.y - 2 * q.w * q.z, 2 * q.x * q.z + 2 * q.w * q.y, 0.f);
	r1 = f4(2 * q.x * q.y + 2 * q.w * q.z, 1 - 2 * q2.x - 2 * q2.z, 2 * q.y * q.z - 2 * q.w * q.x, 0.f);
	r2 = f4(2 * q.x * q.z - 2 * q.w * q.y, 2 * q.y * q.z + 2 * q.w * q.x, 1 - 2 * q2.x - 2 * q2.y, 0.f);
	F4 a0 = f4(fabsf(r0.x), fabsf(r0.y), fabsf(r0.z), 0.f);
	F4 a1 = f4(fabsf(r1.x), fabsf(r1.y), fabsf(r1.z), 0.f);
	F4 a2 = f4(fabsf(r2.x), fabsf(r2.y), fabsf(r2.z), 0.f);
	F4 c = xform(center, pos, q);
	F4 ext = f4(dot3(half, a0), dot3(half, a1), dot3(half, a2), 0.f);
	*mn = sub4(c, ext);
	*mx = add4(c, ext);
}
/* b3TestAabbAgainstAabb (shared/b3Aabb.h:46-54) */
static bool o_aabb_overlap(F4 mn1, F4 mx1, F4 mn2, F4 mx2)
{
	bool overlap = true;
	overlap = (mn1.x > mx2.x || mx1.x < mn2.x) ? false : overlap;
	overlap = (mn1.z > mx2.z || mx1.z < mn2.z) ? false : overlap;
	overlap = (mn1.y > mx2.y || mx1.y < mn2.y) ? false : overlap;
	return overlap;
}
/* MyUnQuantize (sat.cl:926-940) */
static F4 o_unquantize(const unsigned short* q, F4 quant, F4 bmin)
{
	return add4(f4((float)q[0] / quant.x, (float)q[1] / quant.y, (float)q[2] / quant.z, 0.f), bmin);
}
static void o_node_box(const b3hBvhInfo& info, const unsigned short* qmn, const unsigned short* qmx, F4 pos, F4 orn,
					   F4* mn, F4* mx)
{
	o_transform_aabb2(o_unquantize(qmn, info.m_quantization, info.m_aabbMin),
					  o_unquantize(qmx, info.m_quantization, info.m_aabbMin), pos, orn, mn, mx);
}
/* findCompoundPairsKernel (sat.cl:945-1130): the child pairs of body pair
 * (a, b) appended to out, in the walk order of the reference */
static void compound_children(const OShapes& S, const b3oCompoundTables& T, int a, int b, std::vector<b3hInt4>& out)
{
	const b3hRigidBodyData& A = S.bodies[a];
	const b3hRigidBodyData& B = S.bodies[b];
	if (A.m_invMass == 0 && B.m_invMass == 0) return;
	const b3hCollidable& cA = S.cols[A.m_collidableIdx];
	const b3hCollidable& cB = S.cols[B.m_collidableIdx];
	const bool compA = cA.m_shapeType == B3H_SHAPE_COMPOUND_OF_CONVEX_HULLS;
	const bool compB = cB.m_shapeType == B3H_SHAPE_COMPOUND_OF_CONVEX_HULLS;
	if (compA && compB)
	{
		const b3hBvhInfo& iA = T.infos[cA.m_compoundBvhIndex];
		const b3hBvhInfo& iB = T.infos[cB.m_compoundBvhIndex];
		const int maxDepth = 256; /* B3_MAX_STACK_DEPTH, sat.cl:10 */
		for (int p = 0; p < iA.m_numSubTrees; p++)
		{
			const b3hBvhSubtreeInfo& sa = T.subtrees[iA.m_subTreeOffset + p];
			F4 aMn, aMx;
			o_node_box(iA, sa.m_quantizedAabbMin, sa.m_quantizedAabbMax, A.m_pos, A.m_quat, &aMn, &aMx);
			for (int q = 0; q < iB.m_numSubTrees; q++)
			{
				const b3hBvhSubtreeInfo& sb = T.subtrees[iB.m_subTreeOffset + q];
				F4 bMn, bMx;
				o_node_box(iB, sb.m_quantizedAabbMin, sb.m_quantizedAabbMax, B.m_pos, B.m_quat, &bMn, &bMx);
				if (!o_aabb_overlap(aMn, aMx, bMn, bMx)) continue;
				std::vector<std::pair<int, int>> stack;
				stack.push_back({sa.m_rootNodeIndex + iA.m_nodeOffset, sb.m_rootNodeIndex + iB.m_nodeOffset});
				while (!stack.empty())
				{
					std::pair<int, int> node = stack.back();
					stack.pop_back();
					const b3hQuantizedBvhNode& na = T.nodes[node.first];
					const b3hQuantizedBvhNode& nb = T.nodes[node.second];
					F4 mnA, mxA, mnB, mxB;
					o_node_box(iA, na.m_quantizedAabbMin, na.m_quantizedAabbMax, A.m_pos, A.m_quat, &mnA, &mxA);
					o_node_box(iB, nb.m_quantizedAabbMin, nb.m_quantizedAabbMax, B.m_pos, B.m_quat, &mnB, &mxB);
					if (!o_aabb_overlap(mnA, mxA, mnB, mxB)) continue;
					bool leafA = na.m_escapeIndexOrTriangleIndex >= 0, leafB = nb.m_escapeIndexOrTriangleIndex >= 0;
					if ((int)stack.size() + 4 > maxDepth && !(leafA && leafB)) continue;
					auto right = [&](int n) {
						const b3hQuantizedBvhNode& l = T.nodes[n + 1];
						return l.m_escapeIndexOrTriangleIndex >= 0 ? n + 2 : n + 1 - l.m_escapeIndexOrTriangleIndex;
					};
					if (!leafA)
					{
						int la = node.first + 1, ra = right(node.first);
						if (!leafB)
						{
							int lb = node.second + 1, rb = right(node.second);
							stack.push_back({la, lb});
							stack.push_back({ra, lb});
							stack.push_back({la, rb});
							stack.push_back({ra, rb});
						}
						else
						{
							stack.push_back({la, node.second});
							stack.push_back({ra, node.second});
						}
					}
					else if (!leafB)
					{
						int lb = node.second + 1, rb = right(node.second);
						stack.push_back({node.first, lb});
						stack.push_back({node.first, rb});
					}
					else
					{
						/* the leaf's child number (getTriangleIndexGlobal, sat.cl:67-73) */
						unsigned y = (~0u) << (31 - 10);
						out.push_back(b3hInt4{a, b, na.m_escapeIndexOrTriangleIndex & ~(int)y,
											  nb.m_escapeIndexOrTriangleIndex & ~(int)y});
					}
				}
			}
		}
	}
	else if (compA)
		for (int c = 0; c < cA.m_numChildShapes; c++) out.push_back(b3hInt4{a, b, cA.m_shapeIndex + c, -1});
	else if (cA.m_shapeType != B3H_SHAPE_CONCAVE_TRIMESH && compB)
		for (int c = 0; c < cB.m_numChildShapes; c++) out.push_back(b3hInt4{a, b, -1, cB.m_shapeIndex + c});
}
/* a child pair's side: the child's collidable and pose, or the body's
 * (processCompoundPairsKernel, sat.cl:830-857) */
static HullPose compound_side(const OShapes& S, const b3oCompoundTables& T, int body, int child)
{
	const b3hRigidBodyData& B = S.bodies[body];
	if (child < 0) return HullPose{B.m_collidableIdx, B.m_pos, B.m_quat};
	const b3hGpuChildShape& c = T.children[child];
	return HullPose{c.m_shapeIndex, add4(qrot(B.m_quat, c.m_childPosition), B.m_pos), qmul(B.m_quat, c.m_childOrientation)};
}
/* the contact of child pair k: processCompoundPairsKernel (sat.cl:802-925) with
 * clipCompoundsHullHullKernel (satClipHullContacts.cl:1003-1130) for two hulls,
 * processCompoundPairsPrimitivesKernel (primitiveContacts.cl:980-1090) for a
 * hull against a plane or a sphere */
static bool compound_pair(const OShapes& S, const b3oCompoundTables& T, const b3hInt4& cp, int k, b3hContact4Data* c)
{
	HullPose A = compound_side(S, T, cp.x, cp.z), B = compound_side(S, T, cp.y, cp.w);
	const int ta = S.cols[A.col].m_shapeType, tb = S.cols[B.col].m_shapeType;
	memset(c, 0, sizeof(*c));
	if (ta == B3H_SHAPE_PLANE && tb == B3H_SHAPE_CONVEX_HULL) return p_plane_convex(S, k, cp.x, cp.y, c, B);
	if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_PLANE) return p_plane_convex(S, k, cp.y, cp.x, c, A);
	if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_SPHERE) return p_sphere_convex(S, k, cp.y, cp.x, c, A);
	if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_CONVEX_HULL) return p_sphere_convex(S, k, cp.x, cp.y, c, B);
	if (ta != B3H_SHAPE_CONVEX_HULL || tb != B3H_SHAPE_CONVEX_HULL) return false;
	const b3hConvexPolyhedronData& ha = S.convexes[S.cols[A.col].m_shapeIndex];
	const b3hConvexPolyhedronData& hb = S.convexes[S.cols[B.col].m_shapeIndex];
	float dmin = FLT_MAX;
	F4 posA = A.m_pos, posB = B.m_pos;
	posA.w = 0.f;
	posB.w = 0.f;
	F4 c0 = xform(ha.m_localCenter, posA, A.m_quat);
	F4 c1 = xform(hb.m_localCenter, posB, B.m_quat);
	F4 dc = sub4(c0, c1);
	F4 sep = f4(1, 0, 0, 0);
	if (!sat_faces(S, ha, hb, A.m_pos, A.m_quat, B.m_pos, B.m_quat, dc, &sep, &dmin)) return false;
	if (!sat_faces(S, hb, ha, B.m_pos, B.m_quat, A.m_pos, A.m_quat, dc, &sep, &dmin)) return false;
	if (!sat_edges(S, ha, hb, A.m_pos, A.m_quat, B.m_pos, B.m_quat, dc, &sep, &dmin)) return false;
	F4 pts[CLIP_CAP];
	int np = clip_hull_hull(S, sep, ha, hb, A.m_pos, A.m_quat, B.m_pos, B.m_quat, pts);
	if (np <= 0) return false;
	int idx[4] = {-1, -1, -1, -1};
	int nr = extract_manifold(pts, np, neg4(sep), idx);
	c->m_worldNormalOnB = sep;
	c->m_restituitionCoeffCmp = (unsigned short)(0.f * 0xffff);
	c->m_frictionCoeffCmp = (unsigned short)(0.7f * 0xffff);
	c->m_batchIdx = k;
	c->m_bodyAPtrAndSignBit = S.bodies[cp.x].m_invMass == 0 ? -cp.x : cp.x;
	c->m_bodyBPtrAndSignBit = S.bodies[cp.y].m_invMass == 0 ? -cp.y : cp.y;
	c->m_childIndexA = cp.z;
	c->m_childIndexB = cp.w;
	for (int i = 0; i < nr; i++) c->m_worldPosB[i] = pts[idx[i]];
	c->m_worldNormalOnB.w = (float)nr;
	return true;
}

extern "C" int b3o_narrowphase_ex(b3hInt4* pairs, int nPairs, const b3hRigidBodyData* bodies, const b3hCollidable* cols,
								   const b3hConvexPolyhedronData* convexes, const b3hFloat4* vertices,
								   const b3hFloat4* uniqueEdges, const b3hGpuFace* faces, const int* indices,
								   b3hContact4Data* contacts, int capacity, int* hasSepOut, b3hFloat4* sepOut,
								   const b3oCompoundTables* compound, b3hInt4* childPairsOut, int* numChildPairs)
{
	int nc = b3o_narrowphase(pairs, nPairs, bodies, cols, convexes, vertices, uniqueEdges, faces, indices, contacts,
							 capacity, hasSepOut, sepOut);
	if (numChildPairs) *numChildPairs = 0;
	if (!compound || compound->capacity <= 0) return nc;
	OShapes S = {bodies, cols, convexes, vertices, uniqueEdges, faces, indices};
	const b3oCompoundTables& T = *compound;
	std::vector<b3hInt4> cps;
	for (int i = 0; i < nPairs; i++)
	{
		size_t k0 = cps.size();
		compound_children(S, T, pairs[i].x, pairs[i].y, cps);
		std::stable_sort(cps.begin() + (long)k0, cps.end(), [](const b3hInt4& u, const b3hInt4& v) {
			if ((unsigned)u.z != (unsigned)v.z) return (unsigned)u.z < (unsigned)v.z;
			return (unsigned)u.w < (unsigned)v.w;
		});
	}
	if (numChildPairs) *numChildPairs = (int)cps.size();
	int kept = (int)cps.size() < T.capacity ? (int)cps.size() : T.capacity;
	for (int k = 0; k < kept; k++)
	{
		if (childPairsOut) childPairsOut[k] = cps[(size_t)k];
		b3hContact4Data c;
		if (!compound_pair(S, T, cps[(size_t)k], k, &c)) continue;
		int slot = nc++;
		if (slot < capacity) contacts[slot] = c;
	}
	return nc;
}

extern "C" int b3o_narrowphase(b3hInt4* pairs, int nPairs, const b3hRigidBodyData* bodies,
								const b3hCollidable* cols, const b3hConvexPolyhedronData* convexes,
								const b3hFloat4* vertices, const b3hFloat4* uniqueEdges, const b3hGpuFace* faces,
								const int* indices, b3hContact4Data* contacts, int capacity,
								int* hasSepOut, b3hFloat4* sepOut)
{
	OShapes S = {bodies, cols, convexes, vertices, uniqueEdges, faces, indices};
	int nc = 0;
	int threads = b3o_threads();
	if (threads > 1 && nPairs > 1)
	{
		std::vector<b3hContact4Data> tmp(nPairs);
		std::vector<unsigned char> hit(nPairs);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
		for (int i = 0; i < nPairs; i++)
		{
			int state;
			F4 sep;
			hit[i] = narrow_pair(S, pairs[i], i, &tmp[i], &state, &sep) ? 1 : 0;
			if (hasSepOut) hasSepOut[i] = state;
			if (sepOut) sepOut[i] = sep;
		}
		for (int i = 0; i < nPairs; i++)
		{
			pairs[i].z = -1;
			if (!hit[i]) continue;
			int slot = nc++;
			if (slot >= capacity) continue;
			pairs[i].z = slot;
			contacts[slot] = tmp[i];
		}
		return nc;
	}
	for (int i = 0; i < nPairs; i++)
	{
		pairs[i].z = -1;
		int state;
		F4 sep;
		b3hContact4Data c;
		bool h = narrow_pair(S, pairs[i], i, &c, &state, &sep);
		if (hasSepOut) hasSepOut[i] = state;
		if (sepOut) sepOut[i] = sep;
		if (!h) continue;
		int slot = nc++;
		if (slot >= capacity) continue;
		pairs[i].z = slot;
		contacts[slot] = c;
	}
	return nc;
}
