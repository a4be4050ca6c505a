/*
 * k_narrowphase.hip -- convex-convex contact generation: MPR one pair per
 * thread, then the SAT and clipping in three passes (face families; edge
 * family of the pairs still open; clipping of the pairs still touching), each
 * pair handled by a group of L lanes of a wavefront, followed by an
 * order-preserving compaction of the contacts.
 *
 * Per pair (reference kernels, all in src/Bullet3OpenCL/NarrowphaseCollision/kernels/):
 *   1. MPR penetration + 1-point contact      mpr.cl:14-90 -> b3MprPenetration.h:816-886
 *   2. SAT over face normals of A then B      sat.cl:597-651, 1417-1509
 *   3. exact edge-edge SAT if |EA|*|EB| <= 162 sat.cl:719-798, 1512-1607
 *      else the 162-direction approximation   mpr.cl:180-311
 *   4. clip B's incident face against the side planes of A's reference face,
 *      keep depth <= 0.02, reduce to <= 4 points (satClipHullContacts.cl:235-382,
 *      488-541, extractManifoldSequential :720-810); point 0 stays the MPR point
 *      when MPR found one (clipHullHullKernel :890-1000).
 *
 * Work split:
 *   - MPR runs in its own kernel, one pair per thread (the portal iteration
 *     is serial; support queries use the sequential b3MaxDot);
 *   - per step, every body's world face normals with its own hull's interval
 *     on each (k_sat_faces) and its intervals on the 162 fixed directions
 *     (k_sat_dirs) are computed once, instead of once per pair the body is in;
 *   - the SAT passes stage both hulls' vertices in LDS (per-group slice); every
 *     SAT family spreads its axes over the lanes; a lane projects the hull(s)
 *     not covered by the per-body intervals on its axis in vertex order (so
 *     lo/hi are the reference's values), the family stops at the first round
 *     that finds a separating axis, and the minimum overlap is an arg-min with
 *     first-index tie-break, which equals the sequential "if (d < dmin)" scan;
 *   - each Sutherland-Hodgman pass gives one polygon edge per lane; lanes emit
 *     0, 1 or 2 points and a ballot prefix places them in sequential order.
 * The reduction orders therefore reproduce the sequential kernels exactly.
 * The reference runs these stages as 7 launches with a host sync after each;
 * here the passes hand pair lists to each other through prefix sums on the
 * device, each pair gets a pair-indexed contact record and a flag, and a
 * prefix sum over the flags assigns contact slots in pair order
 * (deterministic; the reference uses an atomic counter).
 */
#include "b3h_internal.h"
#include "unit_sphere_162.h"

#define MPR_TOL 1E-6f
#define MPR_MAX_ITER 1000
#define LARGE_FLOAT 1e18f
#define CLIP_CAP 64
#define NP_VCAP 64   /* hull vertices staged in LDS per hull; larger hulls are read from global memory */
#define NP_THREADS 256

__constant__ float c_unitSphere[162][3];
static_assert(B3_UNIT_SPHERE_162_COUNT == B3H_SAT_DIRS, "fixed-direction count");

/* LDS ordering between lanes of one wavefront (groups never span waves) */
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int L>
struct Group
{
	int lane;                 /* 0..L-1 */
	int base;                 /* first wavefront lane of the group */
	unsigned long long mask;  /* the group's lanes */
	unsigned long long below; /* the group's lanes below this one */
	__device__ Group()
	{
		int wl = threadIdx.x & 63;
		lane = wl & (L - 1);
		base = wl & ~(L - 1);
		mask = (L == 64) ? ~0ull : (((1ull << L) - 1ull) << base);
		below = ((1ull << wl) - 1ull) & mask;
	}
	__device__ __forceinline__ bool any(bool p) const { return (__ballot(p) & mask) != 0ull; }
	__device__ __forceinline__ int count(bool p) const { return __popcll(__ballot(p) & mask); }
	__device__ __forceinline__ int count_below(bool p) const { return __popcll(__ballot(p) & below); }
	__device__ __forceinline__ int src(int item) const { return base + (item & (L - 1)); }
	/* (v, i) with the largest v, ties -> smallest i; losers carry (-inf, INT_MAX) */
	__device__ __forceinline__ void argmax(float& v, int& i) const
	{
		for (int m = 1; m < L; m <<= 1)
		{
			float ov = __shfl_xor(v, m);
			int oi = __shfl_xor(i, m);
			if (ov > v || (ov == v && oi < i))
			{
				v = ov;
				i = oi;
			}
		}
	}
	/* (v, i) with the smallest v, ties -> smallest i; losers carry (+inf, INT_MAX) */
	__device__ __forceinline__ void argmin(float& v, int& i) const
	{
		for (int m = 1; m < L; m <<= 1)
		{
			float ov = __shfl_xor(v, m);
			int oi = __shfl_xor(i, m);
			if (ov < v || (ov == v && oi < i))
			{
				v = ov;
				i = oi;
			}
		}
	}
	__device__ __forceinline__ int imax(int v) const
	{
		for (int m = 1; m < L; m <<= 1)
		{
			int o = __shfl_xor(v, m);
			v = o > v ? o : v;
		}
		return v;
	}
	__device__ __forceinline__ F4 bcast(F4 v, int srcLane) const
	{
		return f4(__shfl(v.x, srcLane), __shfl(v.y, srcLane), __shfl(v.z, srcLane), __shfl(v.w, srcLane));
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

__device__ __forceinline__ void portal_dir(const Simplex* p, F4* dir)
{
	F4 v2v1 = sub4(p->ps[2].v, p->ps[1].v);
	F4 v3v1 = sub4(p->ps[3].v, p->ps[1].v);
	*dir = cross3(v2v1, v3v1);
	mpr_normalize(dir);
}
__device__ __forceinline__ int portal_reach_tolerance(const Simplex* p, const Sup* v4, F4 dir)
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
/* conditional copies as value selects (keeps the simplex in registers) */
__device__ __forceinline__ F4 sel4(bool c, F4 a, F4 b) { return f4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w); }
__device__ __forceinline__ void sup_set_if(Sup& dst, const Sup& src, bool c)
{
	dst.v = sel4(c, src.v, dst.v);
	dst.v1 = sel4(c, src.v1, dst.v1);
	dst.v2 = sel4(c, src.v2, dst.v2);
}
__device__ __forceinline__ void expand_portal(Simplex* p, const Sup* v4)
{
	F4 v4v0 = cross3(v4->v, p->ps[0].v);
	float d = dot3(p->ps[1].v, v4v0);
	int slot;
	if (d > 0.f)
	{
		d = dot3(p->ps[2].v, v4v0);
		slot = d > 0.f ? 1 : 3;
	}
	else
	{
		d = dot3(p->ps[3].v, v4v0);
		slot = d > 0.f ? 2 : 1;
	}
	sup_set_if(p->ps[1], *v4, slot == 1);
	sup_set_if(p->ps[2], *v4, slot == 2);
	sup_set_if(p->ps[3], *v4, slot == 3);
}
template <class SP>
__device__ __forceinline__ int discover_portal(SP& S, Simplex* p)
{
	F4 dir, va, vb;
	float d;
	p->ps[0].v1 = S.posA;
	p->ps[0].v2 = S.posB;
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
	S.support(dir, &p->ps[1]);
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
	S.support(dir, &p->ps[2]);
	d = dot3(p->ps[2].v, dir);
	if (mpr_is_zero(d) || d < 0.f) return -1;
	p->last = 2;
	va = sub4(p->ps[1].v, p->ps[0].v);
	vb = sub4(p->ps[2].v, p->ps[0].v);
	dir = cross3(va, vb);
	mpr_normalize(&dir);
	d = dot3(dir, p->ps[0].v);
	{
		bool swap = d > 0.f;
		Sup t = p->ps[1];
		sup_set_if(p->ps[1], p->ps[2], swap);
		sup_set_if(p->ps[2], t, swap);
		dir = sel4(swap, mul4(dir, -1.f), dir);
	}
	while (p->last + 1 < 4)
	{
		S.support(dir, &p->ps[3]);
		d = dot3(p->ps[3].v, dir);
		if (mpr_is_zero(d) || d < 0.f) return -1;
		int cont = 0;
		va = cross3(p->ps[1].v, p->ps[3].v);
		d = dot3(va, p->ps[0].v);
		bool c1 = d < 0.f && !mpr_is_zero(d);
		sup_set_if(p->ps[2], p->ps[3], c1);
		cont = c1 ? 1 : 0;
		if (!cont)
		{
			va = cross3(p->ps[3].v, p->ps[2].v);
			d = dot3(va, p->ps[0].v);
			bool c2 = d < 0.f && !mpr_is_zero(d);
			sup_set_if(p->ps[1], p->ps[3], c2);
			cont = c2 ? 1 : 0;
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
template <class SP>
__device__ __forceinline__ int refine_portal(SP& S, Simplex* p)
{
	F4 dir;
	Sup v4;
	for (int i = 0; i < MPR_MAX_ITER; i++)
	{
		portal_dir(p, &dir);
		float d = dot3(dir, p->ps[1].v);
		if (mpr_is_zero(d) || d > 0.f) return 0;
		S.support(dir, &v4);
		float dv4 = dot3(v4.v, dir);
		int canEnc = mpr_is_zero(dv4) || dv4 > 0.f;
		if (!canEnc || portal_reach_tolerance(p, &v4, dir)) return -1;
		expand_portal(p, &v4);
	}
	return -1;
}
__device__ __forceinline__ void find_pos(const Simplex* p, F4* pos)
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
	/* i = 0..3, written out so the simplex stays in registers */
	vec = mul4(p->ps[0].v1, b[0]);
	vadd3(&p1, vec);
	vec = mul4(p->ps[0].v2, b[0]);
	vadd3(&p2, vec);
	vec = mul4(p->ps[1].v1, b[1]);
	vadd3(&p1, vec);
	vec = mul4(p->ps[1].v2, b[1]);
	vadd3(&p2, vec);
	vec = mul4(p->ps[2].v1, b[2]);
	vadd3(&p1, vec);
	vec = mul4(p->ps[2].v2, b[2]);
	vadd3(&p2, vec);
	vec = mul4(p->ps[3].v1, b[3]);
	vadd3(&p1, vec);
	vec = mul4(p->ps[3].v2, b[3]);
	vadd3(&p2, vec);
	p1 = mul4(p1, inv);
	p2 = mul4(p2, inv);
	*pos = p1;
	vadd3(pos, p2);
	*pos = mul4(*pos, 0.5f);
}
__device__ __forceinline__ float point_segment_dist2(F4 P, F4 x0, F4 b, F4* witness)
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
__device__ __forceinline__ float point_tri_dist2(F4 P, F4 x0, F4 B, F4 Cc, F4* witness)
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
template <class SP>
__device__ __forceinline__ void find_penetr(SP& S, Simplex* p, float* depth, F4* pdir, F4* pos)
{
	F4 dir;
	Sup v4;
	unsigned iterations = 1u;
	F4 zero = f4(0, 0, 0, 0);
	for (int i = 0; i < MPR_MAX_ITER; i++)
	{
		portal_dir(p, &dir);
		S.support(dir, &v4);
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
template <class SP>
__device__ __forceinline__ int mpr_penetration(SP& S, float* depth, F4* dir, F4* pos)
{
	Simplex portal;
	int res = discover_portal(S, &portal);
	if (res == 0)
	{
		if (refine_portal(S, &portal) < 0) return -1;
		find_penetr(S, &portal, depth, dir, pos);
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
/* project (sat.cl): the interval of dot(vertex, localDir) over the hull's
 * vertices, shifted by dot(pos, dir).  Hull vertices are stored with w = 0, so
 * the w term of the reference's float4 dot adds a zero; it is left out, and
 * the interval ends are taken with min/max instead of compare-and-select.
 * Both can only change the sign of a zero end, and the ends are used in
 * comparisons and differences whose outcome a signed zero does not change
 * (the overlap test, the depth comparisons, the axis chosen). */
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) const F4 LdsF4;
#define LDS_PTR(p) ((LdsF4*)(p))
#else
#define LDS_PTR(p) (p)
#endif
/* the interval loop, two vertices per round with the loads issued first;
 * VP is an LDS or a global pointer so the loads are ds_read / global_load */
template <typename VP>
__device__ __forceinline__ float vdot3(VP vs, int i, F4 d)
{
	return vs[i].x * d.x + vs[i].y * d.y + vs[i].z * d.z;
}
template <typename VP>
__device__ __forceinline__ void project_loop(VP vs, int nv, F4 d, float& lo, float& hi)
{
	int i = 0;
	for (; i + 4 <= nv; i += 4)
	{
		float pa = vdot3(vs, i, d), pb = vdot3(vs, i + 1, d), pc = vdot3(vs, i + 2, d), pd = vdot3(vs, i + 3, d);
		lo = fminf(fminf(lo, fminf(pa, pb)), fminf(pc, pd));
		hi = fmaxf(fmaxf(hi, fmaxf(pa, pb)), fmaxf(pc, pd));
	}
	for (; i + 2 <= nv; i += 2)
	{
		float pa = vdot3(vs, i, d), pb = vdot3(vs, i + 1, d);
		lo = fminf(lo, fminf(pa, pb));
		hi = fmaxf(hi, fmaxf(pa, pb));
	}
	for (; i < nv; i++)
	{
		float pa = vdot3(vs, i, d);
		lo = fminf(lo, pa);
		hi = fmaxf(hi, pa);
	}
}
/* UNIFORM: every lane of the wavefront projects the same hull (one pair per
 * wavefront), so the vertex count is read as a scalar */
template <bool UNIFORM>
__device__ __forceinline__ void project(const F4* vs, bool inLds, int nv, F4 pos, F4 orn, F4 dir, float* mn,
										float* mx)
{
	float lo = FLT_MAX, hi = -FLT_MAX;
	F4 localDir = qinvrot(orn, dir);
	float offset = dot4(pos, dir);
	if (UNIFORM) nv = __builtin_amdgcn_readfirstlane(nv);
	if (inLds)
		project_loop(LDS_PTR(vs), nv, localDir, lo, hi);
	else
		project_loop(vs, nv, localDir, lo, hi);
	if (lo > hi)
	{
		float t = lo;
		lo = hi;
		hi = t;
	}
	*mn = lo + offset;
	*mx = hi + offset;
}
struct PairGeom
{
	const F4* vA;
	const F4* vB;
	int nA, nB;
	F4 posA, ornA, posB, ornB;
	bool ldsA, ldsB; /* vA / vB point into LDS */
};
/* TestSepAxis: false if the axis separates, else the overlap */
template <bool UNIFORM>
__device__ __forceinline__ bool test_sep_axis(const PairGeom& P, F4 axis, float* depth)
{
	float mn0, mx0, mn1, mx1;
	project<UNIFORM>(P.vA, P.ldsA, P.nA, P.posA, P.ornA, axis, &mn0, &mx0);
	project<UNIFORM>(P.vB, P.ldsB, P.nB, P.posB, P.ornB, axis, &mn1, &mx1);
	if (mx0 < mn1 || mx1 < mn0) return false;
	float d0 = mx0 - mn1, d1 = mx1 - mn0;
	*depth = d0 < d1 ? d0 : d1;
	return true;
}
__device__ __forceinline__ bool almost_zero(F4 v) { return !(fabsf(v.x) > 1e-6f || fabsf(v.y) > 1e-6f || fabsf(v.z) > 1e-6f); }

/* The SAT families: kind 0: faces of A (tested with A as hull 0), kind 1:
 * faces of B (tested with B as hull 0, sat.cl:1463-1489), kind 2: edge
 * crosses, kind 3: the 162 fixed directions. */
struct Family
{
	int kind;
	int count;
	const b3hGpuFace* faces; /* kind 0/1 */
	const F4* edgesA;        /* kind 2 */
	const F4* edgesB;
	int nEB;
	const b3hSatFace* cache; /* kind 0/1: the face owner's records, or null */
	const float2* dA;        /* kind 3: A's and B's intervals on the fixed directions, or null */
	const float2* dB;
};
/* Edge or fixed-direction axis t; false when the family has no axis t
 * (parallel edges) */
__device__ __forceinline__ bool family_axis(const Family& F, const PairGeom& P, F4 dc, int t, F4* axis)
{
	if (F.kind == 2)
	{
		int e0 = t / F.nEB, e1 = t - e0 * F.nEB;
		F4 ew0 = qrot(P.ornA, F.edgesA[e0]);
		F4 ew1 = qrot(P.ornB, F.edgesB[e1]);
		F4 c = cross3(ew0, ew1);
		if (almost_zero(c)) return false;
		c = nrm3(c);
		if (dot3(dc, c) < 0) c = mul4(c, -1.f);
		*axis = c;
		return true;
	}
	F4 c = f4(c_unitSphere[t][0], c_unitSphere[t][1], c_unitSphere[t][2], 0.f);
	if (dot3(dc, c) > 0) c = mul4(c, -1.f);
	*axis = c;
	return true;
}
/* One SAT family over the group: false if some axis separates; otherwise the
 * sequential min-overlap update of (sep, dmin) and the final orientation flip.
 *
 * A face axis is the owner's world face normal w = qrot(orn, plane), turned
 * towards dc, and the owner's own interval on it.  Both come from the owner's
 * face record (k_sat_faces) when there is one, else they are computed here with
 * the same operations.  Turning the axis negates every product of the
 * projection, so the owner's interval on -w is (-mx, -mn) of its interval on w
 * (the same values up to the sign of a zero end, which no comparison or
 * difference downstream can see).  Only the other hull is projected per pair. */
template <int L>
__device__ bool sat_family(const Group<L>& g, const Family& F, const PairGeom& P, F4 dc, F4* sep, float* dmin)
{
	float bv = INFINITY;
	int bi = INT_MAX;
	F4 bn = f4(0, 0, 0, 0);
	const bool faceKind = F.kind <= 1;
	const bool ownA = F.kind == 0; /* the face owner is hull 0 of the overlap test */
	for (int t0 = 0; t0 < F.count; t0 += L)
	{
		int t = t0 + g.lane;
		bool separated = false;
		F4 axis;
		if (faceKind)
		{
			if (t < F.count)
			{
				b3hSatFace r;
				if (F.cache)
					r = F.cache[t];
				else
				{
					r.w = qrot(ownA ? P.ornA : P.ornB, F.faces[t].m_plane);
					project<L == 64>(ownA ? P.vA : P.vB, ownA ? P.ldsA : P.ldsB, ownA ? P.nA : P.nB,
									 ownA ? P.posA : P.posB, ownA ? P.ornA : P.ornB, r.w, &r.mn, &r.mx);
				}
				float mn0, mx0, mn1, mx1;
				if (dot3(dc, r.w) < 0)
				{
					axis = mul4(r.w, -1.f);
					mn0 = -r.mx;
					mx0 = -r.mn;
				}
				else
				{
					axis = r.w;
					mn0 = r.mn;
					mx0 = r.mx;
				}
				project<L == 64>(ownA ? P.vB : P.vA, ownA ? P.ldsB : P.ldsA, ownA ? P.nB : P.nA, ownA ? P.posB : P.posA,
								 ownA ? P.ornB : P.ornA, axis, &mn1, &mx1);
				if (mx0 < mn1 || mx1 < mn0)
					separated = true;
				else
				{
					float d0 = mx0 - mn1, d1 = mx1 - mn0;
					float d = d0 < d1 ? d0 : d1;
					if (d < bv)
					{
						bv = d;
						bi = t;
						bn = axis;
					}
				}
			}
		}
		else if (F.dA)
		{
			/* fixed directions with both hulls' intervals precomputed on the
			 * unturned direction; the turned axis negates and swaps them */
			if (t < F.count)
			{
				F4 c = f4(c_unitSphere[t][0], c_unitSphere[t][1], c_unitSphere[t][2], 0.f);
				const float2 ia = F.dA[t], ib = F.dB[t];
				float mn0 = ia.x, mx0 = ia.y, mn1 = ib.x, mx1 = ib.y;
				if (dot3(dc, c) > 0)
				{
					c = mul4(c, -1.f);
					mn0 = -ia.y;
					mx0 = -ia.x;
					mn1 = -ib.y;
					mx1 = -ib.x;
				}
				axis = c;
				if (mx0 < mn1 || mx1 < mn0)
					separated = true;
				else
				{
					float d0 = mx0 - mn1, d1 = mx1 - mn0;
					float d = d0 < d1 ? d0 : d1;
					if (d < bv)
					{
						bv = d;
						bi = t;
						bn = axis;
					}
				}
			}
		}
		else if (t < F.count && family_axis(F, P, dc, t, &axis))
		{
			float d;
			if (!test_sep_axis<L == 64>(P, axis, &d))
				separated = true;
			else if (d < bv)
			{
				bv = d;
				bi = t;
				bn = axis;
			}
		}
		if (g.any(separated)) return false;
	}
	g.argmin(bv, bi);
	F4 n = g.bcast(bn, g.src(bi == INT_MAX ? 0 : bi));
	if (bi != INT_MAX && bv < *dmin)
	{
		*dmin = bv;
		*sep = n;
	}
	if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
	return true;
}

/* ---------------- clipping ---------------- */
__device__ __forceinline__ F4 lerp3(F4 a, F4 b, float t)
{
	return f4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, 0.f);
}
/* one Sutherland-Hodgman pass (clipFace, satClipHullContacts.cl:235-290), one
 * polygon edge per lane, outputs placed in the sequential order */
template <int L>
__device__ int clip_pass(const Group<L>& g, const F4* in, int nIn, F4 n, float eq, F4* out)
{
	if (nIn < 2) return 0;
	int total = 0;
	for (int c0 = 0; c0 < nIn; c0 += L)
	{
		int j = c0 + g.lane;
		int cnt = 0;
		F4 o0 = f4(0, 0, 0, 0), o1 = f4(0, 0, 0, 0);
		if (j < nIn)
		{
			F4 first = in[j == 0 ? nIn - 1 : j - 1];
			F4 end = in[j];
			float ds = dot3(n, first) + eq;
			float de = dot3(n, end) + eq;
			if (ds < 0)
			{
				cnt = 1;
				o0 = (de < 0) ? end : lerp3(first, end, ds * 1.f / (ds - de));
			}
			else if (de < 0)
			{
				cnt = 2;
				o0 = lerp3(first, end, ds * 1.f / (ds - de));
				o1 = end;
			}
		}
		int pre = total + g.count_below(cnt >= 1) + g.count_below(cnt == 2);
		if (cnt >= 1 && pre < CLIP_CAP) out[pre] = o0;
		if (cnt == 2 && pre + 1 < CLIP_CAP) out[pre + 1] = o1;
		total += g.count(cnt >= 1) + g.count(cnt == 2);
	}
	wave_sync();
	return total < CLIP_CAP ? total : CLIP_CAP;
}
/* clipHullAgainstHull (satClipHullContacts.cl:488-541): returns the kept points
 * (xyz + depth) in `out` */
template <int L>
__device__ int clip_hull_hull(const Group<L>& g, F4 sep, const PairGeom& P, const b3hSatBody& ha, const b3hSatBody& hb,
							  const b3hShapeTables& sh, F4* bufA, F4* bufB, F4* out, const b3hSatFace* cA,
							  const b3hSatFace* cB)
{
	const float minDist = -1e30f, maxDist = 0.02f;
	/* incident face of B: largest dot with the separating normal */
	float bv = -INFINITY;
	int bi = INT_MAX;
	for (int f = g.lane; f < hb.nF; f += L)
	{
		F4 wn = cB ? cB[f].w : qrot(P.ornB, xyz0(sh.faces[hb.fOff + f].m_plane));
		float d = dot3(wn, sep);
		if (d > -FLT_MAX && (bi == INT_MAX || d > bv))
		{
			bv = d;
			bi = f;
		}
	}
	g.argmax(bv, bi);
	if (bi == INT_MAX) return 0;
	const b3hGpuFace polyB = sh.faces[hb.fOff + bi];
	int nB = polyB.m_numIndices < CLIP_CAP ? polyB.m_numIndices : CLIP_CAP;
	for (int e = g.lane; e < nB; e += L)
		bufA[e] = xform(P.vB[sh.indices[polyB.m_indexOffset + e]], P.posB, P.ornB);
	wave_sync();
	/* reference face of A: smallest dot with the separating normal */
	float av = INFINITY;
	int ai = INT_MAX;
	for (int f = g.lane; f < ha.nF; f += L)
	{
		F4 wn = cA ? cA[f].w : qrot(P.ornA, xyz0(sh.faces[ha.fOff + f].m_plane));
		float d = dot3(wn, sep);
		if (d < FLT_MAX && (ai == INT_MAX || d < av))
		{
			av = d;
			ai = f;
		}
	}
	g.argmin(av, ai);
	if (ai == INT_MAX) return 0;
	const b3hGpuFace polyA = sh.faces[ha.fOff + ai];
	int nvA = polyA.m_numIndices;
	const int* ia = sh.indices + polyA.m_indexOffset;
	F4 wpn = cA ? cA[ai].w : qrot(P.ornA, xyz0(polyA.m_plane));
	F4* pin = bufA;
	F4* pout = bufB;
	int nIn = nB;
	/* the side planes of the reference face, L at a time, one per lane; pass
	 * e0 takes plane e0 from its lane */
	F4 lanePn = f4(0, 0, 0, 0);
	float lanePeq = 0.f;
	for (int e0 = 0; e0 < nvA; e0++)
	{
		if ((e0 & (L - 1)) == 0)
		{
			int e = e0 + g.lane;
			if (e < nvA)
			{
				F4 a = P.vA[ia[e]];
				F4 b = P.vA[ia[(e + 1) % nvA]];
				F4 we0 = qrot(P.ornA, sub4(a, b));
				lanePn = neg4(cross3(we0, wpn));
				F4 wa = xform(a, P.posA, P.ornA);
				lanePeq = -dot3(wa, lanePn);
			}
		}
		F4 pn = g.bcast(lanePn, g.src(e0));
		float peq = __shfl(lanePeq, g.src(e0));
		int nOut = clip_pass(g, pin, nIn, pn, peq, pout);
		F4* t = pout;
		pout = pin;
		pin = t;
		nIn = nOut;
	}
	float peq = polyA.m_plane.w - dot3(wpn, P.posA);
	int nc = 0;
	for (int c0 = 0; c0 < nIn; c0 += L)
	{
		int i = c0 + g.lane;
		bool keep = false;
		F4 p = f4(0, 0, 0, 0);
		if (i < nIn)
		{
			p = pin[i];
			float depth = dot3(wpn, p) + peq;
			if (depth <= minDist) depth = minDist;
			keep = depth <= maxDist;
			p.w = depth;
		}
		int pre = nc + g.count_below(keep);
		if (keep) out[pre] = p;
		nc += g.count(keep);
	}
	wave_sync();
	return nc;
}
/* extractManifoldSequential (satClipHullContacts.cl:720-810); the centre is the
 * sequential sum (every lane runs it), the four extreme points come from an
 * integer max whose low byte is the point index (order-free) */
template <int L>
__device__ int extract_manifold(const Group<L>& g, const F4* p, int n, F4 nearNormal, int* idx)
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
	int b0 = INT_MIN, b1 = INT_MIN, b2 = INT_MIN, b3 = INT_MIN;
	for (int ie = g.lane; ie < n; ie += L)
	{
		F4 r = sub4(p[ie], center);
		unsigned lo = 0xffu & (unsigned)ie;
		int a0 = (int)((f2u(dot3(u, r)) & 0xffffff00u) | lo);
		int a1 = (int)((f2u(dot3(neg4(u), r)) & 0xffffff00u) | lo);
		int a2 = (int)((f2u(dot3(v, r)) & 0xffffff00u) | lo);
		int a3 = (int)((f2u(dot3(neg4(v), r)) & 0xffffff00u) | lo);
		b0 = a0 > b0 ? a0 : b0;
		b1 = a1 > b1 ? a1 : b1;
		b2 = a2 > b2 ? a2 : b2;
		b3 = a3 > b3 ? a3 : b3;
	}
	idx[0] = g.imax(b0) & 0xff;
	idx[1] = g.imax(b1) & 0xff;
	idx[2] = g.imax(b2) & 0xff;
	idx[3] = g.imax(b3) & 0xff;
	return 4;
}

/* MPR support of one pair for one thread: b3MprConvexSupport with the
 * sequential b3MaxDot (first maximum wins), vertices read from the (L1/L2
 * resident) hull table */
struct ThreadSupport
{
	const F4* vA;
	const F4* vB;
	int nA, nB;
	F4 posA, posB, qA, qB;
	__device__ __forceinline__ F4 one(const F4* vs, int nv, F4 pos, F4 q, F4 dirIn) const
	{
		F4 dir = f4(dirIn.x, dirIn.y, dirIn.z, 0.f);
		F4 localDir = qrot(qinv(q), dir);
		float best = -LARGE_FLOAT;
		int idx = -1;
		for (int i = 0; i < nv; i++)
		{
			float d = dot3(vs[i], localDir);
			if (d > best)
			{
				best = d;
				idx = i;
			}
		}
		if (idx < 0) idx = 0;
		return xform(vs[idx], pos, q);
	}
	__device__ __forceinline__ void support(F4 dir, Sup* s) const
	{
		s->v1 = one(vA, nA, posA, qA, dir);
		s->v2 = one(vB, nB, posB, qB, mul4(dir, -1.f));
		s->v = sub4(s->v1, s->v2);
	}
};

__device__ __forceinline__ bool pair_is_hull_hull(const b3hRigidBodyData& A, const b3hRigidBodyData& B,
												  const b3hShapeTables& sh)
{
	return !(A.m_invMass == 0 && B.m_invMass == 0) &&
		   sh.collidables[A.m_collidableIdx].m_shapeType == B3H_SHAPE_CONVEX_HULL &&
		   sh.collidables[B.m_collidableIdx].m_shapeType == B3H_SHAPE_CONVEX_HULL;
}

/* MPR penetration of every pair, one pair per thread (mprPenetrationKernel,
 * mpr.cl:14-90).  The portal iteration is inherently serial, so a thread per
 * pair keeps all lanes busy; the result (direction, point, depth, hit) is
 * picked up by the SAT/clipping kernel. */
/* four blocks of 256 per CU: 128 VGPRs, four waves per SIMD */
__global__ __launch_bounds__(256, 4) void k_np_mpr(const b3hInt4* __restrict__ pairs, const int* __restrict__ numPairsPtr,
												const b3hRigidBodyData* __restrict__ bodies, b3hShapeTables sh,
												F4* __restrict__ mprOut)
{
	int numPairs = *numPairsPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numPairs; i += gridDim.x * blockDim.x)
	{
		b3hInt4 pr = pairs[i];
		const b3hRigidBodyData A = bodies[pr.x];
		const b3hRigidBodyData B = bodies[pr.y];
		F4 dir = f4(0, 0, 0, 0), pos = f4(0, 0, 0, 0);
		float depth = 0.f;
		bool hit = false;
		if (pair_is_hull_hull(A, B, sh))
		{
			const b3hConvexPolyhedronData& ha = sh.convexes[sh.collidables[A.m_collidableIdx].m_shapeIndex];
			const b3hConvexPolyhedronData& hb = sh.convexes[sh.collidables[B.m_collidableIdx].m_shapeIndex];
			ThreadSupport S = {sh.vertices + ha.m_vertexOffset, sh.vertices + hb.m_vertexOffset, ha.m_numVertices,
							   hb.m_numVertices, A.m_pos, B.m_pos, A.m_quat, B.m_quat};
			hit = mpr_penetration(S, &depth, &dir, &pos) == 0;
		}
		mprOut[3 * (size_t)i] = dir;
		mprOut[3 * (size_t)i + 1] = pos;
		mprOut[3 * (size_t)i + 2] = f4(depth, hit ? 1.f : 0.f, 0.f, 0.f);
	}
}

/* ---------------- per-body data of the SAT ---------------- */
/* The descriptor of every body (thread per body).  The face-record base and
 * the fixed-direction slot are kept only when they cover the body's current
 * hull. */
__global__ __launch_bounds__(256) void k_sat_bodies(const b3hRigidBodyData* __restrict__ bodies, b3hShapeTables sh,
													b3hSatCache sc)
{
	for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < sc.numBodies; b += gridDim.x * blockDim.x)
	{
		const b3hRigidBodyData body = bodies[b];
		const b3hCollidable col = sh.collidables[body.m_collidableIdx];
		b3hSatBody d;
		d.pos = body.m_pos;
		d.quat = body.m_quat;
		d.center = f4(0, 0, 0, 0);
		d.vOff = d.nV = d.fOff = d.nF = d.eOff = d.nE = 0;
		d.rec = d.dir = -1;
		d.flags = body.m_invMass == 0 ? B3H_SB_STATIC : 0;
		d.col = body.m_collidableIdx;
		d.pad0 = d.pad1 = 0;
		if (col.m_shapeType == B3H_SHAPE_CONVEX_HULL)
		{
			const b3hConvexPolyhedronData h = sh.convexes[col.m_shapeIndex];
			d.flags |= B3H_SB_HULL;
			d.vOff = h.m_vertexOffset;
			d.nV = h.m_numVertices;
			d.fOff = h.m_faceOffset;
			d.nF = h.m_numFaces;
			d.eOff = h.m_uniqueEdgesOffset;
			d.nE = h.m_numUniqueEdges;
			F4 p = body.m_pos;
			p.w = 0.f;
			d.center = xform(h.m_localCenter, p, body.m_quat);
			if (sc.recs)
			{
				int s0 = sc.base[b], s1 = sc.base[b + 1];
				if (s1 - s0 == h.m_numFaces && s0 >= 0 && s1 <= sc.total) d.rec = s0;
			}
			if (sc.dirs)
			{
				int k = sc.dslot[b];
				if (k >= 0 && k < sc.numDirBodies) d.dir = k;
			}
		}
		sc.desc[b] = d;
	}
}
/* One record per thread: body owner[r], face r - base[owner[r]] of its hull:
 * the world face normal and the body's own interval on it.  Records of a body
 * whose hull changed on the device are not used (its descriptor has rec -1). */
__global__ __launch_bounds__(256) void k_sat_faces(b3hShapeTables sh, b3hSatCache sc)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < sc.total; r += gridDim.x * blockDim.x)
	{
		const b3hSatBody d = sc.desc[sc.owner[r]];
		if (d.rec < 0) continue;
		int f = r - d.rec;
		F4 pos = d.pos;
		pos.w = 0.f;
		b3hSatFace rec;
		rec.w = qrot(d.quat, sh.faces[d.fOff + f].m_plane);
		project<false>(sh.vertices + d.vOff, false, d.nV, pos, d.quat, rec.w, &rec.mn, &rec.mx);
		rec.pad0 = 0.f;
		rec.pad1 = 0.f;
		sc.recs[r] = rec;
	}
}
/* Intervals on the 162 fixed directions: items (body k, direction t) in
 * tiles of 256; the few bodies a tile touches have their vertices staged in
 * LDS first, and each thread projects its body on its direction. */
#define SD_TILE 256
#define SD_BODIES 3 /* bodies a tile of 256 items can touch (162 items each) */
__global__ __launch_bounds__(SD_TILE) void k_sat_dirs(b3hShapeTables sh, b3hSatCache sc)
{
	__shared__ F4 sV[SD_BODIES][NP_VCAP];
	const int total = sc.numDirBodies * B3_UNIT_SPHERE_162_COUNT;
	const int t = threadIdx.x;
	for (int base = blockIdx.x * SD_TILE; base < total; base += gridDim.x * SD_TILE)
	{
		const int k0 = base / B3_UNIT_SPHERE_162_COUNT;
		/* stage: thread t loads vertex t % 64 of the tile's body t / 64 */
		{
			int j = t / NP_VCAP, v = t % NP_VCAP, k = k0 + j;
			if (j < SD_BODIES && k < sc.numDirBodies)
			{
				const b3hSatBody d = sc.desc[sc.downer[k]];
				if (d.nV <= NP_VCAP && v < d.nV) sV[j][v] = sh.vertices[d.vOff + v];
			}
		}
		__syncthreads();
		int r = base + t;
		if (r < total)
		{
			int k = r / B3_UNIT_SPHERE_162_COUNT, dirIdx = r - k * B3_UNIT_SPHERE_162_COUNT;
			const b3hSatBody d = sc.desc[sc.downer[k]];
			if (d.flags & B3H_SB_HULL)
			{
				F4 pos = d.pos;
				pos.w = 0.f;
				F4 c = f4(c_unitSphere[dirIdx][0], c_unitSphere[dirIdx][1], c_unitSphere[dirIdx][2], 0.f);
				float mn, mx;
				bool inLds = d.nV <= NP_VCAP;
				project<false>(inLds ? sV[k - k0] : sh.vertices + d.vOff, inLds, d.nV, pos, d.quat, c, &mn, &mx);
				sc.dirs[r] = make_float2(mn, mx);
			}
		}
		__syncthreads();
	}
}

/* ---------------- the SAT passes ----------------
 * Three kernels over the pair list, each finishing the pairs it decides:
 *   faces:  all pairs; the face families of A and B.  A pair separated there,
 *           or not a hull pair, is finished (its contact is the MPR point
 *           alone, if MPR hit); the others keep (state, sep, dmin) for
 *   edges:  one pair per wavefront, pairs still open; the edge-pair or fixed-
 *           direction family.  Separated pairs are finished, the rest are
 *           flagged for clipping, and a prefix sum lists them in pair order for
 *   clip:   the listed pairs, several per wavefront; clipping and reduction.
 * Every pass stages both hulls' vertices in LDS.  Splitting the stages keeps
 * each wavefront on one kind of work: pairs that need clipping (about a fifth)
 * or an edge family (about a tenth) no longer hold the lanes of pairs that are
 * done after the face families. */
/* vertices of both hulls into LDS (when they fit) and the pair geometry */
template <int L>
__device__ __forceinline__ void stage_pair(const Group<L>& g, const b3hShapeTables& sh, const b3hSatBody& A,
										   const b3hSatBody& B, F4* sA, F4* sB, PairGeom& P)
{
	const F4* gA = sh.vertices + A.vOff;
	const F4* gB = sh.vertices + B.vOff;
	bool ldsA = A.nV <= NP_VCAP, ldsB = B.nV <= NP_VCAP;
	wave_sync(); /* the previous pair's LDS reads are done */
	if (ldsA)
		for (int k = g.lane; k < A.nV; k += L) sA[k] = gA[k];
	if (ldsB)
		for (int k = g.lane; k < B.nV; k += L) sB[k] = gB[k];
	wave_sync();
	P.vA = ldsA ? sA : gA;
	P.vB = ldsB ? sB : gB;
	P.ldsA = ldsA;
	P.ldsB = ldsB;
	P.nA = A.nV;
	P.nB = B.nV;
	P.posA = A.pos;
	P.posB = B.pos;
	P.ornA = A.quat;
	P.ornB = B.quat;
}
/* the SAT frame: positions with w = 0 */
__device__ __forceinline__ void sat_frame(const PairGeom& P, PairGeom& Q)
{
	Q = P;
	Q.posA.w = 0.f;
	Q.posB.w = 0.f;
}
__device__ __forceinline__ const b3hSatFace* body_records(const b3hSatCache& sc, const b3hSatBody& d)
{
	return d.rec >= 0 ? sc.recs + d.rec : nullptr;
}
__device__ __forceinline__ const float2* body_dirs(const b3hSatCache& sc, const b3hSatBody& d)
{
	return d.dir >= 0 ? sc.dirs + (size_t)d.dir * B3_UNIT_SPHERE_162_COUNT : nullptr;
}
/* the pair's final record (lane 0 of the group): SAT state and axis, and for
 * a hull pair its flag and contact (the MPR point, then the reduced clip
 * points); pairs with a sphere or a plane get their contact from k_np_primitive */
__device__ void sat_finish(int i, int a, int b, bool aStatic, bool bStatic, bool prim, int state, F4 sep,
						   const F4* __restrict__ mprIn, int nReduced, const int* idx, const F4* pts,
						   int* __restrict__ flags, b3hContact4Data* __restrict__ staging, int* __restrict__ sepStateOut,
						   b3hFloat4* __restrict__ sepOut)
{
	sepStateOut[i] = state;
	sepOut[i] = sep;
	if (prim) return;
	F4 mprDH = mprIn[3 * (size_t)i + 2];
	bool mprHit = mprDH.y != 0.f;
	int flag = (mprHit || nReduced > 0) ? 1 : 0;
	flags[i] = flag;
	if (!flag) return;
	b3hContact4Data c;
	for (int k = 0; k < 4; k++) c.m_worldPosB[k] = f4(0, 0, 0, 0);
	c.m_restituitionCoeffCmp = (unsigned short)(0.f * 0xffff);
	c.m_frictionCoeffCmp = (unsigned short)(0.7f * 0xffff);
	c.m_batchIdx = i;
	c.m_bodyAPtrAndSignBit = aStatic ? -a : a;
	c.m_bodyBPtrAndSignBit = bStatic ? -b : b;
	c.m_childIndexA = -1;
	c.m_childIndexB = -1;
	c.m_unused1 = 0;
	c.m_unused2 = 0;
	c.m_worldNormalOnB = f4(0, 0, 0, 0);
	if (mprHit)
	{
		F4 mprDir = mprIn[3 * (size_t)i], mprPos = mprIn[3 * (size_t)i + 1];
		c.m_worldNormalOnB = neg4(mprDir);
		F4 p = mprPos;
		p.w = -mprDH.x;
		c.m_worldPosB[0] = p;
		c.m_worldNormalOnB.w = 1.f;
	}
	if (nReduced > 0)
	{
		c.m_worldNormalOnB = sep;
#pragma unroll
		for (int k = 0; k < 4; k++)
			if (k < nReduced && (k > 0 || !mprHit)) c.m_worldPosB[k] = pts[idx[k]];
		c.m_worldNormalOnB.w = (float)nReduced;
	}
	staging[i] = c;
}

/* pass 1: the face families (findSeparatingAxisKernel's faces of A then B).
 * launch bounds: 1024 / THREADS blocks per CU keep the register count at 128,
 * so four waves share each SIMD */
template <int L, int THREADS>
__global__ __launch_bounds__(THREADS, 1024 / THREADS) void k_sat_face_pass(
	const b3hInt4* __restrict__ pairs, const int* __restrict__ numPairsPtr, b3hShapeTables sh,
	b3hContact4Data* __restrict__ staging, int* __restrict__ flags, int* __restrict__ sepStateOut,
	b3hFloat4* __restrict__ sepOut, const F4* __restrict__ mprIn, b3hSatCache sc, float* __restrict__ dminOut,
	int* __restrict__ openFlag)
{
	constexpr int G = THREADS / L;
	__shared__ F4 sVerts[G][2][NP_VCAP];
	const Group<L> g;
	const int gi = threadIdx.x / L;
	int numPairs = *numPairsPtr;
	for (int i = blockIdx.x * G + gi; i < numPairs; i += gridDim.x * G)
	{
		b3hInt4 pr = pairs[i];
		int a = pr.x, b = pr.y;
		const b3hSatBody A = sc.desc[a], B = sc.desc[b];
		const bool hullA = A.flags & B3H_SB_HULL, hullB = B.flags & B3H_SB_HULL;
		const bool aStatic = A.flags & B3H_SB_STATIC, bStatic = B.flags & B3H_SB_STATIC;
		F4 mprDH = mprIn[3 * (size_t)i + 2];
		F4 sep = f4(0, 0, 0, 0);
		if (mprDH.y != 0.f) sep = neg4(mprIn[3 * (size_t)i]);
		if (!(hullA && hullB) || (aStatic && bStatic))
		{
			if (g.lane == 0)
			{
				openFlag[i] = 0;
				sat_finish(i, a, b, aStatic, bStatic, !(hullA && hullB), 0, f4(0, 0, 0, 0), mprIn, 0, nullptr,
						   nullptr, flags, staging, sepStateOut, sepOut);
			}
			continue;
		}
		PairGeom P, Q;
		stage_pair(g, sh, A, B, sVerts[gi][0], sVerts[gi][1], P);
		sat_frame(P, Q);
		F4 dc = sub4(A.center, B.center);
		float dmin = FLT_MAX;
		F4 sepN = sep;
		int state = 0;
		float dminSaved = FLT_MAX;
		Family fa = {0, A.nF, sh.faces + A.fOff, nullptr, nullptr, 1, body_records(sc, A), nullptr, nullptr};
		if (sat_family(g, fa, Q, dc, &sepN, &dmin))
		{
			Family fb = {1, B.nF, sh.faces + B.fOff, nullptr, nullptr, 1, body_records(sc, B), nullptr, nullptr};
			if (sat_family(g, fb, Q, dc, &sepN, &dmin))
			{
				state = 1;
				dminSaved = dmin;
				sep = sepN;
			}
			else
				state = 4;
		}
		if (g.lane == 0)
		{
			openFlag[i] = state != 0;
			if (state == 0)
				sat_finish(i, a, b, aStatic, bStatic, false, 0, sep, mprIn, 0, nullptr, nullptr, flags, staging,
						   sepStateOut, sepOut);
			else
			{
				sepStateOut[i] = state;
				sepOut[i] = sep;
				dminOut[i] = dminSaved;
			}
		}
	}
}

/* pass 2: the edge family (exact edge pairs, sat.cl:719-798, or the 162 fixed
 * directions, mpr.cl:180-311) of the listed pairs (those the face pass left
 * open), one pair per wavefront; each sets its clip flag (1: clip it) */
template <int L, int THREADS>
__global__ __launch_bounds__(THREADS, 1024 / THREADS) void k_sat_edge_pass(
	const b3hInt4* __restrict__ pairs, b3hShapeTables sh, b3hContact4Data* __restrict__ staging, int* __restrict__ flags,
	int* __restrict__ sepStateOut, b3hFloat4* __restrict__ sepOut, const F4* __restrict__ mprIn, b3hSatCache sc,
	const float* __restrict__ dminIn, const int* __restrict__ listCount, const int* __restrict__ list,
	int* __restrict__ clipFlag)
{
	constexpr int G = THREADS / L;
	__shared__ F4 sVerts[G][2][NP_VCAP];
	const Group<L> g;
	const int gi = threadIdx.x / L;
	int n = *listCount;
	for (int k = blockIdx.x * G + gi; k < n; k += gridDim.x * G)
	{
		int i = list[k];
		b3hInt4 pr = pairs[i];
		int a = pr.x, b = pr.y;
		const b3hSatBody A = sc.desc[a], B = sc.desc[b];
		F4 dc = sub4(A.center, B.center);
		F4 sep = sepOut[i];
		float dm = dminIn[i];
		F4 s2 = sep;
		int nEE = A.nE * B.nE;
		Family fe;
		bool cached = false;
		if (nEE <= B3_UNIT_SPHERE_162_COUNT)
			fe = {2, nEE, nullptr, sh.uniqueEdges + A.eOff, sh.uniqueEdges + B.eOff, B.nE, nullptr, nullptr, nullptr};
		else
		{
			const float2* dA = body_dirs(sc, A);
			const float2* dB = body_dirs(sc, B);
			fe = {3, B3_UNIT_SPHERE_162_COUNT, nullptr, nullptr, nullptr, 1, nullptr, dB ? dA : nullptr, dB};
			cached = dA && dB;
		}
		/* the vertices are staged only when an axis needs projecting: the
		 * fixed directions with both hulls' intervals cached read none */
		PairGeom P, Q;
		if (cached)
		{
			P.vA = sh.vertices + A.vOff;
			P.vB = sh.vertices + B.vOff;
			P.ldsA = P.ldsB = false;
			P.nA = A.nV;
			P.nB = B.nV;
			P.posA = A.pos;
			P.posB = B.pos;
			P.ornA = A.quat;
			P.ornB = B.quat;
		}
		else
			stage_pair(g, sh, A, B, sVerts[gi][0], sVerts[gi][1], P);
		sat_frame(P, Q);
		bool open = sat_family(g, fe, Q, dc, &s2, &dm);
		if (g.lane == 0)
		{
			if (!open)
			{
				clipFlag[i] = 0;
				sat_finish(i, a, b, A.flags & B3H_SB_STATIC, B.flags & B3H_SB_STATIC, false, 0, sep, mprIn, 0, nullptr,
						   nullptr, flags, staging, sepStateOut, sepOut);
			}
			else
			{
				clipFlag[i] = 1;
				sepStateOut[i] = 1;
				sepOut[i] = s2;
			}
		}
	}
}

/* the flagged pairs in pair order, and their number */
__global__ __launch_bounds__(256) void k_pair_list(const int* __restrict__ numPairsPtr, const int* __restrict__ flag,
												   const int* __restrict__ scan, int* __restrict__ list,
												   int* __restrict__ count)
{
	int n = *numPairsPtr;
	if (blockIdx.x == 0 && threadIdx.x == 0) *count = scan[n];
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
		if (flag[i]) list[scan[i]] = i;
}

/* pass 3: clipping and contact reduction of the listed pairs (clipHullHullKernel,
 * satClipHullContacts.cl:890-1000), L lanes per pair */
template <int L, int THREADS>
__global__ __launch_bounds__(THREADS, 3) void k_sat_clip_pass(
	const b3hInt4* __restrict__ pairs, const int* __restrict__ clipCount, const int* __restrict__ clipList,
	b3hShapeTables sh, b3hContact4Data* __restrict__ staging, int* __restrict__ flags, int* __restrict__ sepStateOut,
	b3hFloat4* __restrict__ sepOut, const F4* __restrict__ mprIn, b3hSatCache sc)
{
	constexpr int G = THREADS / L;
	__shared__ F4 sClip[G][3][CLIP_CAP];
	const Group<L> g;
	const int gi = threadIdx.x / L;
	int n = *clipCount;
	for (int k = blockIdx.x * G + gi; k < n; k += gridDim.x * G)
	{
		int i = clipList[k];
		b3hInt4 pr = pairs[i];
		int a = pr.x, b = pr.y;
		const b3hSatBody A = sc.desc[a], B = sc.desc[b];
		/* clipping reads a few vertices of two faces: from the (cached) table */
		PairGeom P;
		P.vA = sh.vertices + A.vOff;
		P.vB = sh.vertices + B.vOff;
		P.ldsA = P.ldsB = false;
		P.nA = A.nV;
		P.nB = B.nV;
		P.posA = A.pos;
		P.posB = B.pos;
		P.ornA = A.quat;
		P.ornB = B.quat;
		wave_sync(); /* the previous pair's clip buffers are read */
		F4 sep = sepOut[i];
		F4* pts = sClip[gi][2];
		int idx[4] = {0, 0, 0, 0};
		int nReduced = 0;
		int np = clip_hull_hull(g, sep, P, A, B, sh, sClip[gi][0], sClip[gi][1], pts, body_records(sc, A),
								body_records(sc, B));
		if (np > 0) nReduced = extract_manifold(g, pts, np, neg4(sep), idx);
		if (g.lane == 0)
			sat_finish(i, a, b, A.flags & B3H_SB_STATIC, B.flags & B3H_SB_STATIC, false, 1, sep, mprIn, nReduced, idx,
					   pts, flags, staging, sepStateOut, sepOut);
	}
}

/* ---------------- primitive pairs ----------------
 * primitiveContactsKernel (primitiveContacts.cl:793-977): a plane or a sphere
 * against a hull, a sphere or a plane; one thread per pair.  The reference
 * appends these contacts with an atomic counter before the hull pairs; here
 * they take their pair's slot like every contact. */
__device__ __forceinline__ F4 pc_transform(F4 p, F4 t, F4 q) { return add4(qrot(q, p), t); }
__device__ __forceinline__ void pc_tr_inverse(F4 t, F4 q, F4* ti, F4* qi)
{
	*qi = qinv(q);
	*ti = qrot(*qi, neg4(t));
}
__device__ __forceinline__ void pc_tr_mul(F4 tA, F4 qA, F4 tB, F4 qB, F4* to, F4* qo)
{
	*qo = qmul(qA, qB);
	*to = pc_transform(tB, tA, qA);
}
__device__ __forceinline__ void pc_header(b3hContact4Data& c, int pairIndex, int bA, int bB,
										  const b3hRigidBodyData* bodies)
{
	for (int k = 0; k < 4; k++) c.m_worldPosB[k] = f4(0, 0, 0, 0);
	c.m_restituitionCoeffCmp = (unsigned short)(0.f * 0xffff);
	c.m_frictionCoeffCmp = (unsigned short)(0.7f * 0xffff);
	c.m_batchIdx = pairIndex;
	c.m_bodyAPtrAndSignBit = bodies[bA].m_invMass == 0 ? -bA : bA;
	c.m_bodyBPtrAndSignBit = bodies[bB].m_invMass == 0 ? -bB : bB;
	c.m_childIndexA = -1;
	c.m_childIndexB = -1;
	c.m_unused1 = 0;
	c.m_unused2 = 0;
}

/* extractManifoldSequential (primitiveContacts.cl:489-582): four of n > 4
 * points (P(i) gives point i); the extreme searches start at FLT_MIN, as the
 * reference's do */
template <class PT>
__device__ int pc_extract_manifold(const PT& P, int n, F4 nearNormal, int* idx)
{
	if (n > 64) n = 64;
	F4 center = f4(0.f, 0.f, 0.f, 0.f);
	for (int i = 0; i < n; i++) center = add4(center, P(i));
	const float fn = (float)n;
	center = f4(center.x / fn, center.y / fn, center.z / fn, center.w / fn);
	const F4 a = sub4(P(0), center);
	F4 u = cross3(nearNormal, a);
	F4 v = cross3(nearNormal, u);
	u = nrm3(u);
	v = nrm3(v);
	float minW = FLT_MAX;
	int minIndex = -1;
	float md[4] = {FLT_MIN, FLT_MIN, FLT_MIN, FLT_MIN};
	const F4 nu = neg4(u), nv = neg4(v);
	for (int ie = 0; ie < n; ie++)
	{
		const F4 pe = P(ie);
		if (pe.w < minW)
		{
			minW = pe.w;
			minIndex = ie;
		}
		const F4 r = sub4(pe, center);
		const float f0 = dot3(u, r), f1 = dot3(nu, r), f2 = dot3(v, r), f3 = dot3(nv, r);
		if (f0 < md[0]) md[0] = f0, idx[0] = ie;
		if (f1 < md[1]) md[1] = f1, idx[1] = ie;
		if (f2 < md[2]) md[2] = f2, idx[2] = ie;
		if (f3 < md[3]) md[3] = f3, idx[3] = ie;
	}
	if (idx[0] != minIndex && idx[1] != minIndex && idx[2] != minIndex && idx[3] != minIndex) idx[0] = minIndex;
	return 4;
}

#define PC_MAX_PLANE_POINTS 64
#define PC_THREADS 128
/* computeContactPlaneConvex (:584-727): the hull's vertices below the plane,
 * reduced to four; contact bodies (plane, hull).  The kept points are held
 * as vertex numbers (an LDS column of the thread, stride PC_THREADS) and
 * recomputed where needed, with the same operations. */
/* hullCol, hullPos, hullOrn: the hull's collidable and pose (the body's own,
 * or a compound child's, processCompoundPairsPrimitivesKernel :980-1090) */
__device__ bool pc_plane_convex(int pairIndex, int bPlane, int bHull, const b3hRigidBodyData* bodies,
								const b3hShapeTables& sh, int* col, b3hContact4Data& c, int hullCol, F4 hullPos,
								F4 hullOrn)
{
	const b3hRigidBodyData& A = bodies[bPlane];
	struct
	{
		F4 m_pos, m_quat;
	} B = {hullPos, hullOrn};
	const b3hConvexPolyhedronData hullB = sh.convexes[sh.collidables[hullCol].m_shapeIndex];
	const F4 planeEq = sh.faces[sh.collidables[A.m_collidableIdx].m_shapeIndex].m_plane;
	const F4 planeNormal = f4(planeEq.x, planeEq.y, planeEq.z, 0.f);
	const F4 planeNormalWorld = qrot(A.m_quat, planeNormal);
	const float planeConstant = planeEq.w;
	F4 invPosA, invOrnA, invPosB, invOrnB, pic, qic;
	pc_tr_inverse(A.m_pos, A.m_quat, &invPosA, &invOrnA);
	pc_tr_inverse(B.m_pos, B.m_quat, &invPosB, &invOrnB);
	pc_tr_mul(invPosB, invOrnB, A.m_pos, A.m_quat, &pic, &qic);
	const F4 nInConvex = qrot(qic, neg4(planeNormal));
	const F4* verts = sh.vertices + hullB.m_vertexOffset;
	/* world vertex with w = its distance below the plane */
	auto point = [&](int vi, float* distOut) {
		F4 vw = pc_transform(verts[vi], B.m_pos, B.m_quat);
		const F4 vp = pc_transform(vw, invPosA, invOrnA);
		const float dist = dot4(planeNormal, vp) - planeConstant;
		vw.w = dist;
		*distOut = dist;
		return vw;
	};
	float maxDot = -1e30f;
	int numPoints = 0;
	int idx[4] = {0, 1, 2, 3};
	for (int i = 0; i < hullB.m_numVertices; i++)
	{
		const float curDot = dot4(verts[i], nInConvex);
		if (curDot > maxDot)
		{
			maxDot = curDot;
			if (numPoints == PC_MAX_PLANE_POINTS) numPoints--;
		}
		if (numPoints < PC_MAX_PLANE_POINTS)
		{
			float dist;
			point(i, &dist);
			if (dist < 0.f) col[PC_THREADS * (numPoints++)] = i;
		}
	}
	auto P = [&](int k) {
		float d;
		return point(col[PC_THREADS * k], &d);
	};
	int nr = numPoints;
	if (numPoints > 4) nr = pc_extract_manifold(P, numPoints, nInConvex, idx);
	if (nr <= 0) return false;
	pc_header(c, pairIndex, bPlane, bHull, bodies);
	c.m_worldNormalOnB = neg4(planeNormalWorld);
	for (int k = 0; k < 4; k++)
		if (k < nr) c.m_worldPosB[k] = P(idx[k]);
	c.m_worldNormalOnB.w = (float)nr;
	return true;
}

/* computeContactPlaneSphere (:729-791); contact bodies (plane, sphere) */
__device__ bool pc_plane_sphere(int pairIndex, int bPlane, int bSphere, const b3hRigidBodyData* bodies,
								const b3hShapeTables& sh, b3hContact4Data& c)
{
	const b3hRigidBodyData& A = bodies[bPlane];
	const b3hRigidBodyData& B = bodies[bSphere];
	const F4 planeEq = sh.faces[sh.collidables[A.m_collidableIdx].m_shapeIndex].m_plane;
	const float radius = sh.collidables[B.m_collidableIdx].m_radius;
	const F4 n1 = f4(planeEq.x, planeEq.y, planeEq.z, 0.f);
	const float planeConstant = planeEq.w;
	F4 invPosA, invOrnA, invPosB, invOrnB, cpPos, cpOrn, pcPos, pcOrn;
	pc_tr_inverse(A.m_pos, A.m_quat, &invPosA, &invOrnA);
	pc_tr_mul(invPosA, invOrnA, B.m_pos, B.m_quat, &cpPos, &cpOrn);
	pc_tr_inverse(B.m_pos, B.m_quat, &invPosB, &invOrnB);
	pc_tr_mul(invPosB, invOrnB, A.m_pos, A.m_quat, &pcPos, &pcOrn);
	const F4 vtx1 = mul4(qrot(pcOrn, neg4(n1)), radius);
	const F4 vip = pc_transform(vtx1, cpPos, cpOrn);
	const float distance = dot3(n1, vip) - planeConstant;
	if (!(distance < 0.f)) return false;
	const F4 proj = sub4(vip, smul4(distance, n1));
	const F4 world = pc_transform(proj, A.m_pos, A.m_quat);
	const F4 nB = qrot(A.m_quat, n1);
	F4 p = add4(world, mul4(nB, distance));
	p.w = distance;
	pc_header(c, pairIndex, bPlane, bSphere, bodies);
	c.m_worldNormalOnB = neg4(nB);
	c.m_worldPosB[0] = p;
	c.m_worldNormalOnB.w = 1.f;
	return true;
}

/* IsPointInPolygon (:280-337): true if p lies inside the face's edges;
 * otherwise *out is the closest point of the first edge it lies outside (the
 * reference leaves out.w unset on the edge's interior; 0 here) */
__device__ bool pc_point_in_polygon(F4 p, const b3hGpuFace& face, const F4* base, const int* indices, F4* out)
{
	const F4 plane = f4(face.m_plane.x, face.m_plane.y, face.m_plane.z, 0.f);
	if (face.m_numIndices < 2) return false;
	F4 b = base[indices[face.m_indexOffset + face.m_numIndices - 1]];
	for (int i = 0; i < face.m_numIndices; i++)
	{
		const F4 a = b;
		b = base[indices[face.m_indexOffset + i]];
		const F4 ab = sub4(b, a), ap = sub4(p, a);
		const F4 v = cross3(ab, plane);
		if (dot4(ap, v) > 0.f)
		{
			const float abm2 = dot4(ab, ab);
			const float rt = abm2 != 0.f ? dot4(ab, ap) / abm2 : 0.f;
			if (rt <= 0.f)
				*out = a;
			else if (rt >= 1.f)
				*out = b;
			else
			{
				const float s = 1.f - rt;
				*out = f4(s * a.x + rt * b.x, s * a.y + rt * b.y, s * a.z + rt * b.z, 0.f);
			}
			return false;
		}
	}
	return true;
}

/* computeContactSphereConvex (:339-487); contact bodies (sphere, hull) */
__device__ bool pc_sphere_convex(int pairIndex, int bSphere, int bHull, const b3hRigidBodyData* bodies,
								 const b3hShapeTables& sh, b3hContact4Data& c, int hullCol, F4 hullPos, F4 hullOrn)
{
	const b3hRigidBodyData& S = bodies[bSphere];
	const float radius = sh.collidables[S.m_collidableIdx].m_radius;
	const F4 pos = hullPos, quat = hullOrn;
	F4 invPos, invOrn;
	pc_tr_inverse(pos, quat, &invPos, &invOrn);
	const F4 sp = pc_transform(S.m_pos, invPos, invOrn);
	const b3hConvexPolyhedronData hull = sh.convexes[sh.collidables[hullCol].m_shapeIndex];
	F4 closest = f4(0.f, 0.f, 0.f, 0.f), hitN = f4(0.f, 0.f, 0.f, 0.f);
	float minDist = -1000000.f;
	bool collide = true;
	for (int f = 0; f < hull.m_numFaces; f++)
	{
		const b3hGpuFace face = sh.faces[hull.m_faceOffset + f];
		const F4 planeEqn = face.m_plane;
		const F4 n = f4(planeEqn.x, planeEqn.y, planeEqn.z, 0.f);
		const float dist0 = dot3(n, sp) + planeEqn.w;
		const F4 pnt = sub4(sp, smul4(dist0, n));
		float dist = dist0;
		if (dist > radius)
		{
			collide = false;
			break;
		}
		if (dist > 0.f)
		{
			F4 out = f4(0.f, 0.f, 0.f, 0.f);
			if (pc_point_in_polygon(sp, face, sh.vertices + hull.m_vertexOffset, sh.indices, &out))
			{
				if (dist > minDist)
				{
					minDist = dist;
					closest = pnt;
					hitN = planeEqn;
				}
			}
			else
			{
				const F4 tmp = sub4(sp, out);
				const float l2 = dot4(tmp, tmp);
				if (l2 < radius * radius)
				{
					dist = sqrtf(l2);
					if (dist > minDist)
					{
						minDist = dist;
						closest = out;
						hitN = f4(tmp.x / dist, tmp.y / dist, tmp.z / dist, tmp.w / dist);
					}
				}
				else
				{
					collide = false;
					break;
				}
			}
		}
		else if (dist > minDist)
		{
			minDist = dist;
			closest = pnt;
			hitN = f4(planeEqn.x, planeEqn.y, planeEqn.z, hitN.w);
		}
	}
	if (!(collide && minDist > -10000.f)) return false;
	const F4 nB = qrot(quat, neg4(hitN));
	F4 p = pc_transform(closest, pos, quat);
	const float depth = minDist - radius;
	if (!(depth <= 0.f)) return false;
	p.w = depth;
	pc_header(c, pairIndex, bSphere, bHull, bodies);
	c.m_worldNormalOnB = neg4(nB);
	c.m_worldPosB[0] = p;
	c.m_worldNormalOnB.w = 1.f;
	return true;
}

/* sphere-sphere (primitiveContactsKernel :905-951); contact bodies (A, B) */
__device__ bool pc_sphere_sphere(int pairIndex, int a, int b, const b3hRigidBodyData* bodies, const b3hShapeTables& sh,
								 b3hContact4Data& c)
{
	const float rA = sh.collidables[bodies[a].m_collidableIdx].m_radius;
	const float rB = sh.collidables[bodies[b].m_collidableIdx].m_radius;
	const F4 posA = bodies[a].m_pos, posB = bodies[b].m_pos;
	const F4 diff = sub4(posA, posB);
	const float len = sqrtf(dot4(diff, diff));
	if (!(len <= rA + rB)) return false;
	const float dist = len - (rA + rB);
	F4 n = f4(1.f, 0.f, 0.f, 0.f);
	if ((double)len > 0.00001) n = f4(diff.x / len, diff.y / len, diff.z / len, diff.w / len);
	F4 p = add4(posB, mul4(n, rB));
	p.w = dist;
	pc_header(c, pairIndex, a, b, bodies);
	c.m_worldNormalOnB = n;
	c.m_worldPosB[0] = p;
	c.m_worldNormalOnB.w = 1.f;
	return true;
}

/* every pair that is not hull x hull: its contact (staging) and flag; the
 * hull x hull pairs are the SAT kernel's */
__global__ __launch_bounds__(PC_THREADS) void k_np_primitive(const b3hInt4* __restrict__ pairs,
															 const int* __restrict__ numPairsPtr,
															 const b3hRigidBodyData* __restrict__ bodies,
															 b3hShapeTables sh, b3hContact4Data* __restrict__ staging,
															 int* __restrict__ flags)
{
	__shared__ int sIdx[PC_MAX_PLANE_POINTS * PC_THREADS];
	int* col = sIdx + threadIdx.x;
	const int numPairs = *numPairsPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numPairs; i += gridDim.x * blockDim.x)
	{
		const b3hInt4 pr = pairs[i];
		const int ta = sh.collidables[bodies[pr.x].m_collidableIdx].m_shapeType;
		const int tb = sh.collidables[bodies[pr.y].m_collidableIdx].m_shapeType;
		if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_CONVEX_HULL) continue;
		b3hContact4Data c;
		bool hit = false;
		const b3hRigidBodyData& A = bodies[pr.x];
		const b3hRigidBodyData& B = bodies[pr.y];
		if (ta == B3H_SHAPE_PLANE && tb == B3H_SHAPE_CONVEX_HULL)
			hit = pc_plane_convex(i, pr.x, pr.y, bodies, sh, col, c, B.m_collidableIdx, B.m_pos, B.m_quat);
		else if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_PLANE)
			hit = pc_plane_convex(i, pr.y, pr.x, bodies, sh, col, c, A.m_collidableIdx, A.m_pos, A.m_quat);
		else if (ta == B3H_SHAPE_PLANE && tb == B3H_SHAPE_SPHERE)
			hit = pc_plane_sphere(i, pr.x, pr.y, bodies, sh, c);
		else if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_PLANE)
			hit = pc_plane_sphere(i, pr.y, pr.x, bodies, sh, c);
		else if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_CONVEX_HULL)
			hit = pc_sphere_convex(i, pr.x, pr.y, bodies, sh, c, B.m_collidableIdx, B.m_pos, B.m_quat);
		else if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_SPHERE)
			hit = pc_sphere_convex(i, pr.y, pr.x, bodies, sh, c, A.m_collidableIdx, A.m_pos, A.m_quat);
		else if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_SPHERE)
			hit = pc_sphere_sphere(i, pr.x, pr.y, bodies, sh, c);
		flags[i] = hit ? 1 : 0;
		if (hit) staging[i] = c;
	}
}

/* ---------------- compound pairs ----------------
 * findCompoundPairsKernel (sat.cl:945-1130): a body pair with a compound
 * becomes child pairs (bodyA, bodyB, childA, childB), -1 for a side that is
 * not a compound.  Compound x compound: the children whose quantized boxes
 * (transformed to world) overlap, found by walking both child trees from every
 * overlapping pair of subtrees (the leaf's child number is the child index,
 * used as an index into the child table as the reference does).  Compound x
 * other: every child.  Here the list is built in two passes (count, then emit
 * at prefix-sum offsets), in pair order and walk order, so it is the same on
 * every run; the reference appends with an atomic counter. */
#define CP_STACK 256 /* B3_MAX_STACK_DEPTH, sat.cl:10 */
__device__ __forceinline__ F4 bvh_unquantize(const unsigned short* q, F4 quant, F4 bmin)
{
	F4 v = f4((float)q[0] / quant.x, (float)q[1] / quant.y, (float)q[2] / quant.z, 0.f);
	return add4(v, bmin);
}
__device__ __forceinline__ int bvh_child_index(int escapeOrIndex)
{
	unsigned y = (~0u) << (31 - 10);
	return escapeOrIndex & ~(int)y;
}
/* walks the two child trees of a compound x compound pair; emit(ca, cb) per
 * overlapping leaf pair, in the reference's order */
template <class EMIT>
__device__ void compound_walk(const b3hShapeTables& sh, const b3hBvhInfo& iA, F4 posA, F4 ornA, const b3hBvhInfo& iB,
							  F4 posB, F4 ornB, EMIT emit)
{
	int2 stack[CP_STACK];
	for (int p = 0; p < iA.m_numSubTrees; p++)
	{
		const b3hBvhSubtreeInfo sa = sh.bvhSubtrees[iA.m_subTreeOffset + p];
		F4 aMn, aMx;
		transform_aabb2(bvh_unquantize(sa.m_quantizedAabbMin, iA.m_quantization, iA.m_aabbMin),
						bvh_unquantize(sa.m_quantizedAabbMax, iA.m_quantization, iA.m_aabbMin), posA, ornA, &aMn, &aMx);
		for (int q = 0; q < iB.m_numSubTrees; q++)
		{
			const b3hBvhSubtreeInfo sb = sh.bvhSubtrees[iB.m_subTreeOffset + q];
			F4 bMn, bMx;
			transform_aabb2(bvh_unquantize(sb.m_quantizedAabbMin, iB.m_quantization, iB.m_aabbMin),
							bvh_unquantize(sb.m_quantizedAabbMax, iB.m_quantization, iB.m_aabbMin), posB, ornB, &bMn,
							&bMx);
			if (!aabb_overlap(aMn, aMx, bMn, bMx)) continue;
			int depth = 0;
			stack[depth++] = make_int2(sa.m_rootNodeIndex + iA.m_nodeOffset, sb.m_rootNodeIndex + iB.m_nodeOffset);
			do
			{
				int2 node = stack[--depth];
				const b3hQuantizedBvhNode na = sh.bvhNodes[node.x], nb = sh.bvhNodes[node.y];
				F4 mnA, mxA, mnB, mxB;
				transform_aabb2(bvh_unquantize(na.m_quantizedAabbMin, iA.m_quantization, iA.m_aabbMin),
								bvh_unquantize(na.m_quantizedAabbMax, iA.m_quantization, iA.m_aabbMin), posA, ornA,
								&mnA, &mxA);
				transform_aabb2(bvh_unquantize(nb.m_quantizedAabbMin, iB.m_quantization, iB.m_aabbMin),
								bvh_unquantize(nb.m_quantizedAabbMax, iB.m_quantization, iB.m_aabbMin), posB, ornB,
								&mnB, &mxB);
				if (!aabb_overlap(mnA, mxA, mnB, mxB)) continue;
				bool leafA = na.m_escapeIndexOrTriangleIndex >= 0, leafB = nb.m_escapeIndexOrTriangleIndex >= 0;
				if (depth + 4 > CP_STACK && !(leafA && leafB)) continue;
				if (!leafA)
				{
					int la = node.x + 1;
					int ra = sh.bvhNodes[la].m_escapeIndexOrTriangleIndex >= 0
								 ? node.x + 2
								 : node.x + 1 - sh.bvhNodes[la].m_escapeIndexOrTriangleIndex;
					if (!leafB)
					{
						int lb = node.y + 1;
						int rb = sh.bvhNodes[lb].m_escapeIndexOrTriangleIndex >= 0
									 ? node.y + 2
									 : node.y + 1 - sh.bvhNodes[lb].m_escapeIndexOrTriangleIndex;
						stack[depth++] = make_int2(la, lb);
						stack[depth++] = make_int2(ra, lb);
						stack[depth++] = make_int2(la, rb);
						stack[depth++] = make_int2(ra, rb);
					}
					else
					{
						stack[depth++] = make_int2(la, node.y);
						stack[depth++] = make_int2(ra, node.y);
					}
				}
				else if (!leafB)
				{
					int lb = node.y + 1;
					int rb = sh.bvhNodes[lb].m_escapeIndexOrTriangleIndex >= 0
								 ? node.y + 2
								 : node.y + 1 - sh.bvhNodes[lb].m_escapeIndexOrTriangleIndex;
					stack[depth++] = make_int2(node.x, lb);
					stack[depth++] = make_int2(node.x, rb);
				}
				else
					emit(bvh_child_index(na.m_escapeIndexOrTriangleIndex), bvh_child_index(nb.m_escapeIndexOrTriangleIndex));
			} while (depth);
		}
	}
}
/* the child pairs of body pair (a, b): emit(ca, cb) for each, in order */
template <class EMIT>
__device__ void compound_children(const b3hShapeTables& sh, const b3hRigidBodyData* bodies, int a, int b, EMIT emit)
{
	const b3hRigidBodyData A = bodies[a], B = bodies[b];
	if (A.m_invMass == 0 && B.m_invMass == 0) return;
	const b3hCollidable cA = sh.collidables[A.m_collidableIdx], cB = sh.collidables[B.m_collidableIdx];
	const bool compA = cA.m_shapeType == B3H_SHAPE_COMPOUND_OF_CONVEX_HULLS;
	const bool compB = cB.m_shapeType == B3H_SHAPE_COMPOUND_OF_CONVEX_HULLS;
	if (compA && compB)
		compound_walk(sh, sh.bvhInfos[cA.m_compoundBvhIndex], A.m_pos, A.m_quat, sh.bvhInfos[cB.m_compoundBvhIndex],
					  B.m_pos, B.m_quat, emit);
	else if (compA)
		for (int c = 0; c < cA.m_numChildShapes; c++) emit(cA.m_shapeIndex + c, -1);
	else if (cA.m_shapeType != B3H_SHAPE_CONCAVE_TRIMESH && compB)
		for (int c = 0; c < cB.m_numChildShapes; c++) emit(-1, cB.m_shapeIndex + c);
}
__global__ __launch_bounds__(256) void k_compound_count(const b3hInt4* __restrict__ pairs, const int* __restrict__ numPairsPtr,
														const b3hRigidBodyData* __restrict__ bodies, b3hShapeTables sh,
														int* __restrict__ counts)
{
	int n = *numPairsPtr;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		int cnt = 0;
		compound_children(sh, bodies, pairs[i].x, pairs[i].y, [&](int, int) { cnt++; });
		counts[i] = cnt;
	}
}
/* the child pairs at their offsets (capped at cap); *numOut = min(total, cap),
 * numOut[1] = total */
__global__ __launch_bounds__(256) void k_compound_emit(const b3hInt4* __restrict__ pairs, const int* __restrict__ numPairsPtr,
													   const b3hRigidBodyData* __restrict__ bodies, b3hShapeTables sh,
													   const int* __restrict__ offsets, int cap, b3hInt4* __restrict__ out,
													   int* __restrict__ numOut)
{
	int n = *numPairsPtr;
	if (blockIdx.x == 0 && threadIdx.x == 0)
	{
		int total = offsets[n];
		numOut[0] = total < cap ? total : cap;
		numOut[1] = total;
	}
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		int a = pairs[i].x, b = pairs[i].y;
		const int k0 = offsets[i];
		int k = k0;
		compound_children(sh, bodies, a, b, [&](int ca, int cb) {
			if (k < cap) out[k] = b3hInt4{a, b, ca, cb};
			k++;
		});
		/* the body pair's child pairs ordered by (childA, childB) as unsigned
		 * (the solver's contact order; the walk order of two trees is not it) */
		int k1 = k < cap ? k : cap;
		for (int x = k0 + 1; x < k1; x++)
		{
			b3hInt4 v = out[x];
			int y = x - 1;
			while (y >= k0 && ((unsigned)out[y].z > (unsigned)v.z ||
							   ((unsigned)out[y].z == (unsigned)v.z && (unsigned)out[y].w > (unsigned)v.w)))
			{
				out[y + 1] = out[y];
				y--;
			}
			out[y + 1] = v;
		}
	}
}
/* a side of a child pair: the child's collidable and pose, or the body's
 * (processCompoundPairsKernel, sat.cl:830-857) */
__device__ __forceinline__ void compound_side(const b3hShapeTables& sh, const b3hRigidBodyData& body, int child, int* col,
											  F4* pos, F4* orn)
{
	*pos = body.m_pos;
	*orn = body.m_quat;
	*col = body.m_collidableIdx;
	if (child >= 0)
	{
		const b3hGpuChildShape c = sh.children[child];
		*col = c.m_shapeIndex;
		*pos = add4(qrot(body.m_quat, c.m_childPosition), body.m_pos);
		*orn = qmul(body.m_quat, c.m_childOrientation);
	}
}
/* processCompoundPairsKernel (sat.cl:802-925) and clipCompoundsHullHullKernel
 * (satClipHullContacts.cl:1003-1130), one child pair per L lanes: faces of
 * A, faces of B, then every edge pair (the exact edge test whatever the edge
 * count), then clipping and reduction; the contact carries the child indices
 * and the child pair's index.  Child pairs that are not hull x hull get their
 * flag here (0) and their contact, if any, from k_compound_primitive. */
template <int L, int THREADS>
__global__ __launch_bounds__(THREADS, 1024 / THREADS) void k_compound_sat(const b3hInt4* __restrict__ cpairs,
																		  const int* __restrict__ numPtr,
																		  const b3hRigidBodyData* __restrict__ bodies,
																		  b3hShapeTables sh,
																		  b3hContact4Data* __restrict__ staging,
																		  int* __restrict__ flags)
{
	constexpr int G = THREADS / L;
	__shared__ F4 sVerts[G][2][NP_VCAP];
	__shared__ F4 sClip[G][3][CLIP_CAP];
	const Group<L> g;
	const int gi = threadIdx.x / L;
	int n = *numPtr;
	for (int k = blockIdx.x * G + gi; k < n; k += gridDim.x * G)
	{
		const b3hInt4 cp = cpairs[k];
		const b3hRigidBodyData BA = bodies[cp.x], BB = bodies[cp.y];
		int colA, colB;
		b3hSatBody A, B;
		compound_side(sh, BA, cp.z, &colA, &A.pos, &A.quat);
		compound_side(sh, BB, cp.w, &colB, &B.pos, &B.quat);
		const b3hCollidable cA = sh.collidables[colA], cB = sh.collidables[colB];
		int hit = 0;
		int nReduced = 0;
		int idx[4] = {0, 0, 0, 0};
		F4 sep = f4(1.f, 0.f, 0.f, 0.f);
		F4* pts = sClip[gi][2];
		if (cA.m_shapeType == B3H_SHAPE_CONVEX_HULL && cB.m_shapeType == B3H_SHAPE_CONVEX_HULL)
		{
			const b3hConvexPolyhedronData ha = sh.convexes[cA.m_shapeIndex], hb = sh.convexes[cB.m_shapeIndex];
			A.vOff = ha.m_vertexOffset;
			A.nV = ha.m_numVertices;
			A.fOff = ha.m_faceOffset;
			A.nF = ha.m_numFaces;
			B.vOff = hb.m_vertexOffset;
			B.nV = hb.m_numVertices;
			B.fOff = hb.m_faceOffset;
			B.nF = hb.m_numFaces;
			PairGeom P, Q;
			stage_pair(g, sh, A, B, sVerts[gi][0], sVerts[gi][1], P);
			sat_frame(P, Q);
			F4 dc = sub4(xform(ha.m_localCenter, Q.posA, Q.ornA), xform(hb.m_localCenter, Q.posB, Q.ornB));
			float dmin = FLT_MAX;
			Family fa = {0, ha.m_numFaces, sh.faces + ha.m_faceOffset, nullptr, nullptr, 1, nullptr, nullptr, nullptr};
			Family fb = {1, hb.m_numFaces, sh.faces + hb.m_faceOffset, nullptr, nullptr, 1, nullptr, nullptr, nullptr};
			Family fe = {2,
						 ha.m_numUniqueEdges * hb.m_numUniqueEdges,
						 nullptr,
						 sh.uniqueEdges + ha.m_uniqueEdgesOffset,
						 sh.uniqueEdges + hb.m_uniqueEdgesOffset,
						 hb.m_numUniqueEdges,
						 nullptr,
						 nullptr,
						 nullptr};
			if (sat_family(g, fa, Q, dc, &sep, &dmin) && sat_family(g, fb, Q, dc, &sep, &dmin) &&
				sat_family(g, fe, Q, dc, &sep, &dmin))
			{
				wave_sync();
				int np = clip_hull_hull(g, sep, P, A, B, sh, sClip[gi][0], sClip[gi][1], pts, nullptr, nullptr);
				if (np > 0)
				{
					nReduced = extract_manifold(g, pts, np, neg4(sep), idx);
					hit = 1;
				}
			}
		}
		const bool prim = !(cA.m_shapeType == B3H_SHAPE_CONVEX_HULL && cB.m_shapeType == B3H_SHAPE_CONVEX_HULL);
		if (g.lane == 0 && !prim)
		{
			flags[k] = hit;
			if (hit)
			{
				b3hContact4Data c;
				for (int q = 0; q < 4; q++) c.m_worldPosB[q] = q < nReduced ? pts[idx[q]] : f4(0, 0, 0, 0);
				c.m_worldNormalOnB = sep;
				c.m_worldNormalOnB.w = (float)nReduced;
				c.m_restituitionCoeffCmp = (unsigned short)(0.f * 0xffff);
				c.m_frictionCoeffCmp = (unsigned short)(0.7f * 0xffff);
				c.m_batchIdx = k;
				c.m_bodyAPtrAndSignBit = BA.m_invMass == 0 ? -cp.x : cp.x;
				c.m_bodyBPtrAndSignBit = BB.m_invMass == 0 ? -cp.y : cp.y;
				c.m_childIndexA = cp.z;
				c.m_childIndexB = cp.w;
				c.m_unused1 = 0;
				c.m_unused2 = 0;
				staging[k] = c;
			}
		}
		if (g.lane == 0 && prim) flags[k] = 0;
	}
}
/* processCompoundPairsPrimitivesKernel (primitiveContacts.cl:980-1090): a
 * child hull against a plane or a sphere body, one child pair per thread; the
 * contact carries the child pair's index and child indices -1 */
__global__ __launch_bounds__(PC_THREADS) void k_compound_primitive(const b3hInt4* __restrict__ cpairs,
																   const int* __restrict__ numPtr,
																   const b3hRigidBodyData* __restrict__ bodies,
																   b3hShapeTables sh, b3hContact4Data* __restrict__ staging,
																   int* __restrict__ flags)
{
	__shared__ int sIdx[PC_MAX_PLANE_POINTS * PC_THREADS];
	int* col = sIdx + threadIdx.x;
	const int n = *numPtr;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
	{
		const b3hInt4 cp = cpairs[k];
		int colA, colB;
		F4 posA, ornA, posB, ornB;
		compound_side(sh, bodies[cp.x], cp.z, &colA, &posA, &ornA);
		compound_side(sh, bodies[cp.y], cp.w, &colB, &posB, &ornB);
		const int ta = sh.collidables[colA].m_shapeType, tb = sh.collidables[colB].m_shapeType;
		b3hContact4Data c;
		bool hit = false, mine = true;
		if (ta == B3H_SHAPE_PLANE && tb == B3H_SHAPE_CONVEX_HULL)
			hit = pc_plane_convex(k, cp.x, cp.y, bodies, sh, col, c, colB, posB, ornB);
		else if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_PLANE)
			hit = pc_plane_convex(k, cp.y, cp.x, bodies, sh, col, c, colA, posA, ornA);
		else if (ta == B3H_SHAPE_CONVEX_HULL && tb == B3H_SHAPE_SPHERE)
			hit = pc_sphere_convex(k, cp.y, cp.x, bodies, sh, c, colA, posA, ornA);
		else if (ta == B3H_SHAPE_SPHERE && tb == B3H_SHAPE_CONVEX_HULL)
			hit = pc_sphere_convex(k, cp.x, cp.y, bodies, sh, c, colB, posB, ornB);
		else
			mine = false;
		if (!mine) continue;
		flags[k] = hit ? 1 : 0;
		if (hit) staging[k] = c;
	}
}
/* compound contacts after the pair contacts: slot base + scan[k] for flagged
 * child pair k (staging index stagingBase + k); counts[0] grows by their
 * number (capped at maxContacts), counts[1] (contacts produced) likewise */
__global__ __launch_bounds__(256) void k_compound_compact(const int* __restrict__ numPtr, const int* __restrict__ flags,
														  const int* __restrict__ scan, int* __restrict__ contactPair,
														  int stagingBase, int maxContacts, int* __restrict__ numContacts)
{
	int n = *numPtr;
	int base = numContacts[1];
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
	{
		if (!flags[k]) continue;
		int slot = base + scan[k];
		if (slot < maxContacts) contactPair[slot] = stagingBase + k;
	}
}
__global__ void k_compound_count_contacts(const int* __restrict__ numPtr, const int* __restrict__ scan, int maxContacts,
										  int* __restrict__ numContacts)
{
	int total = numContacts[1] + scan[*numPtr];
	numContacts[1] = total;
	numContacts[0] = total < maxContacts ? total : maxContacts;
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
					 int pairsPerGroup, hipStream_t s, hipEvent_t evAfterKernel, b3hFloat4* mprScratch,
					 hipEvent_t evAfterMpr, bool hasPrimitives, const b3hSatCache* satCache, const b3hSatWork& work,
					 const b3hCompoundWork& cw)
{
	{
		int mb = (maxPairs + 255) / 256;
		if (mb > 16384) mb = 16384;
		if (mb < 1) mb = 1;
		hipLaunchKernelGGL(k_np_mpr, dim3(mb), dim3(256), 0, s, pairs, numPairs, bodies, sh, (F4*)mprScratch);
	}
	if (evAfterMpr) hipEventRecord(evAfterMpr, s);
	b3hSatCache sc = *satCache;
	if (sc.numBodies > 0)
	{
		int nb = (sc.numBodies + 255) / 256;
		if (nb > 16384) nb = 16384;
		hipLaunchKernelGGL(k_sat_bodies, dim3(nb), dim3(256), 0, s, bodies, sh, sc);
	}
	if (sc.recs && sc.total > 0)
	{
		int fb = (sc.total + 255) / 256;
		if (fb > 16384) fb = 16384;
		hipLaunchKernelGGL(k_sat_faces, dim3(fb), dim3(256), 0, s, sh, sc);
	}
	if (sc.dirs && sc.numDirBodies > 0)
	{
		int db = (sc.numDirBodies * B3_UNIT_SPHERE_162_COUNT + SD_TILE - 1) / SD_TILE;
		if (db > 16384) db = 16384;
		hipLaunchKernelGGL(k_sat_dirs, dim3(db), dim3(SD_TILE), 0, s, sh, sc);
	}
	/* grid: four times the blocks the CUs hold at four waves per SIMD.  The
	 * blocks loop over pairs, so a larger grid only adds block start-ups, each
	 * of which writes its loop-invariant register spills to fresh scratch
	 * lines (HBM traffic); the factor four leaves the block scheduler room to
	 * balance pairs of unequal cost. */
	const int maxBlocks256 = 16 * b3h_cu_count();
	auto blocksFor = [&](int items, int perBlock, int threads) {
		int blocks = (items + perBlock - 1) / perBlock;
		int cap = maxBlocks256 * (256 / threads);
		if (blocks > cap) blocks = cap;
		return blocks < 1 ? 1 : blocks;
	};
	/* pass 1, faces: lanes per pair 64 (one pair per wavefront), 32 or 16 */
	auto facePass = [&](auto kern, int lanes, int threads) {
		hipLaunchKernelGGL(kern, dim3(blocksFor(maxPairs, threads / lanes, threads)), dim3(threads), 0, s, pairs,
						   numPairs, sh, staging, flags, sepState, sepNormal, (const F4*)mprScratch, sc, work.dmin,
						   work.flag);
	};
	if (pairsPerGroup == 16)
		facePass(k_sat_face_pass<16, 128>, 16, 128);
	else if (pairsPerGroup == 32)
		facePass(k_sat_face_pass<32, 256>, 32, 256);
	else
		facePass(k_sat_face_pass<64, NP_THREADS>, 64, NP_THREADS);
	/* pass 2, edges, over the open pairs */
	int lb = blocksFor(maxPairs, 256, 256);
	b3h_exclusive_scan(work.flag, work.scan, numPairs, maxPairs, scanTmp, s);
	hipLaunchKernelGGL(k_pair_list, dim3(lb), dim3(256), 0, s, numPairs, work.flag, work.scan, work.list, work.count);
	hipLaunchKernelGGL((k_sat_edge_pass<32, 256>), dim3(blocksFor(maxPairs, 8, 256)), dim3(256), 0, s, pairs, sh, staging, flags,
					   sepState, sepNormal, (const F4*)mprScratch, sc, (const float*)work.dmin, (const int*)work.count,
					   (const int*)work.list, work.flag);
	/* pass 3, clipping, over the pairs pass 2 flagged (the flags of pairs it
	 * did not visit are still 0 from pass 1) */
	b3h_exclusive_scan(work.flag, work.scan, numPairs, maxPairs, scanTmp, s);
	hipLaunchKernelGGL(k_pair_list, dim3(lb), dim3(256), 0, s, numPairs, work.flag, work.scan, work.list, work.count);
	hipLaunchKernelGGL((k_sat_clip_pass<16, 256>), dim3(blocksFor(maxPairs, 16, 256)), dim3(256), 0, s, pairs,
					   (const int*)work.count, (const int*)work.list, sh, staging, flags, sepState, sepNormal,
					   (const F4*)mprScratch, sc);
	if (hasPrimitives)
	{
		int pb = (maxPairs + PC_THREADS - 1) / PC_THREADS;
		if (pb > 16384) pb = 16384;
		if (pb < 1) pb = 1;
		hipLaunchKernelGGL(k_np_primitive, dim3(pb), dim3(PC_THREADS), 0, s, pairs, numPairs, bodies, sh, staging,
						   flags);
	}
	if (cw.cap > 0)
	{
		/* compound child pairs: list, SAT and clipping, primitives */
		int pb = blocksFor(maxPairs, 256, 256);
		hipLaunchKernelGGL(k_compound_count, dim3(pb), dim3(256), 0, s, pairs, numPairs, bodies, sh, cw.counts);
		b3h_exclusive_scan(cw.counts, cw.offsets, numPairs, maxPairs, cw.scanTmp, s);
		hipLaunchKernelGGL(k_compound_emit, dim3(pb), dim3(256), 0, s, pairs, numPairs, bodies, sh,
						   (const int*)cw.offsets, cw.cap, cw.cpairs, cw.num);
		/* two child pairs per wavefront: most children are boxes, whose
		 * families fill a fraction of 64 lanes */
		hipLaunchKernelGGL((k_compound_sat<32, 256>), dim3(blocksFor(cw.cap, 8, 256)), dim3(256), 0, s,
						   (const b3hInt4*)cw.cpairs, (const int*)cw.num, bodies, sh, staging + cw.stagingBase,
						   cw.flags);
		if (hasPrimitives)
		{
			int cb = (cw.cap + PC_THREADS - 1) / PC_THREADS;
			if (cb > 16384) cb = 16384;
			hipLaunchKernelGGL(k_compound_primitive, dim3(cb), dim3(PC_THREADS), 0, s, (const b3hInt4*)cw.cpairs,
							   (const int*)cw.num, bodies, sh, staging + cw.stagingBase, cw.flags);
		}
	}
	if (evAfterKernel) hipEventRecord(evAfterKernel, s);
	b3h_exclusive_scan(flags, flagScan, numPairs, maxPairs, scanTmp, s);
	int cblocks = (maxPairs + 255) / 256;
	if (cblocks > 16384) cblocks = 16384;
	if (cblocks < 1) cblocks = 1;
	hipLaunchKernelGGL(k_np_compact, dim3(cblocks), dim3(256), 0, s, pairs, numPairs, flags, flagScan, contactPair,
					   maxContacts);
	hipLaunchKernelGGL(k_np_count, dim3(1), dim3(1), 0, s, numPairs, flagScan, maxContacts, numContacts, numContacts + 1);
	if (cw.cap > 0)
	{
		/* the child pairs' contacts after the pair contacts, in child-pair order */
		b3h_exclusive_scan(cw.flags, cw.scan, cw.num, cw.cap, cw.scanTmp, s);
		int cb = (cw.cap + 255) / 256;
		if (cb > 16384) cb = 16384;
		hipLaunchKernelGGL(k_compound_compact, dim3(cb), dim3(256), 0, s, (const int*)cw.num, (const int*)cw.flags,
						   (const int*)cw.scan, contactPair, cw.stagingBase, maxContacts, numContacts);
		hipLaunchKernelGGL(k_compound_count_contacts, dim3(1), dim3(1), 0, s, (const int*)cw.num, (const int*)cw.scan,
						   maxContacts, numContacts);
	}
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
