/*
 * b3h_math.h -- float4 / quaternion helpers for the HIP kernels and host code.
 *
 * Float conventions (the whole library is built with -ffp-contract=off):
 * left-to-right evaluation, dot3 = (x*x' + y*y') + z*z', cross3 as the OpenCL
 * builtin, correctly rounded division and square root, and short polynomial
 * sine/cosine for the integration half angle.  Fixing the operation order makes
 * every kernel reproducible run to run and lets the results be compared bit for
 * bit against a scalar CPU evaluation of the same formulas.
 */
#ifndef B3H_MATH_H
#define B3H_MATH_H

#include <hip/hip_runtime.h>
#include <float.h>
#include "../../include/b3hip_types.h"

#define B3H_HD __host__ __device__ __forceinline__

typedef b3hFloat4 F4;

B3H_HD F4 f4(float x, float y, float z, float w)
{
	F4 r;
	r.x = x;
	r.y = y;
	r.z = z;
	r.w = w;
	return r;
}
B3H_HD F4 add4(F4 a, F4 b) { return f4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
B3H_HD F4 sub4(F4 a, F4 b) { return f4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
B3H_HD F4 mul4(F4 a, float s) { return f4(a.x * s, a.y * s, a.z * s, a.w * s); }
B3H_HD F4 smul4(float s, F4 a) { return f4(s * a.x, s * a.y, s * a.z, s * a.w); }
B3H_HD F4 neg4(F4 a) { return f4(-a.x, -a.y, -a.z, -a.w); }
/* OpenCL C's common max/min: max(x,y) = x < y ? y : x, min(x,y) = y < x ? y : x
 * (this fixes which zero a clamp at +-0 returns) */
B3H_HD float cl_max(float x, float y) { return x < y ? y : x; }
B3H_HD float cl_min(float x, float y) { return y < x ? y : x; }
B3H_HD float dot3(F4 a, F4 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
B3H_HD float dot4(F4 a, F4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
B3H_HD F4 cross3(F4 a, F4 b)
{
	return f4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.f);
}
B3H_HD F4 xyz0(F4 a) { return f4(a.x, a.y, a.z, 0.f); }

/* Hamilton product with the cross/dot split of the OpenCL pipeline's qtMul */
B3H_HD F4 qmul(F4 a, F4 b)
{
	F4 c = cross3(a, b);
	F4 r;
	r.x = c.x + (a.w * b.x + b.w * a.x);
	r.y = c.y + (a.w * b.y + b.w * a.y);
	r.z = c.z + (a.w * b.z + b.w * a.z);
	r.w = a.w * b.w - dot3(a, b);
	return r;
}
B3H_HD F4 qinv(F4 q) { return f4(-q.x, -q.y, -q.z, q.w); }
/* rotate v by q: q * (v,0) * conj(q) */
B3H_HD F4 qrot(F4 q, F4 v) { return qmul(qmul(q, xyz0(v)), qinv(q)); }
B3H_HD F4 qinvrot(F4 q, F4 v) { return qrot(qinv(q), v); }
B3H_HD F4 xform(F4 p, F4 pos, F4 q) { return add4(qrot(q, p), pos); }
/* b3TransformAabb2 (shared/b3Aabb.h:24-43) with margin 0 and the rotation
 * matrix of b3QuatGetRotationMatrix (shared/b3Mat3x3.h:41-61) */
B3H_HD void transform_aabb2(F4 lmin, F4 lmax, F4 pos, F4 q, F4* mn, F4* mx)
{
	F4 half = smul4(0.5f, sub4(lmax, lmin));
	half = add4(half, f4(0.f, 0.f, 0.f, 0.f));
	F4 center = smul4(0.5f, add4(lmax, lmin));
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
	F4 a0 = f4(fabsf(r0.x), fabsf(r0.y), fabsf(r0.z), 0.f);
	F4 a1 = f4(fabsf(r1.x), fabsf(r1.y), fabsf(r1.z), 0.f);
	F4 a2 = f4(fabsf(r2.x), fabsf(r2.y), fabsf(r2.z), 0.f);
	F4 c = xform(center, pos, q);
	F4 ext = f4(dot3(half, a0), dot3(half, a1), dot3(half, a2), 0.f);
	*mn = sub4(c, ext);
	*mx = add4(c, ext);
}
/* b3TestAabbAgainstAabb (shared/b3Aabb.h:46-54) */
B3H_HD bool aabb_overlap(F4 mn1, F4 mx1, F4 mn2, F4 mx2)
{
	bool overlap = true;
	overlap = (mn1.x > mx2.x || mx1.x < mn2.x) ? false : overlap;
	overlap = (mn1.z > mx2.z || mx1.z < mn2.z) ? false : overlap;
	overlap = (mn1.y > mx2.y || mx1.y < mn2.y) ? false : overlap;
	return overlap;
}
B3H_HD F4 nrm3(F4 v)
{
	float l2 = dot3(v, v);
	float k = 1.f / sqrtf(l2);
	return f4(v.x * k, v.y * k, v.z * k, 0.f);
}

/* sine / cosine on |x| <= pi/8: Taylor polynomials to x^9 and x^8 in Horner
 * form, truncation error below 1e-10 (under float resolution). */
B3H_HD float b3h_sin_small(float x)
{
	float x2 = x * x;
	float p = -1.f / 5040.f + x2 * (1.f / 362880.f);
	p = 1.f / 120.f + x2 * p;
	p = -1.f / 6.f + x2 * p;
	return x + (x * x2) * p;
}
B3H_HD float b3h_cos_small(float x)
{
	float x2 = x * x;
	float p = -1.f / 720.f + x2 * (1.f / 40320.f);
	p = 1.f / 24.f + x2 * p;
	p = -0.5f + x2 * p;
	return 1.f + x2 * p;
}

B3H_HD int f2i(float f) { return __builtin_bit_cast(int, f); }
B3H_HD float i2f(int i) { return __builtin_bit_cast(float, i); }
B3H_HD unsigned f2u(float f) { return __builtin_bit_cast(unsigned, f); }

/* matrix * column vector */
B3H_HD F4 mtMul1(const b3hMat3x3& a, F4 b)
{
	return f4(dot3(a.m_row[0], b), dot3(a.m_row[1], b), dot3(a.m_row[2], b), 0.f);
}
/* row vector * matrix */
B3H_HD F4 mtMul3(F4 a, const b3hMat3x3& b)
{
	F4 colx = f4(b.m_row[0].x, b.m_row[1].x, b.m_row[2].x, 0);
	F4 coly = f4(b.m_row[0].y, b.m_row[1].y, b.m_row[2].y, 0);
	F4 colz = f4(b.m_row[0].z, b.m_row[1].z, b.m_row[2].z, 0);
	return f4(dot3(a, colx), dot3(a, coly), dot3(a, colz), 0.f);
}

/* 16-byte vector loads of the AoS records */
__device__ __forceinline__ F4 ld4(const F4* p) { return *p; }

/* integrateSingleTransform (b3IntegrateTransforms.h:5-55): angular damping,
 * quaternion step (pi/4 clamp, Taylor branch below 0.001 rad/s), explicit
 * Euler position then gravity.  Static bodies (invMass 0) are left alone. */
B3H_HD void b3h_integrate_body(b3hRigidBodyData& b, float dt, float damp, F4 g)
{
	if (b.m_invMass == 0.f) return;
	const float thr = 0.25f * 3.14159254f;
	b.m_angVel.x *= damp;
	b.m_angVel.y *= damp;
	b.m_angVel.z *= damp;
	F4 av = b.m_angVel;
	float ang = sqrtf(dot3(av, av));
	if (ang * dt > thr) ang = thr / dt;
	F4 axis;
	if (ang < 0.001f)
		axis = mul4(av, 0.5f * dt - (dt * dt * dt) * 0.020833333333f * ang * ang);
	else
		axis = mul4(av, b3h_sin_small(0.5f * ang * dt) / ang);
	F4 dorn = f4(axis.x, axis.y, axis.z, b3h_cos_small(ang * dt * 0.5f));
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

/* b3PlaneSpace1: two unit tangents orthogonal to the unit normal n */
B3H_HD void plane_space(F4 n, F4* p, F4* q)
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


/* arc cosine of x in [-1, 1], from the arcsine series
 * asin(y) = sum_n (2n)! / (4^n (n!)^2 (2n+1)) y^(2n+1) on |y| <= 1/2 (terms
 * shrink at least fourfold, 30 of them reach double precision) with
 * acos(x) = 2 asin(sqrt((1-x)/2)) for x >= 1/2, pi - 2 asin(sqrt((1+x)/2)) for
 * x <= -1/2 and pi/2 - asin(x) between; evaluated in double and rounded once,
 * so host and device give the same float */
B3H_HD double b3h_asin_series(double y)
{
	double y2 = y * y, term = y, sum = y;
	for (int n = 1; n <= 30; n++)
	{
		term = term * (y2 * (double)((2 * n - 1) * (2 * n - 1)) / (double)((2 * n) * (2 * n + 1)));
		sum = sum + term;
	}
	return sum;
}
B3H_HD float b3h_acos(float xf)
{
	const double pi = 3.141592653589793;
	double x = (double)xf;
	if (x >= 0.5) return (float)(2.0 * b3h_asin_series(sqrt((1.0 - x) * 0.5)));
	if (x <= -0.5) return (float)(pi - 2.0 * b3h_asin_series(sqrt((1.0 + x) * 0.5)));
	return (float)(pi * 0.5 - b3h_asin_series(x));
}

#endif
