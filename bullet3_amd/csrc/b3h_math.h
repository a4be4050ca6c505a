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

#endif
