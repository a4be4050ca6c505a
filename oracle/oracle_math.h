/*
 * ORACLE -- test infrastructure only.  Scalar CPU restatement of the Bullet3
 * GPU rigid-body step, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in the product links it.
 *
 * Float conventions (identical to the device code so results can be compared
 * bit for bit): no FMA contraction (built with -ffp-contract=off), left-to-right
 * evaluation, dot3 = (x*x' + y*y') + z*z', cross3 as in the OpenCL builtin,
 * correctly rounded / and sqrtf, and the short polynomial sine/cosine below.
 *
 * What pins each part (tests/):
 *   world AABBs, integration      the reference's own C++ (oracle/_ref), bit for bit
 *                                 or within 2e-6; closed forms of free fall
 *   MPR                           the reference's known-answer kernel dumps
 *                                 (counts 0, 1, 46, 98) and its own C++ MPR
 *   SAT (faces, edge-edge, 162)   float64 brute force of the minimum overlap over
 *                                 the same candidate axes; box closed forms
 *   clipping + manifold reduction box/ground and tilted-box closed forms (points
 *                                 are the penetrating corners, depth -delta).
 *                                 Which 3 of 4 equal-depth corners the reduction
 *                                 keeps is parity unpinned (only the GPU kernel
 *                                 is compared with it, bit for bit)
 *   grid / SAP pairs              brute force, as sets; the stable key sort
 *                                 against numpy's stable argsort
 *   batching                      invariants (every contact batched, no shared
 *                                 dynamic body in a 64-row, cells partition)
 *   constraint setup              the reference's own C++ setConstraint4
 *   PGS sweeps                    closed forms: weight support, penetration bias,
 *                                 no pull, friction bound, momentum conservation
 *   OpenCL native_* rounding      parity unpinned: the reference kernels cannot
 *                                 run here; correctly rounded operations are used
 */
#ifndef B3ORACLE_MATH_H
#define B3ORACLE_MATH_H

#include <math.h>
#include <float.h>
#include <stdint.h>
#include <string.h>
#include "../include/b3hip_types.h"

typedef b3hFloat4 F4;

static inline F4 f4(float x, float y, float z, float w) { F4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
static inline F4 add4(F4 a, F4 b) { return f4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline F4 sub4(F4 a, F4 b) { return f4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static inline F4 mul4(F4 a, float s) { return f4(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline F4 smul4(float s, F4 a) { return f4(s * a.x, s * a.y, s * a.z, s * a.w); }
static inline F4 neg4(F4 a) { return f4(-a.x, -a.y, -a.z, -a.w); }
/* OpenCL C's common max/min: max(x,y) = x < y ? y : x, min(x,y) = y < x ? y : x
 * (this fixes which zero a clamp at +-0 returns) */
static inline float cl_max(float x, float y) { return x < y ? y : x; }
static inline float cl_min(float x, float y) { return y < x ? y : x; }
static inline float dot3(F4 a, F4 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float dot4(F4 a, F4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
static inline F4 cross3(F4 a, F4 b)
{
	return f4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.f);
}
static inline F4 xyz0(F4 a) { return f4(a.x, a.y, a.z, 0.f); }

/* b3QuatMul / qtMul: src/Bullet3Common/shared/b3Quat.h:43-51 */
static inline F4 qmul(F4 a, F4 b)
{
	F4 c = cross3(a, b);
	F4 r;
	r.x = c.x + (a.w * b.x + b.w * a.x);
	r.y = c.y + (a.w * b.y + b.w * a.y);
	r.z = c.z + (a.w * b.z + b.w * a.z);
	r.w = a.w * b.w - dot3(a, b);
	return r;
}
static inline F4 qinv(F4 q) { return f4(-q.x, -q.y, -q.z, q.w); }
/* b3QuatRotate / qtRotate: src/Bullet3Common/shared/b3Quat.h:68-75 */
static inline F4 qrot(F4 q, F4 v)
{
	return qmul(qmul(q, xyz0(v)), qinv(q));
}
static inline F4 qinvrot(F4 q, F4 v) { return qrot(qinv(q), v); }
/* b3TransformPoint / transform: src/Bullet3Common/shared/b3Quat.h:87-90 */
static inline F4 xform(F4 p, F4 pos, F4 q) { return add4(qrot(q, p), pos); }

/* normalize of a 3-vector (w forced to 0) */
static inline F4 nrm3(F4 v)
{
	float l2 = dot3(v, v);
	float k = 1.f / sqrtf(l2);
	return f4(v.x * k, v.y * k, v.z * k, 0.f);
}

/* Sine and cosine for |x| <= pi/8 (the integration step clamps the half angle
 * to that range): Taylor polynomials to x^9 / x^8, evaluated in Horner form.
 * Truncation error < 1e-10 on the range, i.e. below float resolution. */
static inline float sin_small(float x)
{
	float x2 = x * x;
	float p = -1.f / 5040.f + x2 * (1.f / 362880.f);
	p = 1.f / 120.f + x2 * p;
	p = -1.f / 6.f + x2 * p;
	return x + (x * x2) * p;
}
static inline float cos_small(float x)
{
	float x2 = x * x;
	float p = -1.f / 720.f + x2 * (1.f / 40320.f);
	p = 1.f / 24.f + x2 * p;
	p = -0.5f + x2 * p;
	return 1.f + x2 * p;
}

static inline int f2i(float f) { int i; memcpy(&i, &f, 4); return i; }
static inline float i2f(int i) { float f; memcpy(&f, &i, 4); return f; }
static inline unsigned f2u(float f) { unsigned i; memcpy(&i, &f, 4); return i; }

/* mtMul1: matrix * column vector (solveContact.cl:141-148) */
static inline F4 mtMul1(const b3hMat3x3& a, F4 b)
{
	return f4(dot3(a.m_row[0], b), dot3(a.m_row[1], b), dot3(a.m_row[2], b), 0.f);
}
/* mtMul3: row vector * matrix (b3Mat3x3.h:150-161) */
static inline F4 mtMul3(F4 a, const b3hMat3x3& b)
{
	F4 colx = f4(b.m_row[0].x, b.m_row[1].x, b.m_row[2].x, 0);
	F4 coly = f4(b.m_row[0].y, b.m_row[1].y, b.m_row[2].y, 0);
	F4 colz = f4(b.m_row[0].z, b.m_row[1].z, b.m_row[2].z, 0);
	return f4(dot3(a, colx), dot3(a, coly), dot3(a, colz), 0.f);
}


/* arc cosine of x in [-1, 1], from the arcsine series
 * asin(y) = sum_n (2n)! / (4^n (n!)^2 (2n+1)) y^(2n+1) on |y| <= 1/2 (terms
 * shrink at least fourfold, 30 of them reach double precision) with
 * acos(x) = 2 asin(sqrt((1-x)/2)) for x >= 1/2, pi - 2 asin(sqrt((1+x)/2)) for
 * x <= -1/2 and pi/2 - asin(x) between; evaluated in double and rounded once,
 * so host and device give the same float */
static inline double b3h_asin_series(double y)
{
	double y2 = y * y, term = y, sum = y;
	for (int n = 1; n <= 30; n++)
	{
		term = term * (y2 * (double)((2 * n - 1) * (2 * n - 1)) / (double)((2 * n) * (2 * n + 1)));
		sum = sum + term;
	}
	return sum;
}
static inline float b3h_acos(float xf)
{
	const double pi = 3.141592653589793;
	double x = (double)xf;
	if (x >= 0.5) return (float)(2.0 * b3h_asin_series(sqrt((1.0 - x) * 0.5)));
	if (x <= -0.5) return (float)(pi - 2.0 * b3h_asin_series(sqrt((1.0 + x) * 0.5)));
	return (float)(pi * 0.5 - b3h_asin_series(x));
}

#endif
