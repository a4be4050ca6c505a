/*
 * k_body.hip -- one lane per body: world-space AABBs and integration.
 *
 * World AABB: initializeGpuAabbsFull -> b3ComputeWorldAabb -> b3TransformAabb2
 * (src/Bullet3OpenCL/RigidBody/kernels/updateAabbsKernel.cl:6-13,
 *  src/Bullet3Collision/NarrowPhaseCollision/shared/b3UpdateAabbs.h:8-33,
 *  src/Bullet3Collision/BroadPhaseCollision/shared/b3Aabb.h:24-43), margin 0.
 *  min.w <- body index; max.w <- (invMass == 0 ? 0 : 1) of the body itself
 *  (the reference indexes body[bodyId] there, out of range; nothing reads it).
 * Integration: integrateTransformsKernel -> integrateSingleTransform
 * (src/Bullet3OpenCL/RigidBody/kernels/integrateKernel.cl:23-32,
 *  src/Bullet3Dynamics/shared/b3IntegrateTransforms.h:5-55).
 * Both are HBM-bound streams: 112 B/body read + 32 B written for the AABBs,
 * 80 B read + 80 B written for integration.
 */
#include "b3h_internal.h"

__global__ __launch_bounds__(256) void k_world_aabbs(int n, const b3hRigidBodyData* __restrict__ bodies,
													 const b3hCollidable* __restrict__ cols,
													 const b3hAabb* __restrict__ localAabbs,
													 b3hAabb* __restrict__ world)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const b3hRigidBodyData& b = bodies[i];
	F4 pos = b.m_pos, q = b.m_quat;
	int ci = b.m_collidableIdx;
	if (cols[ci].m_shapeIndex < 0) return;
	b3hAabb la = localAabbs[ci];
	F4 half = smul4(0.5f, sub4(la.m_max, la.m_min));
	half = add4(half, f4(0.f, 0.f, 0.f, 0.f));
	F4 center = smul4(0.5f, add4(la.m_max, la.m_min));
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
	b3hAabb w;
	w.m_min = sub4(c, ext);
	w.m_max = add4(c, ext);
	w.m_min.w = i2f(i);
	w.m_max.w = i2f(b.m_invMass == 0.f ? 0 : 1);
	world[i] = w;
}

void b3h_launch_world_aabbs(int n, const b3hRigidBodyData* bodies, const b3hCollidable* cols, const b3hAabb* localAabbs,
							b3hAabb* world, hipStream_t s)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(k_world_aabbs, dim3((n + 255) / 256), dim3(256), 0, s, n, bodies, cols, localAabbs, world);
}

__global__ __launch_bounds__(256) void k_integrate(int n, b3hRigidBodyData* __restrict__ bodies, float dt, float damp, F4 g,
													const int* __restrict__ skip)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	if (skip && skip[i] >= 0) return; /* integrated by the last PGS sweep */
	b3hRigidBodyData b = bodies[i];
	if (b.m_invMass == 0.f) return;
	b3h_integrate_body(b, dt, damp, g);
	bodies[i].m_pos = b.m_pos;
	bodies[i].m_quat = b.m_quat;
	bodies[i].m_linVel = b.m_linVel;
	bodies[i].m_angVel = b.m_angVel;
}

void b3h_launch_integrate(int n, b3hRigidBodyData* bodies, float dt, float damping, F4 gravity, hipStream_t s,
						  const int* skip)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(k_integrate, dim3((n + 255) / 256), dim3(256), 0, s, n, bodies, dt, damping, gravity, skip);
}
