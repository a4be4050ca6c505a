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
	F4 mn, mx;
	transform_aabb2(la.m_min, la.m_max, pos, q, &mn, &mx);
	b3hAabb w;
	w.m_min = mn;
	w.m_max = mx;
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
