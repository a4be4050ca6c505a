/*
 * ORACLE -- test infrastructure only.
 * Harness that exposes pieces of the reference's own C++ code (compiled from
 * /root/reference by oracle/ref.mk into oracle/_ref/libb3ref.so) through a C
 * interface, so tests can pin the oracle against them:
 *   b3MprPenetration      src/Bullet3Collision/NarrowPhaseCollision/shared/b3MprPenetration.h:816-886
 *   integrateSingleTransform  src/Bullet3Dynamics/shared/b3IntegrateTransforms.h:5-55
 *   b3ComputeWorldAabb    src/Bullet3Collision/NarrowPhaseCollision/shared/b3UpdateAabbs.h:8-33
 *   setConstraint4        src/Bullet3Dynamics/shared/b3ConvertConstraint4.h:62-148
 *   b3ConvexUtility::initializePolyhedralFeatures  src/Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.cpp:26-248
 *   registerRigidBody's inertia  src/Bullet3OpenCL/RigidBody/b3GpuNarrowPhase.cpp:860-901, evaluated with
 *                         the reference's b3Vector3 / b3Matrix3x3 / b3Quaternion classes
 *   registerCompoundShape's local AABB and b3QuantizedBvh  src/Bullet3OpenCL/RigidBody/b3GpuNarrowPhase.cpp:370-517
 * The headers are the reference's dual-compile ("shared") headers, built here
 * as the C++ they are written to be.
 */
#include "Bullet3Common/b3Vector3.h"
#include "Bullet3Common/b3Quaternion.h"
#include "Bullet3Common/b3Matrix3x3.h"
#include "Bullet3Collision/NarrowPhaseCollision/shared/b3RigidBodyData.h"
#include "Bullet3Collision/NarrowPhaseCollision/shared/b3Collidable.h"
#include "Bullet3Collision/NarrowPhaseCollision/shared/b3ConvexPolyhedronData.h"
#include "Bullet3Collision/NarrowPhaseCollision/shared/b3MprPenetration.h"
#include "Bullet3Collision/NarrowPhaseCollision/shared/b3UpdateAabbs.h"
#include "Bullet3Collision/NarrowPhaseCollision/shared/b3Contact4Data.h"
#include "Bullet3Dynamics/shared/b3IntegrateTransforms.h"
#include "Bullet3Dynamics/shared/b3ConvertConstraint4.h"
#include "Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.h"
#include "Bullet3Common/b3Transform.h"
#include "Bullet3Geometry/b3AabbUtil.h"
#include "Bullet3OpenCL/NarrowphaseCollision/b3QuantizedBvh.h"
#include <string.h>

extern "C" {

/* per pair: res (0 contact, -1 none), depth, dir[4], pos[4] */
int ref_mpr(const int* pairs4, int nPairs, const void* bodies, const void* cols, const void* convexes,
			const void* verts, int* resOut, float* depthOut, float* dirOut, float* posOut)
{
	const b3RigidBodyData* b = (const b3RigidBodyData*)bodies;
	const b3Collidable* c = (const b3Collidable*)cols;
	const b3ConvexPolyhedronData* h = (const b3ConvexPolyhedronData*)convexes;
	const b3Float4* v = (const b3Float4*)verts;
	b3AlignedObjectArray<b3Float4> sep;
	sep.resize(nPairs);
	b3AlignedObjectArray<int> has;
	has.resize(nPairs);
	int count = 0;
	for (int i = 0; i < nPairs; i++)
	{
		int a = pairs4[4 * i], bb = pairs4[4 * i + 1];
		resOut[i] = -1;
		if (b[a].m_invMass == 0 && b[bb].m_invMass == 0) continue;
		if (c[b[a].m_collidableIdx].m_shapeType != SHAPE_CONVEX_HULL || c[b[bb].m_collidableIdx].m_shapeType != SHAPE_CONVEX_HULL) continue;
		float depth = 0;
		b3Float4 dir, pos;
		int r = b3MprPenetration(i, a, bb, b, h, c, v, &sep[0], &has[0], &depth, &dir, &pos);
		resOut[i] = r;
		if (r == 0)
		{
			count++;
			depthOut[i] = depth;
			for (int k = 0; k < 4; k++)
			{
				dirOut[4 * i + k] = dir[k];
				posOut[4 * i + k] = pos[k];
			}
		}
	}
	return count;
}

void ref_integrate(void* bodies, int n, float dt, float damp, const float* g3)
{
	b3RigidBodyData* b = (b3RigidBodyData*)bodies;
	b3Vector3 g = b3MakeVector3(g3[0], g3[1], g3[2]);
	for (int i = 0; i < n; i++) integrateSingleTransform(b, i, dt, damp, g);
}

void ref_world_aabbs(int n, const void* bodies, const void* cols, const void* local, void* world)
{
	for (int i = 0; i < n; i++)
		b3ComputeWorldAabb(i, (const b3RigidBodyData*)bodies, (const b3Collidable*)cols, (const b3Aabb*)local, (b3Aabb*)world);
}

/* one contact -> one constraint (inverse inertias given as 3x4 row blocks) */
void ref_set_constraint4(const void* bodyA, const void* invInertiaA, const void* bodyB, const void* invInertiaB,
						 const void* contact, float dt, float drift, float coeff, void* out)
{
	const b3RigidBodyData* A = (const b3RigidBodyData*)bodyA;
	const b3RigidBodyData* B = (const b3RigidBodyData*)bodyB;
	b3ContactConstraint4_t cs;
	memset(&cs, 0, sizeof(cs));
	setConstraint4(A->m_pos, A->m_linVel, A->m_angVel, A->m_invMass, *(const b3Matrix3x3*)invInertiaA,
				   B->m_pos, B->m_linVel, B->m_angVel, B->m_invMass, *(const b3Matrix3x3*)invInertiaB,
				   (b3Contact4Data*)contact, dt, drift, coeff, &cs);
	memcpy(out, &cs, sizeof(cs));
}

/* Hull features of a point cloud.  Writes counts[0..3] = numVertices,
 * numFaces, numIndices, numUniqueEdges; arrays sized by the caller
 * (vertices 4*cap, planes 4*cap, faceIndexCounts cap, indices cap, edges 4*cap). */
int ref_hull(const float* pts3, int n, int cap, int* counts, float* vertsOut, float* planesOut, int* faceCountsOut,
			 int* indicesOut, float* edgesOut)
{
	b3AlignedObjectArray<b3Vector3> v;
	for (int i = 0; i < n; i++) v.push_back(b3MakeVector3(pts3[3 * i], pts3[3 * i + 1], pts3[3 * i + 2]));
	b3ConvexUtility u;
	u.initializePolyhedralFeatures(&v[0], n, true);
	int nv = u.m_vertices.size(), nf = u.m_faces.size(), ne = u.m_uniqueEdges.size();
	int ni = 0;
	for (int f = 0; f < nf; f++) ni += u.m_faces[f].m_indices.size();
	counts[0] = nv; counts[1] = nf; counts[2] = ni; counts[3] = ne;
	if (nv > cap || nf > cap || ni > cap || ne > cap) return -1;
	for (int i = 0; i < nv; i++)
		for (int k = 0; k < 4; k++) vertsOut[4 * i + k] = k < 3 ? u.m_vertices[i][k] : 0.f;
	int w = 0;
	for (int f = 0; f < nf; f++)
	{
		for (int k = 0; k < 4; k++) planesOut[4 * f + k] = u.m_faces[f].m_plane[k];
		faceCountsOut[f] = u.m_faces[f].m_indices.size();
		for (int j = 0; j < u.m_faces[f].m_indices.size(); j++) indicesOut[w++] = u.m_faces[f].m_indices[j];
	}
	for (int i = 0; i < ne; i++)
		for (int k = 0; k < 4; k++) edgesOut[4 * i + k] = k < 3 ? u.m_uniqueEdges[i][k] : 0.f;
	return 0;
}

/* b3InertiaData of a body registered with this mass, orientation and
 * (margin-inflated) AABB: the statements of b3GpuNarrowPhase::registerRigidBody
 * (b3GpuNarrowPhase.cpp:860-901) on the reference's own math classes. */
void ref_inertia(float mass, const float* aabbMinPtr, const float* aabbMaxPtr, const float* orn, void* out)
{
	b3InertiaData shapeInfo;
	memset(&shapeInfo, 0, sizeof(shapeInfo));
	if (mass == 0.f)
	{
		shapeInfo.m_initInvInertia.setValue(0, 0, 0, 0, 0, 0, 0, 0, 0);
		shapeInfo.m_invInertiaWorld.setValue(0, 0, 0, 0, 0, 0, 0, 0, 0);
	}
	else
	{
		b3Vector3 aabbMin = b3MakeVector3(aabbMinPtr[0], aabbMinPtr[1], aabbMinPtr[2]);
		b3Vector3 aabbMax = b3MakeVector3(aabbMaxPtr[0], aabbMaxPtr[1], aabbMaxPtr[2]);
		b3Vector3 halfExtents = (aabbMax - aabbMin);
		b3Vector3 localInertia;
		float lx = 2.f * halfExtents[0];
		float ly = 2.f * halfExtents[1];
		float lz = 2.f * halfExtents[2];
		localInertia.setValue((mass / 12.0f) * (ly * ly + lz * lz), (mass / 12.0f) * (lx * lx + lz * lz),
							  (mass / 12.0f) * (lx * lx + ly * ly));
		b3Vector3 invLocalInertia;
		invLocalInertia[0] = 1.f / localInertia[0];
		invLocalInertia[1] = 1.f / localInertia[1];
		invLocalInertia[2] = 1.f / localInertia[2];
		invLocalInertia[3] = 0.f;
		shapeInfo.m_initInvInertia.setValue(invLocalInertia[0], 0, 0, 0, invLocalInertia[1], 0, 0, 0, invLocalInertia[2]);
		b3Quaternion q(orn[0], orn[1], orn[2], orn[3]);
		b3Matrix3x3 m(q);
		shapeInfo.m_invInertiaWorld = m.scaled(invLocalInertia) * m.transpose();
	}
	memcpy(out, &shapeInfo, sizeof(shapeInfo));
}

/* broadphase AABB of registerPhysicsInstance (b3GpuRigidBodyPipeline.cpp:605-618) */
void ref_instance_aabb(const float* lmin, const float* lmax, const float* pos, const float* orn, float* mn, float* mx)
{
	b3Vector3 localAabbMin = b3MakeVector3(lmin[0], lmin[1], lmin[2]);
	b3Vector3 localAabbMax = b3MakeVector3(lmax[0], lmax[1], lmax[2]);
	b3Scalar margin = 0.01f;
	b3Transform t;
	t.setIdentity();
	t.setOrigin(b3MakeVector3(pos[0], pos[1], pos[2]));
	t.setRotation(b3Quaternion(orn[0], orn[1], orn[2], orn[3]));
	b3Vector3 aabbMin, aabbMax;
	b3TransformAabb(localAabbMin, localAabbMax, margin, t, aabbMin, aabbMax);
	for (int k = 0; k < 3; k++)
	{
		mn[k] = aabbMin[k];
		mx[k] = aabbMax[k];
	}
}

/* registerCompoundShape (b3GpuNarrowPhase.cpp:370-517) for n children:
 * childPos/childOrn (4 floats each), childMin/childMax the children's local
 * AABBs (4 floats each).  Out: the compound AABB (min, max: 3 floats each),
 * the tree's aabbMin, aabbMax, quantization (3 floats each), its 2n nodes (16
 * bytes each; the last one is never written by the build and is zeroed here)
 * and its subtree headers (32 bytes each, at most subCap). */
int ref_compound_bvh(int n, const float* childPos, const float* childOrn, const float* childMin, const float* childMax,
					 float* aabbMin, float* aabbMax, float* bvhMin, float* bvhMax, float* bvhQuant, void* nodesOut,
					 void* subtreesOut, int subCap)
{
	b3Vector3 myAabbMin = b3MakeVector3(1e30f, 1e30f, 1e30f);
	b3Vector3 myAabbMax = b3MakeVector3(-1e30f, -1e30f, -1e30f);
	b3AlignedObjectArray<b3Vector3> cmin, cmax;
	cmin.resize(n);
	cmax.resize(n);
	for (int i = 0; i < n; i++)
	{
		b3Vector3 lmin = b3MakeVector3(childMin[4 * i], childMin[4 * i + 1], childMin[4 * i + 2]);
		b3Vector3 lmax = b3MakeVector3(childMax[4 * i], childMax[4 * i + 1], childMax[4 * i + 2]);
		b3Transform childTr;
		childTr.setIdentity();
		childTr.setOrigin(b3MakeVector3(childPos[4 * i], childPos[4 * i + 1], childPos[4 * i + 2]));
		childTr.setRotation(b3Quaternion(childOrn[4 * i], childOrn[4 * i + 1], childOrn[4 * i + 2], childOrn[4 * i + 3]));
		b3Vector3 aMin, aMax;
		b3Scalar margin(0.f);
		b3TransformAabb(lmin, lmax, margin, childTr, aMin, aMax);
		myAabbMin.setMin(aMin);
		myAabbMax.setMax(aMax);
		cmin[i] = aMin;
		cmax[i] = aMax;
	}
	for (int k = 0; k < 3; k++)
	{
		aabbMin[k] = myAabbMin[k];
		aabbMax[k] = myAabbMax[k];
	}
	b3QuantizedBvh* bvh = new b3QuantizedBvh;
	bvh->setQuantizationValues(myAabbMin, myAabbMax);
	QuantizedNodeArray& nodes = bvh->getLeafNodeArray();
	for (int i = 0; i < n; i++)
	{
		b3QuantizedBvhNode node;
		bvh->quantize(&node.m_quantizedAabbMin[0], cmin[i], 0);
		bvh->quantize(&node.m_quantizedAabbMax[0], cmax[i], 1);
		int partId = 0;
		node.m_escapeIndexOrTriangleIndex = (partId << (31 - MAX_NUM_PARTS_IN_BITS)) | i;
		nodes.push_back(node);
	}
	int numLeaves = n;
	bvh->buildInternal();
	for (int k = 0; k < 3; k++)
	{
		bvhMin[k] = bvh->m_bvhAabbMin[k];
		bvhMax[k] = bvh->m_bvhAabbMax[k];
		bvhQuant[k] = bvh->m_bvhQuantization[k];
	}
	const QuantizedNodeArray& all = bvh->getQuantizedNodeArray();
	memset(nodesOut, 0, sizeof(b3QuantizedBvhNode) * 2 * (size_t)numLeaves);
	memcpy(nodesOut, &all[0], sizeof(b3QuantizedBvhNode) * (2 * (size_t)numLeaves - 1));
	const BvhSubtreeInfoArray& st = bvh->getSubtreeInfoArray();
	int ns = st.size();
	for (int i = 0; i < ns && i < subCap; i++) memcpy((char*)subtreesOut + 32 * i, &st[i], 32);
	delete bvh;
	return ns;
}
}
