/*
 * b3hip_types.h -- plain-old-data records exchanged across the b3hip C ABI.
 *
 * Every record has the byte layout of the Bullet3 GPU pipeline's shared
 * structs, so buffers produced by either implementation can be read by the
 * other (and by the reference's own tools):
 *
 *   b3hRigidBodyData        80 B  src/Bullet3Collision/NarrowPhaseCollision/shared/b3RigidBodyData.h:10-21
 *   b3hInertiaData          96 B  src/Bullet3Collision/NarrowPhaseCollision/shared/b3RigidBodyData.h:25-29
 *   b3hCollidable           16 B  src/Bullet3Collision/NarrowPhaseCollision/shared/b3Collidable.h:22-38
 *   b3hConvexPolyhedronData 96 B  src/Bullet3Collision/NarrowPhaseCollision/shared/b3ConvexPolyhedronData.h:21-36
 *   b3hGpuFace              32 B  src/Bullet3Collision/NarrowPhaseCollision/shared/b3ConvexPolyhedronData.h:7-14
 *   b3hAabb                 32 B  src/Bullet3Collision/BroadPhaseCollision/shared/b3Aabb.h:10-22
 *   b3hInt4 (pair)          16 B  src/Bullet3Collision/BroadPhaseCollision/b3OverlappingPair.h:24
 *   b3hContact4Data        112 B  src/Bullet3Collision/NarrowPhaseCollision/shared/b3Contact4Data.h:8-24
 *   b3hContactConstraint4  176 B  src/Bullet3Dynamics/shared/b3ContactConstraint4.h:8-23
 *   b3hSortData              8 B  src/Bullet3OpenCL/ParallelPrimitives/b3RadixSort32CL.h:7-18
 *   b3hConfig               48 B  src/Bullet3Collision/NarrowPhaseCollision/b3Config.h:4-37
 *
 * This header is C and C++ (host and HIP device) compatible.
 */
#ifndef B3HIP_TYPES_H
#define B3HIP_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct __attribute__((aligned(16))) b3hFloat4
{
	float x, y, z, w;
} b3hFloat4;

typedef struct __attribute__((aligned(16))) b3hInt4
{
	int x, y, z, w;
} b3hInt4;

typedef struct b3hMat3x3
{
	b3hFloat4 m_row[3];
} b3hMat3x3;

typedef struct b3hRigidBodyData
{
	b3hFloat4 m_pos;
	b3hFloat4 m_quat; /* x,y,z,w */
	b3hFloat4 m_linVel;
	b3hFloat4 m_angVel;
	int m_collidableIdx;
	float m_invMass;
	float m_restituitionCoeff;
	float m_frictionCoeff;
} b3hRigidBodyData;

typedef struct b3hInertiaData
{
	b3hMat3x3 m_invInertiaWorld;
	b3hMat3x3 m_initInvInertia;
} b3hInertiaData;

/* collidable shape types, shared/b3Collidable.h:8-18 (SHAPE_CONVEX_HULL = 3,
 * SHAPE_PLANE = 4, SHAPE_CONCAVE_TRIMESH = 5, SHAPE_COMPOUND_OF_CONVEX_HULLS = 6,
 * SHAPE_SPHERE = 7) */
enum b3hShapeTypes
{
	B3H_SHAPE_CONVEX_HULL = 3,
	B3H_SHAPE_PLANE = 4,
	B3H_SHAPE_CONCAVE_TRIMESH = 5,
	B3H_SHAPE_COMPOUND_OF_CONVEX_HULLS = 6,
	B3H_SHAPE_SPHERE = 7
};

typedef struct b3hCollidable
{
	int m_numChildShapes; /* compound: its children */
	union
	{
		float m_radius;         /* sphere radius, plane constant */
		int m_compoundBvhIndex; /* compound: its b3hBvhInfo */
	};
	int m_shapeType;
	int m_shapeIndex; /* hull: convex index; plane: face; compound: first child */
} b3hCollidable;

/* b3GpuChildShape, shared/b3Collidable.h:40-58: a compound's child (a
 * registered convex-hull collidable placed by a child transform) */
typedef struct b3hGpuChildShape
{
	b3hFloat4 m_childPosition;
	b3hFloat4 m_childOrientation;
	int m_shapeIndex; /* the child's collidable */
	int m_numChildShapes;
	int m_collidableShapeIndex;
	int m_shapeType;
} b3hGpuChildShape;

/* the quantized AABB tree over a compound's children (b3QuantizedBvhNodeData,
 * b3BvhSubtreeInfoData, b3BvhInfo: src/Bullet3Collision/NarrowPhaseCollision/
 * shared/b3QuantizedBvhNodeData.h:14-21, shared/b3BvhSubtreeInfoData.h:7-17,
 * src/Bullet3OpenCL/NarrowphaseCollision/b3BvhInfo.h:6-15) */
typedef struct b3hQuantizedBvhNode
{
	unsigned short m_quantizedAabbMin[3];
	unsigned short m_quantizedAabbMax[3];
	int m_escapeIndexOrTriangleIndex; /* leaf: child index (>= 0); internal: -(subtree size) */
} b3hQuantizedBvhNode;
typedef struct b3hBvhSubtreeInfo
{
	unsigned short m_quantizedAabbMin[3];
	unsigned short m_quantizedAabbMax[3];
	int m_rootNodeIndex;
	int m_subtreeSize;
	int m_padding[3];
} b3hBvhSubtreeInfo;
typedef struct b3hBvhInfo
{
	b3hFloat4 m_aabbMin;
	b3hFloat4 m_aabbMax;
	b3hFloat4 m_quantization;
	int m_numNodes;
	int m_numSubTrees;
	int m_nodeOffset;
	int m_subTreeOffset;
} b3hBvhInfo;

typedef struct b3hGpuFace
{
	b3hFloat4 m_plane; /* xyz normal, w offset: dot(n,p)+w = 0 on the face */
	int m_indexOffset;
	int m_numIndices;
	int m_unusedPadding1;
	int m_unusedPadding2;
} b3hGpuFace;

typedef struct b3hConvexPolyhedronData
{
	b3hFloat4 m_localCenter;
	b3hFloat4 m_extents;
	b3hFloat4 mC;
	b3hFloat4 mE;
	float m_radius;
	int m_faceOffset;
	int m_numFaces;
	int m_numVertices;
	int m_vertexOffset;
	int m_uniqueEdgesOffset;
	int m_numUniqueEdges;
	int m_unused;
} b3hConvexPolyhedronData;

/* A precomputed convex polyhedron: the members of b3ConvexUtility
 * (src/Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.h:28-45) as flat host
 * arrays, for registerConvexHullShape(b3ConvexUtility*).  vertices and
 * uniqueEdges are b3Vector3 arrays (xyz, w unused); face f has the plane
 * facePlanes[f] and the faceIndexCounts[f] next entries of indices.  mC, mE,
 * m_extents and m_radius are copied into the hull record as the reference
 * copies them.  localCenterOut (float[4], may be NULL) receives the local
 * centre the registration computes, as the reference writes it into the
 * utility's m_localCenter. */
typedef struct b3hConvexUtility
{
	const b3hFloat4* vertices;
	int numVertices;
	int numFaces;
	const b3hFloat4* facePlanes;
	const int* faceIndexCounts;
	const int* indices;
	const b3hFloat4* uniqueEdges;
	int numUniqueEdges;
	int pad;
	b3hFloat4 mC;
	b3hFloat4 mE;
	b3hFloat4 m_extents;
	float m_radius;
	float* localCenterOut;
} b3hConvexUtility;

/* min.w carries the body index as int bits (m_minIndices[3]); max.w carries
 * an int (m_signedMaxIndices[3]). */
typedef struct b3hAabb
{
	b3hFloat4 m_min;
	b3hFloat4 m_max;
} b3hAabb;

typedef struct b3hContact4Data
{
	b3hFloat4 m_worldPosB[4]; /* w: penetration depth (negative = penetrating) */
	b3hFloat4 m_worldNormalOnB; /* w: number of points */
	unsigned short m_restituitionCoeffCmp;
	unsigned short m_frictionCoeffCmp;
	int m_batchIdx;
	int m_bodyAPtrAndSignBit;
	int m_bodyBPtrAndSignBit;
	int m_childIndexA;
	int m_childIndexB;
	int m_unused1;
	int m_unused2;
} b3hContact4Data;

typedef struct b3hContactConstraint4
{
	b3hFloat4 m_linear; /* w: friction coefficient */
	b3hFloat4 m_worldPos[4];
	b3hFloat4 m_center;
	float m_jacCoeffInv[4];
	float m_b[4];
	float m_appliedRambdaDt[4];
	float m_fJacCoeffInv[2];
	float m_fAppliedRambdaDt[2];
	unsigned int m_bodyA;
	unsigned int m_bodyB;
	int m_batchIdx;
	unsigned int m_paddings;
} b3hContactConstraint4;

/* b3GpuGenericConstraint, src/Bullet3OpenCL/RigidBody/b3GpuGenericConstraint.h:
 * 20-84: a joint between bodies m_rbA and m_rbB (point-to-point: type 3, three
 * rows; fixed: type 4, six rows), enabled while m_flags has bit 1 set (the
 * solver clears it when a row's impulse reaches m_breakingImpulseThreshold). */
#define B3H_CONSTRAINT_FLAG_ENABLED 1
#define B3H_POINT2POINT_CONSTRAINT_TYPE 3
#define B3H_FIXED_CONSTRAINT_TYPE 4
typedef struct __attribute__((aligned(16))) b3hGpuGenericConstraint
{
	int m_constraintType;
	int m_rbA;
	int m_rbB;
	float m_breakingImpulseThreshold;
	b3hFloat4 m_pivotInA;
	b3hFloat4 m_pivotInB;
	b3hFloat4 m_relTargetAB; /* quaternion */
	int m_flags;
	int m_uid;
	int m_padding[2];
} b3hGpuGenericConstraint;

/* b3RayInfo / b3RayHit, src/Bullet3Collision/NarrowPhaseCollision/b3RaycastInfo.h:7-23 */
typedef struct b3hRayInfo
{
	b3hFloat4 m_from;
	b3hFloat4 m_to;
} b3hRayInfo;
typedef struct __attribute__((aligned(16))) b3hRayHit
{
	float m_hitFraction;
	int m_hitBody;
	int m_hitResult1;
	int m_hitResult2; /* input: a body the ray ignores (-1: none) */
	b3hFloat4 m_hitPoint;
	b3hFloat4 m_hitNormal;
} b3hRayHit;

typedef struct b3hSortData
{
	unsigned int m_key;
	unsigned int m_value;
} b3hSortData;

typedef struct b3hConfig
{
	int m_maxConvexBodies;
	int m_maxConvexShapes;
	int m_maxBroadphasePairs;
	int m_maxContactCapacity;
	int m_compoundPairCapacity;
	int m_maxVerticesPerFace;
	int m_maxFacesPerShape;
	int m_maxConvexVertices;
	int m_maxConvexIndices;
	int m_maxConvexUniqueEdges;
	int m_maxCompoundChildShapes;
	int m_maxTriConvexPairCapacity;
} b3hConfig;

/* Solver and integration parameters.  The reference hard-codes these values
 * (defaults in brackets); the configurations of this project set some of them.
 *   dt 1/60          src/Bullet3OpenCL/RigidBody/b3GpuPgsContactSolver.cpp:672
 *   iterations 4     src/Bullet3OpenCL/RigidBody/b3GpuPgsContactSolver.cpp:1049
 *   friction 0.7     src/Bullet3OpenCL/RigidBody/kernels/solveFriction.cl:317,
 *                    src/Bullet3Dynamics/shared/b3ConvertConstraint4.h:78
 *   drift 0.005, Baumgarte 0.2  src/Bullet3OpenCL/RigidBody/b3GpuPgsContactSolver.cpp:247
 *   batch cell 6 m   src/Bullet3OpenCL/RigidBody/b3GpuPgsContactSolver.cpp:675
 *   angular damping 0.99        src/Bullet3OpenCL/RigidBody/b3GpuRigidBodyPipeline.cpp:469
 *   gravity (0,-9.8,0)          src/Bullet3OpenCL/RigidBody/b3GpuRigidBodyPipeline.cpp:92
 */
typedef struct b3hSolverParams
{
	float m_solverDt;
	int m_numIterations;
	float m_frictionCoeff;
	float m_positionDrift;
	float m_positionConstraintCoeff;
	float m_batchCellSize;
	float m_angularDamping;
	int m_warmStartMode; /* 0 cold start (reference), 1 previous frame, 2 learned */
} b3hSolverParams;

#ifdef __cplusplus
}
#endif

#endif /* B3HIP_TYPES_H */
