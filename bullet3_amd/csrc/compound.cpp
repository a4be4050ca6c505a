/*
 * compound.cpp -- host tables of a compound shape (a list of convex-hull
 * children with child transforms): the compound's local AABB and the
 * quantized AABB tree over its children that the compound-compound pair search
 * walks.  Follows b3GpuNarrowPhase::registerCompoundShape
 * (src/Bullet3OpenCL/RigidBody/b3GpuNarrowPhase.cpp:370-517), which builds a
 * b3QuantizedBvh (src/Bullet3OpenCL/NarrowphaseCollision/b3QuantizedBvh.cpp:
 * buildInternal :34-70, setQuantizationValues :79-89, buildTree :99-176,
 * updateSubtreeHeaders :178-208, sortAndCalcSplittingIndex :210-261,
 * calcSplittingAxis :263-288; quantize / unQuantize / mergeInternalNodeAabb
 * b3QuantizedBvh.h:241-262,314-398).  The float operations are those of the
 * reference's b3Vector3 scalar path, in the same order, so the tables come out
 * identical (checked against the reference's own class in the tests).
 */
#include "b3h_internal.h"
#include <cmath>
#include <cstring>

namespace
{
struct V3
{
	float x, y, z;
};
V3 v3(float x, float y, float z) { return V3{x, y, z}; }
V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
V3 scale(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* the quantized tree of one compound (b3QuantizedBvh with quantization on) */
struct QBvh
{
	V3 bvhMin, bvhMax, quant;
	std::vector<b3hQuantizedBvhNode> leaves, nodes;
	std::vector<b3hBvhSubtreeInfo> subtrees;
	int cur = 0;

	void set_quantization(V3 mn, V3 mx, float margin)
	{
		V3 clamp = v3(margin, margin, margin);
		bvhMin = mn - clamp;
		bvhMax = mx + clamp;
		V3 size = bvhMax - bvhMin;
		quant = v3(65533.0f / size.x, 65533.0f / size.y, 65533.0f / size.z);
	}
	void quantize(unsigned short* out, V3 p, int isMax) const
	{
		V3 v = (p - bvhMin) * quant;
		if (isMax)
		{
			out[0] = (unsigned short)(((unsigned short)(v.x + 1.0f)) | 1);
			out[1] = (unsigned short)(((unsigned short)(v.y + 1.0f)) | 1);
			out[2] = (unsigned short)(((unsigned short)(v.z + 1.0f)) | 1);
		}
		else
		{
			out[0] = (unsigned short)(((unsigned short)(v.x)) & 0xfffe);
			out[1] = (unsigned short)(((unsigned short)(v.y)) & 0xfffe);
			out[2] = (unsigned short)(((unsigned short)(v.z)) & 0xfffe);
		}
	}
	V3 unquantize(const unsigned short* q) const
	{
		V3 v = v3((float)q[0] / quant.x, (float)q[1] / quant.y, (float)q[2] / quant.z);
		return v + bvhMin;
	}
	V3 leaf_min(int i) const { return unquantize(leaves[(size_t)i].m_quantizedAabbMin); }
	V3 leaf_max(int i) const { return unquantize(leaves[(size_t)i].m_quantizedAabbMax); }
	static bool is_leaf(const b3hQuantizedBvhNode& n) { return n.m_escapeIndexOrTriangleIndex >= 0; }
	static int escape(const b3hQuantizedBvhNode& n) { return -n.m_escapeIndexOrTriangleIndex; }

	void merge(int node, V3 mn, V3 mx)
	{
		unsigned short qmn[3], qmx[3];
		quantize(qmn, mn, 0);
		quantize(qmx, mx, 1);
		b3hQuantizedBvhNode& n = nodes[(size_t)node];
		for (int i = 0; i < 3; i++)
		{
			if (n.m_quantizedAabbMin[i] > qmn[i]) n.m_quantizedAabbMin[i] = qmn[i];
			if (n.m_quantizedAabbMax[i] < qmx[i]) n.m_quantizedAabbMax[i] = qmx[i];
		}
	}
	int split_axis(int start, int end) const
	{
		V3 means = v3(0.f, 0.f, 0.f), var = v3(0.f, 0.f, 0.f);
		int n = end - start;
		for (int i = start; i < end; i++) means = means + scale(0.5f, leaf_max(i) + leaf_min(i));
		means = scale(1.f / (float)n, means);
		for (int i = start; i < end; i++)
		{
			V3 d = scale(0.5f, leaf_max(i) + leaf_min(i)) - means;
			var = var + d * d;
		}
		var = scale(1.f / ((float)n - 1), var);
		return var.x < var.y ? (var.y < var.z ? 2 : 1) : (var.x < var.z ? 2 : 0);
	}
	int split_index(int start, int end, int axis)
	{
		int split = start, n = end - start;
		V3 means = v3(0.f, 0.f, 0.f);
		for (int i = start; i < end; i++) means = means + scale(0.5f, leaf_max(i) + leaf_min(i));
		means = scale(1.f / (float)n, means);
		float splitValue = comp(means, axis);
		for (int i = start; i < end; i++)
		{
			V3 c = scale(0.5f, leaf_max(i) + leaf_min(i));
			if (comp(c, axis) > splitValue)
			{
				std::swap(leaves[(size_t)i], leaves[(size_t)split]);
				split++;
			}
		}
		int balanced = n / 3;
		if (split <= start + balanced || split >= end - 1 - balanced) split = start + (n >> 1);
		return split;
	}
	void subtree_from(int nodeIndex)
	{
		const b3hQuantizedBvhNode& n = nodes[(size_t)nodeIndex];
		b3hBvhSubtreeInfo s;
		memset(&s, 0, sizeof(s));
		memcpy(s.m_quantizedAabbMin, n.m_quantizedAabbMin, sizeof(s.m_quantizedAabbMin));
		memcpy(s.m_quantizedAabbMax, n.m_quantizedAabbMax, sizeof(s.m_quantizedAabbMax));
		s.m_rootNodeIndex = nodeIndex;
		s.m_subtreeSize = is_leaf(n) ? 1 : escape(n);
		subtrees.push_back(s);
	}
	void update_subtrees(int left, int right)
	{
		const int limit = 2048; /* MAX_SUBTREE_SIZE_IN_BYTES */
		const b3hQuantizedBvhNode& l = nodes[(size_t)left];
		const b3hQuantizedBvhNode& r = nodes[(size_t)right];
		int ls = is_leaf(l) ? 1 : escape(l), rs = is_leaf(r) ? 1 : escape(r);
		if (ls * (int)sizeof(b3hQuantizedBvhNode) <= limit) subtree_from(left);
		if (rs * (int)sizeof(b3hQuantizedBvhNode) <= limit) subtree_from(right);
	}
	void build(int start, int end)
	{
		int n = end - start;
		int curIndex = cur;
		if (n == 1)
		{
			nodes[(size_t)cur] = leaves[(size_t)start];
			cur++;
			return;
		}
		int axis = split_axis(start, end);
		int split = split_index(start, end, axis);
		int internal = cur;
		quantize(nodes[(size_t)cur].m_quantizedAabbMin, bvhMax, 0);
		quantize(nodes[(size_t)cur].m_quantizedAabbMax, bvhMin, 1);
		for (int i = start; i < end; i++) merge(cur, leaf_min(i), leaf_max(i));
		cur++;
		int left = cur;
		build(start, split);
		int right = cur;
		build(split, end);
		int escapeIndex = cur - curIndex;
		if (escapeIndex * (int)sizeof(b3hQuantizedBvhNode) > 2048) update_subtrees(left, right);
		nodes[(size_t)internal].m_escapeIndexOrTriangleIndex = -escapeIndex;
	}
	void build_all()
	{
		int n = (int)leaves.size();
		nodes.assign((size_t)2 * n, b3hQuantizedBvhNode());
		for (b3hQuantizedBvhNode& q : nodes) memset(&q, 0, sizeof(q));
		cur = 0;
		build(0, n);
		if (subtrees.empty()) subtree_from(0);
	}
};
} // namespace

/* b3TransformAabb (b3AabbUtil.h:182-198) of a child's local AABB by the child
 * transform, margin 0 */
static void child_aabb(const b3hGpuChildShape& c, const b3hAabb& la, V3* mn, V3* mx)
{
	float m[3][3];
	const float q[4] = {c.m_childOrientation.x, c.m_childOrientation.y, c.m_childOrientation.z, c.m_childOrientation.w};
	b3h_quat_to_basis(q, m);
	V3 lmin = v3(la.m_min.x, la.m_min.y, la.m_min.z), lmax = v3(la.m_max.x, la.m_max.y, la.m_max.z);
	V3 half = scale(0.5f, lmax - lmin);
	half = half + v3(0.f, 0.f, 0.f);
	V3 center = scale(0.5f, lmax + lmin);
	V3 r0 = v3(m[0][0], m[0][1], m[0][2]), r1 = v3(m[1][0], m[1][1], m[1][2]), r2 = v3(m[2][0], m[2][1], m[2][2]);
	V3 a0 = v3(fabsf(r0.x), fabsf(r0.y), fabsf(r0.z)), a1 = v3(fabsf(r1.x), fabsf(r1.y), fabsf(r1.z)),
	   a2 = v3(fabsf(r2.x), fabsf(r2.y), fabsf(r2.z));
	V3 wc = v3(dot(r0, center), dot(r1, center), dot(r2, center)) +
			v3(c.m_childPosition.x, c.m_childPosition.y, c.m_childPosition.z);
	V3 ext = v3(dot(half, a0), dot(half, a1), dot(half, a2));
	*mn = wc - ext;
	*mx = wc + ext;
}

int b3h_build_compound(int n, const b3hGpuChildShape* children, const b3hAabb* childLocal, b3hAabb* compoundAabb,
					   b3hBvhInfo* info, std::vector<b3hQuantizedBvhNode>& nodes,
					   std::vector<b3hBvhSubtreeInfo>& subtrees)
{
	if (n <= 0) return B3H_ERR_INVALID;
	V3 mn = v3(1e30f, 1e30f, 1e30f), mx = v3(-1e30f, -1e30f, -1e30f);
	std::vector<V3> cmin((size_t)n), cmax((size_t)n);
	for (int i = 0; i < n; i++)
	{
		child_aabb(children[i], childLocal[i], &cmin[(size_t)i], &cmax[(size_t)i]);
		/* b3Vector3::setMin / setMax: b3SetMin(a, b) is "if (b < a) a = b" per component */
		if (cmin[(size_t)i].x < mn.x) mn.x = cmin[(size_t)i].x;
		if (cmin[(size_t)i].y < mn.y) mn.y = cmin[(size_t)i].y;
		if (cmin[(size_t)i].z < mn.z) mn.z = cmin[(size_t)i].z;
		if (mx.x < cmax[(size_t)i].x) mx.x = cmax[(size_t)i].x;
		if (mx.y < cmax[(size_t)i].y) mx.y = cmax[(size_t)i].y;
		if (mx.z < cmax[(size_t)i].z) mx.z = cmax[(size_t)i].z;
	}
	memset(compoundAabb, 0, sizeof(*compoundAabb));
	compoundAabb->m_min = f4(mn.x, mn.y, mn.z, 0.f);
	compoundAabb->m_max = f4(mx.x, mx.y, mx.z, 0.f);
	QBvh bvh;
	bvh.set_quantization(mn, mx, 1.0f);
	bvh.leaves.resize((size_t)n);
	for (int i = 0; i < n; i++)
	{
		b3hQuantizedBvhNode& q = bvh.leaves[(size_t)i];
		memset(&q, 0, sizeof(q));
		bvh.quantize(q.m_quantizedAabbMin, cmin[(size_t)i], 0);
		bvh.quantize(q.m_quantizedAabbMax, cmax[(size_t)i], 1);
		q.m_escapeIndexOrTriangleIndex = i; /* part 0 */
	}
	bvh.build_all();
	memset(info, 0, sizeof(*info));
	info->m_aabbMin = f4(bvh.bvhMin.x, bvh.bvhMin.y, bvh.bvhMin.z, 0.f);
	info->m_aabbMax = f4(bvh.bvhMax.x, bvh.bvhMax.y, bvh.bvhMax.z, 0.f);
	info->m_quantization = f4(bvh.quant.x, bvh.quant.y, bvh.quant.z, 0.f);
	info->m_numNodes = n;
	info->m_numSubTrees = (int)bvh.subtrees.size();
	nodes = bvh.nodes;
	subtrees = bvh.subtrees;
	return B3H_OK;
}
