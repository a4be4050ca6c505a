/*
 * hull.cpp -- convex hull of a point cloud and the polyhedral features the
 * narrowphase consumes (host code, used once per registered shape).
 *
 * Follows b3ConvexUtility::initializePolyhedralFeatures / initialize
 * (src/Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.cpp:26-248, 325-365):
 *   hull triangles -> face normal from the first two (normalised) edges, plane
 *   offset = -min over the face of dot(v, n) -> coplanar faces (normal dot >
 *   0.999) merged through a 2-D hull unless a dropped vertex is used by another
 *   face -> unique edge directions (within 1e-6, either sign).
 * The hull itself comes from an exact enumeration of supporting planes (in
 * double precision) instead of the reference's b3ConvexHullComputer; the
 * polyhedron is the same, the enumeration order of vertices and faces is ours
 * (vertices keep the order in which they appear in the input).
 */
#include <vector>
#include <cmath>
#include <algorithm>
#include <cstdio>
#include "hull.h"

namespace
{
struct D3
{
	double x, y, z;
};
D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct V3
{
	float x, y, z;
};
V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 vscale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 vcross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 vnorm(V3 a)
{
	float l = std::sqrt(vdot(a, a));
	float k = 1.f / l;
	return vscale(a, k);
}

struct Face
{
	std::vector<int> idx;
	float plane[4];
};

/* b3PlaneSpace1 (float), used to orient the 2-D merge hull */
void plane_space(V3 n, V3& p, V3& q)
{
	if (std::fabs(n.z) > 0.70710678f)
	{
		float a = n.y * n.y + n.z * n.z;
		float k = 1.f / std::sqrt(a);
		p = {0, -n.z * k, n.y * k};
		q = {a * k, -n.x * p.z, n.x * p.y};
	}
	else
	{
		float a = n.x * n.x + n.y * n.y;
		float k = 1.f / std::sqrt(a);
		p = {-n.y * k, n.x * k, 0};
		q = {-n.z * p.y, n.z * p.x, a * k};
	}
}

/* 2-D convex hull of points (indices into v) in the plane with normal n:
 * anchor = lowest projection on the first tangent, angular sort, convex
 * turns kept (b3GrahamScanConvexHull2D's scheme). */
std::vector<int> hull2d(const std::vector<V3>& v, std::vector<int> pts, V3 n)
{
	if (pts.size() <= 2) return pts;
	V3 a0, a1;
	plane_space(n, a0, a1);
	for (size_t i = 0; i < pts.size(); i++)
		if (vdot(v[pts[i]], a0) < vdot(v[pts[0]], a0)) std::swap(pts[0], pts[i]);
	V3 anchor = v[pts[0]];
	std::vector<std::pair<float, int>> ang;
	for (size_t i = 1; i < pts.size(); i++)
	{
		V3 ar = vsub(v[pts[i]], anchor);
		float l = std::sqrt(vdot(ar, ar));
		ang.push_back({vdot(vcross(a0, ar), n) / l, pts[i]});
	}
	std::stable_sort(ang.begin(), ang.end(), [&](const std::pair<float, int>& p, const std::pair<float, int>& q) {
		if (p.first != q.first) return p.first < q.first;
		float lp = vdot(vsub(v[p.second], anchor), vsub(v[p.second], anchor));
		float lq = vdot(vsub(v[q.second], anchor), vsub(v[q.second], anchor));
		if (lp != lq) return lp < lq;
		return p.second < q.second;
	});
	std::vector<int> h;
	h.push_back(pts[0]);
	for (auto& e : ang)
	{
		int c = e.second;
		while (h.size() >= 2)
		{
			V3 a = v[h[h.size() - 2]], b = v[h[h.size() - 1]];
			bool convex = vdot(vcross(vsub(a, b), vsub(a, v[c])), n) > 0;
			if (convex) break;
			h.pop_back();
		}
		h.push_back(c);
	}
	return h;
}
}  // namespace

/* The reference's hull computer works on integer coordinates: points are
 * centred on their bounding box and truncated to a grid of extent/10216 per
 * axis, and the hull vertices it returns are those grid points mapped back
 * (b3ConvexHullInternal::compute / getCoordinates,
 * src/Bullet3Geometry/b3ConvexHullComputer.cpp:1943-2085).  The same snapping is
 * applied here so the registered vertices coincide with the reference's. */
static void snap_to_hull_grid(const float* pts, int n, int stride, std::vector<float>& outPts)
{
	float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
	for (int i = 0; i < n; i++)
		for (int k = 0; k < 3; k++)
		{
			float v = pts[i * stride + k];
			mn[k] = v < mn[k] ? v : mn[k];
			mx[k] = v > mx[k] ? v : mx[k];
		}
	float s[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
	int maxAxis = s[0] < s[1] ? (s[1] < s[2] ? 2 : 1) : (s[0] < s[2] ? 2 : 0);
	int minAxis = s[0] < s[1] ? (s[0] < s[2] ? 0 : 2) : (s[1] < s[2] ? 1 : 2);
	if (minAxis == maxAxis) minAxis = (maxAxis + 1) % 3;
	int medAxis = 3 - maxAxis - minAxis;
	float r = 1.0f / 10216.0f;
	for (int k = 0; k < 3; k++) s[k] *= r;
	if (((medAxis + 1) % 3) != maxAxis)
		for (int k = 0; k < 3; k++) s[k] *= -1.f;
	float scaling[3] = {s[0], s[1], s[2]};
	for (int k = 0; k < 3; k++)
		if (s[k] != 0) s[k] = 1.0f / s[k];
	float center[3];
	for (int k = 0; k < 3; k++) center[k] = (mn[k] + mx[k]) * 0.5f;
	outPts.resize(3 * (size_t)n);
	for (int i = 0; i < n; i++)
		for (int k = 0; k < 3; k++)
		{
			float q = (pts[i * stride + k] - center[k]) * s[k];
			int iq = (int)q;
			outPts[3 * i + k] = (float)iq * scaling[k] + center[k];
		}
}

int b3h_build_polyhedron(const float* ptsIn, int n, int strideFloats, b3hPolyhedron* out)
{
	out->vertices.clear();
	out->faces.clear();
	out->uniqueEdges.clear();
	if (n < 4) return -1;
	std::vector<float> snapped;
	snap_to_hull_grid(ptsIn, n, strideFloats, snapped);
	const float* pts = snapped.data();
	strideFloats = 3;
	std::vector<D3> p;
	std::vector<int> orig;
	for (int i = 0; i < n; i++)
	{
		D3 q = {pts[i * strideFloats], pts[i * strideFloats + 1], pts[i * strideFloats + 2]};
		bool dup = false;
		for (auto& r : p)
			if (r.x == q.x && r.y == q.y && r.z == q.z) dup = true;
		if (!dup)
		{
			p.push_back(q);
			orig.push_back(i);
		}
	}
	int m = (int)p.size();
	if (m < 4) return -1;
	double scale = 0;
	for (auto& q : p) scale = std::max(scale, std::max(std::fabs(q.x), std::max(std::fabs(q.y), std::fabs(q.z))));
	if (scale == 0) return -1;
	double eps = 1e-6 * scale;
	/* supporting planes through point triples */
	struct Pl
	{
		D3 n;
		double d;
	};
	std::vector<Pl> planes;
	for (int i = 0; i < m; i++)
		for (int j = i + 1; j < m; j++)
			for (int k = j + 1; k < m; k++)
			{
				D3 nn = dcross(dsub(p[j], p[i]), dsub(p[k], p[i]));
				double l = std::sqrt(ddot(nn, nn));
				if (l < 1e-12 * scale * scale) continue;
				nn = {nn.x / l, nn.y / l, nn.z / l};
				double d = ddot(nn, p[i]);
				double mx = -1e300, mn = 1e300;
				for (int t = 0; t < m; t++)
				{
					double s = ddot(nn, p[t]) - d;
					mx = std::max(mx, s);
					mn = std::min(mn, s);
				}
				if (mx <= eps) {}
				else if (mn >= -eps)
				{
					nn = {-nn.x, -nn.y, -nn.z};
					d = -d;
				}
				else
					continue;
				bool seen = false;
				for (auto& q : planes)
					if (ddot(q.n, nn) > 1 - 1e-9 && std::fabs(q.d - d) <= eps) seen = true;
				if (!seen) planes.push_back({nn, d});
			}
	if (planes.size() < 4) return -1;
	/* polygon of each supporting plane, vertices = union, in input order */
	std::vector<std::vector<int>> polys;
	std::vector<int> used(m, 0);
	std::vector<V3> pf(m);
	for (int i = 0; i < m; i++) pf[i] = {(float)p[i].x, (float)p[i].y, (float)p[i].z};
	for (auto& pl : planes)
	{
		std::vector<int> on;
		for (int t = 0; t < m; t++)
			if (std::fabs(ddot(pl.n, p[t]) - pl.d) <= eps) on.push_back(t);
		V3 nf = {(float)pl.n.x, (float)pl.n.y, (float)pl.n.z};
		std::vector<int> poly = hull2d(pf, on, nf);
		if (poly.size() < 3) continue;
		for (int t : poly) used[t] = 1;
		polys.push_back(poly);
	}
	std::vector<int> remap(m, -1);
	for (int i = 0; i < m; i++)
		if (used[i])
		{
			remap[i] = (int)out->vertices.size() / 3;
			out->vertices.push_back(pf[i].x);
			out->vertices.push_back(pf[i].y);
			out->vertices.push_back(pf[i].z);
		}
	int nv = (int)out->vertices.size() / 3;
	std::vector<V3> vv(nv);
	for (int i = 0; i < nv; i++) vv[i] = {out->vertices[3 * i], out->vertices[3 * i + 1], out->vertices[3 * i + 2]};
	/* hull triangles (fans of the supporting polygons) */
	std::vector<Face> tmp;
	for (auto& poly : polys)
		for (size_t t = 1; t + 1 < poly.size(); t++)
		{
			Face f;
			f.idx = {remap[poly[0]], remap[poly[t]], remap[poly[t + 1]]};
			tmp.push_back(f);
		}
	/* face planes (b3ConvexUtility.cpp:51-98) */
	for (auto& f : tmp)
	{
		V3 e0 = vnorm(vsub(vv[f.idx[1]], vv[f.idx[0]]));
		V3 e1 = vnorm(vsub(vv[f.idx[2]], vv[f.idx[1]]));
		V3 nn = vnorm(vcross(e0, e1));
		float pe = 1e30f;
		for (int vi : f.idx)
		{
			float e = vdot(vv[vi], nn);
			if (pe > e) pe = e;
		}
		f.plane[0] = nn.x;
		f.plane[1] = nn.y;
		f.plane[2] = nn.z;
		f.plane[3] = -pe;
	}
	/* merge coplanar faces (b3ConvexUtility.cpp:100-243) */
	std::vector<int> todo;
	for (int i = 0; i < (int)tmp.size(); i++) todo.push_back(i);
	std::vector<Face> faces;
	while (!todo.empty())
	{
		std::vector<int> group;
		int ref = todo.back();
		todo.pop_back();
		group.push_back(ref);
		V3 na = {tmp[ref].plane[0], tmp[ref].plane[1], tmp[ref].plane[2]};
		for (int j = (int)todo.size() - 1; j >= 0; j--)
		{
			int i = todo[j];
			V3 nb = {tmp[i].plane[0], tmp[i].plane[1], tmp[i].plane[2]};
			if (vdot(na, nb) > 0.999f)
			{
				group.push_back(i);
				todo.erase(todo.begin() + j);
			}
		}
		bool merged = false;
		if (group.size() > 1)
		{
			std::vector<int> pts2;
			V3 avg = {0, 0, 0};
			for (int gi : group)
			{
				avg = vadd(avg, V3{tmp[gi].plane[0], tmp[gi].plane[1], tmp[gi].plane[2]});
				for (int vi : tmp[gi].idx)
					if (std::find(pts2.begin(), pts2.end(), vi) == pts2.end()) pts2.push_back(vi);
			}
			avg = vnorm(avg);
			std::vector<int> h = hull2d(vv, pts2, avg);
			bool reject = false;
			for (int vi : pts2)
			{
				if (std::find(h.begin(), h.end(), vi) != h.end()) continue;
				for (int fj = 0; fj < (int)tmp.size() && !reject; fj++)
				{
					if (std::find(group.begin(), group.end(), fj) != group.end()) continue;
					for (int u : tmp[fj].idx)
						if (u == vi) reject = true;
				}
				if (reject) break;
			}
			if (!reject)
			{
				Face c;
				c.idx = h;
				for (int k = 0; k < 4; k++) c.plane[k] = tmp[group[0]].plane[k];
				faces.push_back(c);
				merged = true;
			}
		}
		if (!merged)
			for (int gi : group) faces.push_back(tmp[gi]);
	}
	/* unique edges (b3ConvexUtility.cpp:330-365) */
	std::vector<V3> ue;
	for (auto& f : faces)
	{
		int nvf = (int)f.idx.size();
		for (int j = 0; j < nvf; j++)
		{
			int a = f.idx[j], b = f.idx[(j + 1) % nvf];
			int v0 = a, v1 = b;
			if (v1 > v0) std::swap(v0, v1);
			V3 e = vnorm(vsub(vv[v1], vv[v0]));
			bool found = false;
			for (auto& u : ue)
			{
				V3 d1 = vsub(u, e), d2 = vadd(u, e);
				bool z1 = !(std::fabs(d1.x) > 1e-6f || std::fabs(d1.y) > 1e-6f || std::fabs(d1.z) > 1e-6f);
				bool z2 = !(std::fabs(d2.x) > 1e-6f || std::fabs(d2.y) > 1e-6f || std::fabs(d2.z) > 1e-6f);
				if (z1 || z2)
				{
					found = true;
					break;
				}
			}
			if (!found) ue.push_back(e);
		}
	}
	for (auto& f : faces)
	{
		b3hPolyFace pf2;
		for (int k = 0; k < 4; k++) pf2.plane[k] = f.plane[k];
		pf2.indices = f.idx;
		out->faces.push_back(pf2);
	}
	for (auto& e : ue)
	{
		out->uniqueEdges.push_back(e.x);
		out->uniqueEdges.push_back(e.y);
		out->uniqueEdges.push_back(e.z);
	}
	return 0;
}
