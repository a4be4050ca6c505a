/*
 * hull.cpp -- convex hull of a point cloud and the polyhedral features the
 * narrowphase consumes (host code, run once per registered shape).
 *
 * The narrowphase takes support vertices with a first-maximum rule, so the
 * order in which a shape lists its vertices, faces and face indices decides
 * which of several tied vertices a contact uses.  Registration therefore has to
 * produce the reference's tables in the reference's order, not only the same
 * polyhedron.  Two stages, each a restatement of the reference's algorithm:
 *
 * 1. Exact 3-D hull (b3ConvexHullInternal / b3ConvexHullComputer::compute,
 *    src/Bullet3Geometry/b3ConvexHullComputer.cpp:1943-2085, 2640-2745):
 *    points are centred on their bounding box and truncated to an integer grid
 *    of extent/10216 per axis, sorted by (y, x, z) in a permuted axis frame,
 *    and hulled by divide and conquer.  Each merge first joins the two
 *    hulls' 2-D projections (mergeProjection, :1024-1207), then wraps the
 *    bridge faces around them by gift wrapping with exact rational angle
 *    comparisons (merge / findMaxAngle / findEdgeForCoplanarFaces,
 *    :1390-1936).  The result is a half-edge graph; its vertices are numbered
 *    breadth first from the hull's min-x vertex following each vertex's edge
 *    ring, and its faces are listed by walking the rings in that order.  All
 *    comparisons are exact (64-bit coordinates products, 128-bit rational
 *    cross products), so any exact arithmetic gives the reference's decisions.
 *
 * 2. Polyhedral features (b3ConvexUtility::initializePolyhedralFeatures and
 *    initialize, src/Bullet3Collision/NarrowPhaseCollision/b3ConvexUtility.cpp:
 *    26-248, 325-365): face normal from the face's first two normalised edges,
 *    plane offset -min dot; faces whose normals agree within 0.999 are merged
 *    through a 2-D Graham scan (src/Bullet3Geometry/b3GrahamScan2dConvexHull.h)
 *    unless a dropped vertex is used by another face; unique edge directions
 *    within 1e-6.  Float operations follow the reference's scalar b3Vector3
 *    code (normalise = multiply by 1/sqrt, left-to-right dots), and the work
 *    list of faces is edited the way b3AlignedObjectArray::remove edits it
 *    (swap with the last element), since that order decides the face order.
 */
#include <vector>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include "hull.h"

namespace
{
typedef long long i64;
typedef unsigned long long u64;
typedef unsigned __int128 u128;

/* ---------------------------------------------------------------- stage 1 */

struct Q3 /* grid point (coordinates within +-5109, so 32-bit products fit) */
{
	int x, y, z;
	bool operator==(const Q3& o) const { return x == o.x && y == o.y && z == o.z; }
	bool operator!=(const Q3& o) const { return !(*this == o); }
};
struct W3
{
	i64 x, y, z;
};
inline Q3 qsub(const Q3& a, const Q3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline W3 qcross(const Q3& a, const Q3& b)
{
	return {(i64)a.y * b.z - (i64)a.z * b.y, (i64)a.z * b.x - (i64)a.x * b.z, (i64)a.x * b.y - (i64)a.y * b.x};
}
inline W3 qcross(const Q3& a, const W3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline i64 qdot(const Q3& a, const Q3& b) { return (i64)a.x * b.x + (i64)a.y * b.y + (i64)a.z * b.z; }
inline i64 qdot(const Q3& a, const W3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline i64 wdot(const W3& a, const W3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* p/q with sign kept apart; q = 0 is an infinity (p != 0) or undefined (p = 0) */
struct Ratio
{
	u64 num = 0, den = 0;
	int sign = 0;
	Ratio() {}
	Ratio(i64 p, i64 q)
	{
		sign = p > 0 ? 1 : (p < 0 ? -1 : 0);
		num = p > 0 ? (u64)p : (u64)(-(u128)(u64)p);
		if (p == 0) num = 0;
		if (q > 0)
			den = (u64)q;
		else if (q < 0)
		{
			sign = -sign;
			den = (u64)(-(u128)(u64)q);
		}
		else
			den = 0;
	}
	bool negInf() const { return sign < 0 && den == 0; }
	bool undefined() const { return sign == 0 && den == 0; }
	int cmp(const Ratio& o) const
	{
		if (sign != o.sign) return sign - o.sign;
		if (sign == 0) return 0;
		u128 l = (u128)num * o.den, r = (u128)den * o.num;
		return l > r ? sign : (l < r ? -sign : 0);
	}
};

struct HVert
{
	int next = -1, prev = -1; /* ring of the 2-D projection hull */
	int edges = -1;           /* one outgoing half-edge */
	Q3 p;
	int copy = -1; /* output index */
};
struct HEdge
{
	int next = -1, prev = -1; /* ring of half-edges leaving the same vertex */
	int rev = -1;
	int target = -1;
	int stamp = 0; /* merge that created it (negative), then output index */
};
struct Span /* min / max vertices of an intermediate hull's projection */
{
	int minXy = -1, maxXy = -1, minYx = -1, maxYx = -1;
};

enum Turn
{
	TURN_NONE,
	TURN_CW,
	TURN_CCW
};

class Hull3
{
public:
	std::vector<HVert> V;
	std::vector<HEdge> E;
	int stamp = -3;
	int start = -1;

	HEdge& e(int i) { return E[i]; }
	HVert& v(int i) { return V[i]; }
	Q3 diff(int a, int b) const { return qsub(V[a].p, V[b].p); }

	int new_pair(int from, int to)
	{
		int a = (int)E.size();
		E.resize(E.size() + 2);
		E[a].rev = a + 1;
		E[a + 1].rev = a;
		E[a].stamp = E[a + 1].stamp = stamp;
		E[a].target = to;
		E[a + 1].target = from;
		return a;
	}
	void link(int a, int b)
	{
		E[a].next = b;
		E[b].prev = a;
	}
	/* unlink an edge and its reverse from both rings; the rings' vertices point
	 * at the following edge (b3ConvexHullComputer.cpp:785-818) */
	void drop_pair(int a)
	{
		int r = E[a].rev;
		int n = E[a].next;
		if (n != a)
		{
			E[n].prev = E[a].prev;
			E[E[a].prev].next = n;
			V[E[r].target].edges = n;
		}
		else
			V[E[r].target].edges = -1;
		n = E[r].next;
		if (n != r)
		{
			E[n].prev = E[r].prev;
			E[E[r].prev].next = n;
			V[E[a].target].edges = n;
		}
		else
			V[E[a].target].edges = -1;
	}

	/* rotation sense of two edges leaving one vertex (:1390-1414) */
	Turn turn(int prev, int next, const Q3& s, const Q3& t)
	{
		if (E[prev].next == next)
		{
			if (E[prev].prev == next)
			{
				W3 n = qcross(t, s);
				int o = E[E[next].rev].target;
				W3 m = qcross(diff(E[prev].target, o), diff(E[next].target, o));
				return wdot(n, m) > 0 ? TURN_CCW : TURN_CW;
			}
			return TURN_CCW;
		}
		if (E[prev].prev == next) return TURN_CW;
		return TURN_NONE;
	}

	/* the older edge at `from` that makes the smallest angle with the
	 * plane through the bridge (:1416-1466) */
	int max_angle(bool ccw, int from, const Q3& s, const W3& rxs, const W3& sxrxs, Ratio& best)
	{
		int bestEdge = -1;
		int e0 = V[from].edges;
		if (e0 < 0) return -1;
		int x = e0;
		do
		{
			if (E[x].stamp > stamp)
			{
				Q3 t = diff(E[x].target, from);
				Ratio c(qdot(t, sxrxs), qdot(t, rxs));
				if (!c.undefined())
				{
					int k;
					if (bestEdge < 0)
					{
						best = c;
						bestEdge = x;
					}
					else if ((k = c.cmp(best)) < 0)
					{
						best = c;
						bestEdge = x;
					}
					else if (k == 0 && (ccw == (turn(bestEdge, x, s, t) == TURN_CCW)))
						bestEdge = x;
				}
			}
			x = E[x].next;
		} while (x != e0);
		return bestEdge;
	}

	/* advance the bridge across faces that are coplanar with it (:1468-1649) */
	void coplanar_advance(int c0, int c1, int& e0, int& e1, int stop0, int stop1)
	{
		int start0 = e0, start1 = e1;
		Q3 et0 = start0 >= 0 ? V[E[start0].target].p : V[c0].p;
		Q3 et1 = start1 >= 0 ? V[E[start1].target].p : V[c1].p;
		Q3 s = diff(c1, c0);
		W3 normal = qcross(qsub(V[E[start0 >= 0 ? start0 : start1].target].p, V[c0].p), s);
		i64 dist = qdot(V[c0].p, normal);
		W3 perp = qcross(s, normal);

		i64 maxDot0 = qdot(et0, perp);
		if (e0 >= 0)
		{
			while (E[e0].target != stop0)
			{
				int x = E[E[e0].rev].prev;
				if (qdot(V[E[x].target].p, normal) < dist) break;
				if (E[x].stamp == stamp) break;
				i64 d = qdot(V[E[x].target].p, perp);
				if (d <= maxDot0) break;
				maxDot0 = d;
				e0 = x;
				et0 = V[E[x].target].p;
			}
		}
		i64 maxDot1 = qdot(et1, perp);
		if (e1 >= 0)
		{
			while (E[e1].target != stop1)
			{
				int x = E[E[e1].rev].next;
				if (qdot(V[E[x].target].p, normal) < dist) break;
				if (E[x].stamp == stamp) break;
				i64 d = qdot(V[E[x].target].p, perp);
				if (d <= maxDot1) break;
				maxDot1 = d;
				e1 = x;
				et1 = V[E[x].target].p;
			}
		}

		i64 dx = maxDot1 - maxDot0;
		if (dx > 0)
		{
			while (true)
			{
				i64 dy = qdot(qsub(et1, et0), s);
				if (e0 >= 0 && E[e0].target != stop0)
				{
					int f0 = E[E[e0].next].rev;
					if (E[f0].stamp > stamp)
					{
						Q3 d0 = qsub(V[E[f0].target].p, et0);
						i64 dx0 = qdot(d0, perp);
						i64 dy0 = qdot(d0, s);
						if ((dx0 == 0) ? (dy0 < 0) : ((dx0 < 0) && (Ratio(dy0, dx0).cmp(Ratio(dy, dx)) >= 0)))
						{
							et0 = V[E[f0].target].p;
							dx = qdot(qsub(et1, et0), perp);
							e0 = (e0 == start0) ? -1 : f0;
							continue;
						}
					}
				}
				if (e1 >= 0 && E[e1].target != stop1)
				{
					int f1 = E[E[e1].rev].next;
					if (E[f1].stamp > stamp)
					{
						Q3 d1 = qsub(V[E[f1].target].p, et1);
						if (qdot(d1, normal) == 0)
						{
							i64 dx1 = qdot(d1, perp);
							i64 dy1 = qdot(d1, s);
							i64 dxn = qdot(qsub(V[E[f1].target].p, et0), perp);
							if ((dxn > 0) &&
								((dx1 == 0) ? (dy1 < 0) : ((dx1 < 0) && (Ratio(dy1, dx1).cmp(Ratio(dy, dx)) > 0))))
							{
								e1 = f1;
								et1 = V[E[e1].target].p;
								dx = dxn;
								continue;
							}
						}
					}
				}
				break;
			}
		}
		else if (dx < 0)
		{
			while (true)
			{
				i64 dy = qdot(qsub(et1, et0), s);
				if (e1 >= 0 && E[e1].target != stop1)
				{
					int f1 = E[E[e1].prev].rev;
					if (E[f1].stamp > stamp)
					{
						Q3 d1 = qsub(V[E[f1].target].p, et1);
						i64 dx1 = qdot(d1, perp);
						i64 dy1 = qdot(d1, s);
						if ((dx1 == 0) ? (dy1 > 0) : ((dx1 < 0) && (Ratio(dy1, dx1).cmp(Ratio(dy, dx)) <= 0)))
						{
							et1 = V[E[f1].target].p;
							dx = qdot(qsub(et1, et0), perp);
							e1 = (e1 == start1) ? -1 : f1;
							continue;
						}
					}
				}
				if (e0 >= 0 && E[e0].target != stop0)
				{
					int f0 = E[E[e0].rev].prev;
					if (E[f0].stamp > stamp)
					{
						Q3 d0 = qsub(V[E[f0].target].p, et0);
						if (qdot(d0, normal) == 0)
						{
							i64 dx0 = qdot(d0, perp);
							i64 dy0 = qdot(d0, s);
							i64 dxn = qdot(qsub(et1, V[E[f0].target].p), perp);
							if ((dxn < 0) &&
								((dx0 == 0) ? (dy0 > 0) : ((dx0 < 0) && (Ratio(dy0, dx0).cmp(Ratio(dy, dx)) < 0))))
							{
								e0 = f0;
								et0 = V[E[e0].target].p;
								dx = dxn;
								continue;
							}
						}
					}
				}
				break;
			}
		}
	}

	/* join the 2-D projection hulls of h0 (left) and h1 (right); returns false
	 * when h1 projects to a single point above h0's last vertex, with the
	 * bridge's ends in c0, c1 (:1024-1207) */
	bool join_projections(Span& h0, Span& h1, int& c0, int& c1)
	{
		int v0 = h0.maxYx, v1 = h1.minYx;
		if (V[v0].p.x == V[v1].p.x && V[v0].p.y == V[v1].p.y)
		{
			int v1p = V[v1].prev;
			if (v1p == v1)
			{
				c0 = v0;
				if (V[v1].edges >= 0) v1 = E[V[v1].edges].target;
				c1 = v1;
				return false;
			}
			int v1n = V[v1].next;
			V[v1p].next = v1n;
			V[v1n].prev = v1p;
			const Q3 &pn = V[v1n].p, &pp = V[v1p].p;
			if (v1 == h1.minXy) h1.minXy = (pn.x < pp.x || (pn.x == pp.x && pn.y < pp.y)) ? v1n : v1p;
			if (v1 == h1.maxXy) h1.maxXy = (pn.x > pp.x || (pn.x == pp.x && pn.y > pp.y)) ? v1n : v1p;
		}

		v0 = h0.maxXy;
		v1 = h1.maxXy;
		int v00 = -1, v10 = -1;
		int sign = 1;
		for (int side = 0; side <= 1; side++)
		{
			int dx = (V[v1].p.x - V[v0].p.x) * sign;
			if (dx > 0)
			{
				while (true)
				{
					int dy = V[v1].p.y - V[v0].p.y;
					int w0 = side ? V[v0].next : V[v0].prev;
					if (w0 != v0)
					{
						int dx0 = (V[w0].p.x - V[v0].p.x) * sign;
						int dy0 = V[w0].p.y - V[v0].p.y;
						if ((dy0 <= 0) && ((dx0 == 0) || ((dx0 < 0) && (dy0 * dx <= dy * dx0))))
						{
							v0 = w0;
							dx = (V[v1].p.x - V[v0].p.x) * sign;
							continue;
						}
					}
					int w1 = side ? V[v1].next : V[v1].prev;
					if (w1 != v1)
					{
						int dx1 = (V[w1].p.x - V[v1].p.x) * sign;
						int dy1 = V[w1].p.y - V[v1].p.y;
						int dxn = (V[w1].p.x - V[v0].p.x) * sign;
						if ((dxn > 0) && (dy1 < 0) && ((dx1 == 0) || ((dx1 < 0) && (dy1 * dx < dy * dx1))))
						{
							v1 = w1;
							dx = dxn;
							continue;
						}
					}
					break;
				}
			}
			else if (dx < 0)
			{
				while (true)
				{
					int dy = V[v1].p.y - V[v0].p.y;
					int w1 = side ? V[v1].prev : V[v1].next;
					if (w1 != v1)
					{
						int dx1 = (V[w1].p.x - V[v1].p.x) * sign;
						int dy1 = V[w1].p.y - V[v1].p.y;
						if ((dy1 >= 0) && ((dx1 == 0) || ((dx1 < 0) && (dy1 * dx <= dy * dx1))))
						{
							v1 = w1;
							dx = (V[v1].p.x - V[v0].p.x) * sign;
							continue;
						}
					}
					int w0 = side ? V[v0].prev : V[v0].next;
					if (w0 != v0)
					{
						int dx0 = (V[w0].p.x - V[v0].p.x) * sign;
						int dy0 = V[w0].p.y - V[v0].p.y;
						int dxn = (V[v1].p.x - V[w0].p.x) * sign;
						if ((dxn < 0) && (dy0 > 0) && ((dx0 == 0) || ((dx0 < 0) && (dy0 * dx < dy * dx0))))
						{
							v0 = w0;
							dx = dxn;
							continue;
						}
					}
					break;
				}
			}
			else
			{
				int x = V[v0].p.x;
				int y0 = V[v0].p.y;
				int w0 = v0, t;
				while (((t = side ? V[w0].next : V[w0].prev) != v0) && (V[t].p.x == x) && (V[t].p.y <= y0))
				{
					w0 = t;
					y0 = V[t].p.y;
				}
				v0 = w0;
				int y1 = V[v1].p.y;
				int w1 = v1;
				while (((t = side ? V[w1].prev : V[w1].next) != v1) && (V[t].p.x == x) && (V[t].p.y >= y1))
				{
					w1 = t;
					y1 = V[t].p.y;
				}
				v1 = w1;
			}
			if (side == 0)
			{
				v00 = v0;
				v10 = v1;
				v0 = h0.minXy;
				v1 = h1.minXy;
				sign = -1;
			}
		}
		V[v0].prev = v1;
		V[v1].next = v0;
		V[v00].next = v10;
		V[v10].prev = v00;
		if (V[h1.minXy].p.x < V[h0.minXy].p.x) h0.minXy = h1.minXy;
		if (V[h1.maxXy].p.x >= V[h0.maxXy].p.x) h0.maxXy = h1.maxXy;
		h0.maxYx = h1.maxYx;
		c0 = v00;
		c1 = v10;
		return true;
	}

	/* merge h1 into h0: bridge, then wrap new faces around both (:1651-1936) */
	void merge(Span& h0, Span& h1)
	{
		if (h1.maxXy < 0) return;
		if (h0.maxXy < 0)
		{
			h0 = h1;
			return;
		}
		stamp--;
		int c0 = -1, c1 = -1;
		int toPrev0 = -1, firstNew0 = -1, pendHead0 = -1, pendTail0 = -1;
		int toPrev1 = -1, firstNew1 = -1, pendHead1 = -1, pendTail1 = -1;
		Q3 prevPoint;
		const Q3 down = {0, 0, -1};

		if (join_projections(h0, h1, c0, c1))
		{
			Q3 s = diff(c1, c0);
			W3 normal = qcross(down, s);
			W3 t = qcross(s, normal);
			int start0 = -1;
			int x = V[c0].edges;
			if (x >= 0)
			{
				do
				{
					Q3 d = diff(E[x].target, c0);
					if (qdot(d, normal) == 0 && qdot(d, t) > 0)
						if (start0 < 0 || turn(start0, x, s, down) == TURN_CW) start0 = x;
					x = E[x].next;
				} while (x != V[c0].edges);
			}
			int start1 = -1;
			x = V[c1].edges;
			if (x >= 0)
			{
				do
				{
					Q3 d = diff(E[x].target, c1);
					if (qdot(d, normal) == 0 && qdot(d, t) > 0)
						if (start1 < 0 || turn(start1, x, s, down) == TURN_CCW) start1 = x;
					x = E[x].next;
				} while (x != V[c1].edges);
			}
			if (start0 >= 0 || start1 >= 0)
			{
				coplanar_advance(c0, c1, start0, start1, -1, -1);
				if (start0 >= 0) c0 = E[start0].target;
				if (start1 >= 0) c1 = E[start1].target;
			}
			prevPoint = V[c1].p;
			prevPoint.z++;
		}
		else
		{
			prevPoint = V[c1].p;
			prevPoint.x++;
		}

		int first0 = c0, first1 = c1;
		bool firstRun = true;
		while (true)
		{
			Q3 s = diff(c1, c0);
			Q3 r = qsub(prevPoint, V[c0].p);
			W3 rxs = qcross(r, s);
			W3 sxrxs = qcross(s, rxs);
			Ratio best0, best1;
			int min0 = max_angle(false, c0, s, rxs, sxrxs, best0);
			int min1 = max_angle(true, c1, s, rxs, sxrxs, best1);
			if (min0 < 0 && min1 < 0)
			{
				int a = new_pair(c0, c1);
				link(a, a);
				V[c0].edges = a;
				a = E[a].rev;
				link(a, a);
				V[c1].edges = a;
				return;
			}
			int cmp = min0 < 0 ? 1 : (min1 < 0 ? -1 : best0.cmp(best1));
			if (firstRun || ((cmp >= 0) ? !best1.negInf() : !best0.negInf()))
			{
				int a = new_pair(c0, c1);
				if (pendTail0 >= 0)
					E[pendTail0].prev = a;
				else
					pendHead0 = a;
				E[a].next = pendTail0;
				pendTail0 = a;
				a = E[a].rev;
				if (pendTail1 >= 0)
					E[pendTail1].next = a;
				else
					pendHead1 = a;
				E[a].prev = pendTail1;
				pendTail1 = a;
			}
			int e0 = min0, e1 = min1;
			if (cmp == 0) coplanar_advance(c0, c1, e0, e1, -1, -1);

			if (cmp >= 0 && e1 >= 0)
			{
				if (toPrev1 >= 0)
				{
					for (int x = E[toPrev1].next, n; x != min1; x = n)
					{
						n = E[x].next;
						drop_pair(x);
					}
				}
				if (pendTail1 >= 0)
				{
					if (toPrev1 >= 0)
						link(toPrev1, pendHead1);
					else
					{
						link(E[min1].prev, pendHead1);
						firstNew1 = pendHead1;
					}
					link(pendTail1, min1);
					pendHead1 = pendTail1 = -1;
				}
				else if (toPrev1 < 0)
					firstNew1 = min1;
				prevPoint = V[c1].p;
				c1 = E[e1].target;
				toPrev1 = E[e1].rev;
			}
			if (cmp <= 0 && e0 >= 0)
			{
				if (toPrev0 >= 0)
				{
					for (int x = E[toPrev0].prev, n; x != min0; x = n)
					{
						n = E[x].prev;
						drop_pair(x);
					}
				}
				if (pendTail0 >= 0)
				{
					if (toPrev0 >= 0)
						link(pendHead0, toPrev0);
					else
					{
						link(pendHead0, E[min0].next);
						firstNew0 = pendHead0;
					}
					link(min0, pendTail0);
					pendHead0 = pendTail0 = -1;
				}
				else if (toPrev0 < 0)
					firstNew0 = min0;
				prevPoint = V[c0].p;
				c0 = E[e0].target;
				toPrev0 = E[e0].rev;
			}

			if (c0 == first0 && c1 == first1)
			{
				if (toPrev0 < 0)
				{
					link(pendHead0, pendTail0);
					V[c0].edges = pendTail0;
				}
				else
				{
					for (int x = E[toPrev0].prev, n; x != firstNew0; x = n)
					{
						n = E[x].prev;
						drop_pair(x);
					}
					if (pendTail0 >= 0)
					{
						link(pendHead0, toPrev0);
						link(firstNew0, pendTail0);
					}
				}
				if (toPrev1 < 0)
				{
					link(pendTail1, pendHead1);
					V[c1].edges = pendTail1;
				}
				else
				{
					for (int x = E[toPrev1].next, n; x != firstNew1; x = n)
					{
						n = E[x].next;
						drop_pair(x);
					}
					if (pendTail1 >= 0)
					{
						link(toPrev1, pendHead1);
						link(pendTail1, firstNew1);
					}
				}
				return;
			}
			firstRun = false;
		}
	}

	/* hull of the sorted points [lo, hi) (:1209-1324) */
	void build(int lo, int hi, Span& out)
	{
		int n = hi - lo;
		if (n == 0)
		{
			out = Span();
			return;
		}
		if (n == 2 && V[lo].p != V[lo + 1].p)
		{
			int a = lo, b = lo + 1;
			int dx = V[a].p.x - V[b].p.x, dy = V[a].p.y - V[b].p.y;
			if (dx == 0 && dy == 0)
			{
				if (V[a].p.z > V[b].p.z) std::swap(a, b);
				V[a].next = V[a].prev = a;
				out.minXy = out.maxXy = out.minYx = out.maxYx = a;
			}
			else
			{
				V[a].next = V[a].prev = b;
				V[b].next = V[b].prev = a;
				if (dx < 0 || (dx == 0 && dy < 0))
				{
					out.minXy = a;
					out.maxXy = b;
				}
				else
				{
					out.minXy = b;
					out.maxXy = a;
				}
				if (dy < 0 || (dy == 0 && dx < 0))
				{
					out.minYx = a;
					out.maxYx = b;
				}
				else
				{
					out.minYx = b;
					out.maxYx = a;
				}
			}
			int x = new_pair(a, b);
			link(x, x);
			V[a].edges = x;
			x = E[x].rev;
			link(x, x);
			V[b].edges = x;
			return;
		}
		if (n <= 2)
		{
			V[lo].edges = -1;
			V[lo].next = V[lo].prev = lo;
			out.minXy = out.maxXy = out.minYx = out.maxYx = lo;
			return;
		}
		int split0 = lo + n / 2;
		Q3 p = V[split0 - 1].p;
		int split1 = split0;
		while (split1 < hi && V[split1].p == p) split1++;
		build(lo, split0, out);
		Span h1;
		build(split1, hi, h1);
		merge(out, h1);
	}
};

struct F3
{
	float x, y, z;
};
inline F3 fsub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 fadd(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 fmul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 fcross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline F3 fnormalized(F3 a) { return fmul(a, 1.0f / std::sqrt(fdot(a, a))); } /* b3Vector3::normalize, scalar path */
inline float comp(const F3& a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline void set_comp(F3& a, int k, float v) { (k == 0 ? a.x : (k == 1 ? a.y : a.z)) = v; }

/* Stage 1 driver: quantise, sort, hull, then number vertices and list faces
 * as b3ConvexHullComputer::compute does.  Returns false when fewer than
 * three distinct points remain (no face). */
bool exact_hull(const float* pts, int n, int strideFloats, std::vector<F3>& verts, std::vector<std::vector<int>>& faces)
{
	/* bounding box, axis order and grid (b3ConvexHullComputer.cpp:1943-2002) */
	F3 mn = {1e30f, 1e30f, 1e30f}, mx = {-1e30f, -1e30f, -1e30f};
	for (int i = 0; i < n; i++)
	{
		const float* q = pts + (size_t)i * strideFloats;
		for (int k = 0; k < 3; k++)
		{
			if (q[k] < comp(mn, k)) set_comp(mn, k, q[k]);
			if (comp(mx, k) < q[k]) set_comp(mx, k, q[k]);
		}
	}
	F3 s = fsub(mx, mn);
	int maxAxis = s.x < s.y ? (s.y < s.z ? 2 : 1) : (s.x < s.z ? 2 : 0);
	int minAxis = s.x < s.y ? (s.x < s.z ? 0 : 2) : (s.y < s.z ? 1 : 2);
	if (minAxis == maxAxis) minAxis = (maxAxis + 1) % 3;
	int medAxis = 3 - maxAxis - minAxis;
	s = fmul(s, 1.0f / 10216.0f);
	if (((medAxis + 1) % 3) != maxAxis) s = fmul(s, -1.0f);
	F3 scaling = s;
	for (int k = 0; k < 3; k++)
		if (comp(s, k) != 0) set_comp(s, k, 1.0f / comp(s, k));
	F3 center = fmul(fadd(mn, mx), 0.5f);

	Hull3 h;
	h.V.resize(n);
	std::vector<Q3> grid(n);
	for (int i = 0; i < n; i++)
	{
		const float* q = pts + (size_t)i * strideFloats;
		F3 p = {(q[0] - center.x) * s.x, (q[1] - center.y) * s.y, (q[2] - center.z) * s.z};
		grid[i] = {(int)comp(p, medAxis), (int)comp(p, maxAxis), (int)comp(p, minAxis)};
	}
	/* order by (y, x, z); equal points are interchangeable (only coordinates
	 * reach the output), so any sort gives the reference's sequence */
	std::sort(grid.begin(), grid.end(), [](const Q3& a, const Q3& b) {
		return a.y < b.y || (a.y == b.y && (a.x < b.x || (a.x == b.x && a.z < b.z)));
	});
	for (int i = 0; i < n; i++) h.V[i].p = grid[i];
	h.E.reserve(12 * (size_t)n);
	Span all;
	h.build(0, n, all);
	int first = all.minXy;
	if (first < 0) return false;

	/* number vertices breadth first along each vertex's edge ring, mark every
	 * half-edge with its output slot (b3ConvexHullComputer.cpp:2666-2712) */
	std::vector<int> order;
	auto copy_of = [&](int vi) {
		if (h.V[vi].copy < 0)
		{
			h.V[vi].copy = (int)order.size();
			order.push_back(vi);
		}
		return h.V[vi].copy;
	};
	copy_of(first);
	int slots = 0;
	for (size_t c = 0; c < order.size(); c++)
	{
		int vi = order[c];
		int e0 = h.V[vi].edges;
		if (e0 < 0) continue;
		int x = e0;
		do
		{
			if (h.E[x].stamp < 0)
			{
				h.E[x].stamp = slots;
				h.E[h.E[x].rev].stamp = slots + 1;
				slots += 2;
				copy_of(h.E[x].target);
			}
			x = h.E[x].next;
		} while (x != e0);
	}
	verts.resize(order.size());
	for (size_t c = 0; c < order.size(); c++)
	{
		const Q3& g = h.V[order[c]].p;
		F3 p = {0, 0, 0};
		set_comp(p, medAxis, (float)g.x);
		set_comp(p, maxAxis, (float)g.y);
		set_comp(p, minAxis, (float)g.z);
		verts[c] = {p.x * scaling.x + center.x, p.y * scaling.y + center.y, p.z * scaling.z + center.z};
	}
	/* faces: first unused half-edge of each ring, in vertex order; a face's
	 * vertices are the sources of its half-edges (:2714-2742) */
	faces.clear();
	for (size_t c = 0; c < order.size(); c++)
	{
		int e0 = h.V[order[c]].edges;
		if (e0 < 0) continue;
		int x = e0;
		do
		{
			if (h.E[x].stamp >= 0)
			{
				std::vector<int> f;
				int y = x;
				do
				{
					f.push_back(h.V[h.E[h.E[y].rev].target].copy);
					h.E[y].stamp = -1;
					y = h.E[h.E[y].rev].prev;
				} while (y != x);
				faces.push_back(f);
			}
			x = h.E[x].next;
		} while (x != e0);
	}
	return !faces.empty();
}

/* ---------------------------------------------------------------- stage 2 */

struct Face
{
	std::vector<int> idx;
	float plane[4];
};

/* b3PlaneSpace1 (b3Vector3.h:1184-1212) */
void plane_space(F3 n, F3& p, F3& q)
{
	if (std::fabs(n.z) > 0.7071067811865475244008443621048490f)
	{
		float a = n.y * n.y + n.z * n.z;
		float k = 1.0f / std::sqrt(a);
		p = {0, -n.z * k, n.y * k};
		q = {a * k, -n.x * p.z, n.x * p.y};
	}
	else
	{
		float a = n.x * n.x + n.y * n.y;
		float k = 1.0f / std::sqrt(a);
		p = {-n.y * k, n.x * k, 0};
		q = {-n.z * p.y, n.z * p.x, a * k};
	}
}

struct GPoint
{
	F3 p;
	float angle;
	int id;
};

/* 2-D Graham scan in the plane of normal n (b3GrahamScan2dConvexHull.h:59-114):
 * anchor = least projection on the plane's first axis (ties keep the earlier
 * point), angular order with length and index as tie breaks; a point that
 * meets a hull reduced to the anchor alone is dropped, as the reference's
 * backtracking loop does. */
std::vector<int> graham(std::vector<GPoint>& pts, F3 n)
{
	std::vector<int> out;
	F3 a0, a1;
	plane_space(n, a0, a1);
	if (pts.size() <= 1)
	{
		for (size_t i = 0; i < pts.size(); i++) out.push_back(pts[0].id);
		return out;
	}
	for (size_t i = 0; i < pts.size(); i++)
		if (fdot(pts[i].p, a0) < fdot(pts[0].p, a0)) std::swap(pts[0], pts[i]);
	pts[0].angle = -1e30f;
	for (size_t i = 1; i < pts.size(); i++)
	{
		F3 ar = fsub(pts[i].p, pts[0].p);
		pts[i].angle = fdot(fcross(a0, ar), n) / std::sqrt(fdot(ar, ar));
	}
	F3 anchor = pts[0].p;
	std::sort(pts.begin() + 1, pts.end(), [&](const GPoint& a, const GPoint& b) {
		if (a.angle != b.angle) return a.angle < b.angle;
		F3 da = fsub(a.p, anchor), db = fsub(b.p, anchor);
		float la = fdot(da, da), lb = fdot(db, db);
		if (la != lb) return la < lb;
		return a.id < b.id;
	});
	std::vector<GPoint> h;
	h.push_back(pts[0]);
	h.push_back(pts[1]);
	for (size_t i = 2; i < pts.size(); i++)
	{
		bool convex = false;
		while (!convex && h.size() > 1)
		{
			F3 a = h[h.size() - 2].p, b = h[h.size() - 1].p;
			convex = fdot(fcross(fsub(a, b), fsub(a, pts[i].p)), n) > 0;
			if (!convex)
				h.pop_back();
			else
				h.push_back(pts[i]);
		}
	}
	for (auto& g : h) out.push_back(g.id);
	return out;
}

inline bool almost_zero(F3 d) /* b3ConvexUtility.cpp:250-255: float against double 1e-6 */
{
	return !(std::fabs(d.x) > 1e-6 || std::fabs(d.y) > 1e-6 || std::fabs(d.z) > 1e-6);
}
}  // namespace

int b3h_build_polyhedron(const float* ptsIn, int n, int strideFloats, b3hPolyhedron* out)
{
	out->vertices.clear();
	out->faces.clear();
	out->uniqueEdges.clear();
	if (n <= 0) return -1;
	std::vector<F3> vv;
	std::vector<std::vector<int>> hf;
	if (!exact_hull(ptsIn, n, strideFloats, vv, hf)) return -1;

	/* face planes (b3ConvexUtility.cpp:51-98) */
	std::vector<Face> tmp(hf.size());
	for (size_t i = 0; i < hf.size(); i++)
	{
		Face& f = tmp[i];
		f.idx = hf[i];
		int m = (int)f.idx.size();
		F3 e0 = fnormalized(fsub(vv[f.idx[1 % m]], vv[f.idx[0]]));
		F3 e1 = fnormalized(fsub(vv[f.idx[2 % m]], vv[f.idx[1 % m]]));
		F3 nn = fnormalized(fcross(e0, e1));
		float pe = 1e30f;
		for (int vi : f.idx)
		{
			float d = fdot(vv[vi], nn);
			if (pe > d) pe = d;
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
		F3 na = {tmp[ref].plane[0], tmp[ref].plane[1], tmp[ref].plane[2]};
		for (int j = (int)todo.size() - 1; j >= 0; j--)
		{
			int i = todo[j];
			F3 nb = {tmp[i].plane[0], tmp[i].plane[1], tmp[i].plane[2]};
			if (fdot(na, nb) > 0.999f)
			{
				group.push_back(i);
				/* removal by swap with the last element */
				std::swap(todo[j], todo.back());
				todo.pop_back();
			}
		}
		bool merged = false;
		if (group.size() > 1)
		{
			std::vector<GPoint> pts2;
			F3 avg = {0, 0, 0};
			for (int gi : group)
			{
				avg = fadd(avg, F3{tmp[gi].plane[0], tmp[gi].plane[1], tmp[gi].plane[2]});
				for (int vi : tmp[gi].idx)
				{
					bool seen = false;
					for (auto& g : pts2)
						if (g.id == vi)
						{
							seen = true;
							break;
						}
					if (!seen) pts2.push_back({vv[vi], 0.f, vi});
				}
			}
			avg = fnormalized(avg);
			std::vector<int> h = graham(pts2, avg);
			bool reject = false;
			for (auto& g : pts2)
			{
				if (std::find(h.begin(), h.end(), g.id) != h.end()) continue;
				for (int fj = 0; fj < (int)tmp.size() && !reject; fj++)
				{
					if (std::find(group.begin(), group.end(), fj) != group.end()) continue;
					for (int u : tmp[fj].idx)
						if (u == g.id)
						{
							reject = true;
							break;
						}
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

	/* unique edge directions, smaller index minus larger (b3ConvexUtility.cpp:325-365) */
	std::vector<F3> ue;
	for (auto& f : faces)
	{
		int m = (int)f.idx.size();
		for (int j = 0; j < m; j++)
		{
			int a = f.idx[j], b = f.idx[(j + 1) % m];
			int hi = a > b ? a : b, lo = a > b ? b : a;
			F3 e = fnormalized(fsub(vv[lo], vv[hi]));
			bool found = false;
			for (auto& u : ue)
				if (almost_zero(fsub(u, e)) || almost_zero(fadd(u, e)))
				{
					found = true;
					break;
				}
			if (!found) ue.push_back(e);
		}
	}

	for (auto& v : vv)
	{
		out->vertices.push_back(v.x);
		out->vertices.push_back(v.y);
		out->vertices.push_back(v.z);
	}
	for (auto& f : faces)
	{
		b3hPolyFace pf;
		for (int k = 0; k < 4; k++) pf.plane[k] = f.plane[k];
		pf.indices = f.idx;
		out->faces.push_back(pf);
	}
	for (auto& e : ue)
	{
		out->uniqueEdges.push_back(e.x);
		out->uniqueEdges.push_back(e.y);
		out->uniqueEdges.push_back(e.z);
	}
	return 0;
}
