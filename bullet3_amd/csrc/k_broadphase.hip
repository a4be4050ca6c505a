/*
 * k_broadphase.hip -- uniform-grid and 1-axis sweep-and-prune broadphase.
 *
 * Grid (b3GpuGridBroadphase::calculateOverlappingPairs, b3GpuGridBroadphase.cpp:117-259;
 *       kernels gridBroadphase.cl:3-211):
 *   hash of the AABB centre cell (wrapped to the grid dims) -> stable radix sort
 *   of (hash, proxy) -> first index of every hash run -> per sorted proxy a scan
 *   of the 27 neighbour cells (bucket capped at maxBodiesPerCell, inclusive AABB
 *   test, pair emitted once as (lower, higher) body index).
 * SAP (b3GpuSapBroadphase::calculateOverlappingPairs, b3GpuSapBroadphase.cpp:997-1231;
 *      kernels sap.cl:231-389): axis of largest centre variance, stable sort by
 *   the order-preserving bits of min[axis], sweep j = i+1.. while
 *   min_j[axis] <= max_i[axis]; pairs are (sorted-earlier, later) as in the reference.
 * Large x small pairs (the static ground): computePairsKernelTwoArrays, sap.cl:66-105.
 *
 * Pair compaction is atomic-free: a counting pass, a prefix sum and a writing
 * pass.  The pair buffer therefore has a fixed order -- large x small pairs
 * (small-major) first, then per sorted proxy in neighbour/bucket order -- that
 * does not depend on scheduling.
 */
#include "b3h_internal.h"

__device__ __forceinline__ bool bp_overlap(const b3hAabb& a, const b3hAabb& b)
{
	return (a.m_min.x <= b.m_max.x) && (b.m_min.x <= a.m_max.x) && (a.m_min.y <= b.m_max.y) &&
		   (b.m_min.y <= a.m_max.y) && (a.m_min.z <= b.m_max.z) && (b.m_min.z <= a.m_max.z);
}
__device__ __forceinline__ int bp_handle(const b3hAabb& a) { return f2i(a.m_min.w); }
/* batched scenes (b3h_pipeline_set_scenes): proxies of different scenes never pair */
__device__ __forceinline__ bool bp_same_scene(const int* sceneOf, int ha, int hb)
{
	return !sceneOf || sceneOf[ha] == sceneOf[hb];
}

/* ---------------- large x small ---------------- */
template <bool WRITE>
__global__ __launch_bounds__(256) void k_large_small(int nSmall, int nLarge, const b3hAabb* __restrict__ all,
													 const int* __restrict__ smallMap, const int* __restrict__ largeMap,
													 int* __restrict__ counts, const int* __restrict__ offsets,
													 b3hInt4* __restrict__ pairs, int maxPairs,
													 const int* __restrict__ sceneOf, int normalise)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nSmall) return;
	b3hAabb s = all[smallMap[i]];
	int hs = bp_handle(s);
	int c = 0;
	int o = WRITE ? offsets[i] : 0;
	for (int j = 0; j < nLarge; j++)
	{
		b3hAabb l = all[largeMap[j]];
		if (bp_overlap(l, s) && bp_same_scene(sceneOf, bp_handle(l), hs))
		{
			if (WRITE)
			{
				int x = bp_handle(l), y = hs;
				if (normalise && x > y)
				{
					int t = x;
					x = y;
					y = t;
				}
				int at = o + c;
				if (at < maxPairs) pairs[at] = b3hInt4{x, y, -1, -1};
			}
			c++;
		}
	}
	if (!WRITE) counts[i] = c;
}

/* Many large proxies (batched scenes: one ground per scene): sort them by
 * min.x once per step; a small proxy then tests only the large proxies whose
 * min.x lies in [small.min.x - 2*maxExtent - 1, small.max.x] (every overlapping
 * large proxy is in that window; the exact overlap test decides), and emits
 * its matches in ascending large-proxy order, as the all-pairs loop does. */
__global__ __launch_bounds__(256) void k_large_keys(int nLarge, const b3hAabb* __restrict__ all,
													const int* __restrict__ largeMap, unsigned* __restrict__ keys,
													unsigned* __restrict__ vals, unsigned* __restrict__ maxExt)
{
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nLarge) return;
	b3hAabb l = all[largeMap[j]];
	unsigned u = f2u(l.m_min.x);
	unsigned mask = (unsigned)(-(int)(u >> 31)) | 0x80000000u;
	keys[j] = u ^ mask;
	vals[j] = (unsigned)j;
	float e = l.m_max.x - l.m_min.x;
	if (e > 0.f) atomicMax(maxExt, f2u(e));
}
__global__ __launch_bounds__(256) void k_large_minx(int nLarge, const b3hAabb* __restrict__ all,
													const int* __restrict__ largeMap, const unsigned* __restrict__ vals,
													float* __restrict__ minx)
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nLarge) return;
	minx[k] = all[largeMap[vals[k]]].m_min.x;
}
#define LS_LOCAL 16
template <bool WRITE>
__global__ __launch_bounds__(256) void k_large_small_sorted(int nSmall, int nLarge, const b3hAabb* __restrict__ all,
															const int* __restrict__ smallMap,
															const int* __restrict__ largeMap,
															const unsigned* __restrict__ lvals,
															const float* __restrict__ lminx,
															const unsigned* __restrict__ maxExt,
															int* __restrict__ counts, const int* __restrict__ offsets,
															b3hInt4* __restrict__ pairs, int maxPairs,
															const int* __restrict__ sceneOf, int normalise)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nSmall) return;
	b3hAabb s = all[smallMap[i]];
	int hs = bp_handle(s);
	float ext = __builtin_bit_cast(float, *maxExt);
	float t = s.m_min.x - 2.f * ext - 1.f;
	/* first sorted position with min.x >= t */
	int lo = 0, hi = nLarge;
	while (lo < hi)
	{
		int mid = (lo + hi) >> 1;
		if (lminx[mid] < t)
			lo = mid + 1;
		else
			hi = mid;
	}
	int found[LS_LOCAL];
	int nf = 0;
	bool overflow = false;
	for (int k = lo; k < nLarge && lminx[k] <= s.m_max.x; k++)
	{
		int j = (int)lvals[k];
		b3hAabb l = all[largeMap[j]];
		if (bp_overlap(l, s) && bp_same_scene(sceneOf, bp_handle(l), hs))
		{
			if (nf < LS_LOCAL)
				found[nf++] = j;
			else
				overflow = true;
		}
	}
	int c = 0;
	int o = WRITE ? offsets[i] : 0;
	if (!overflow)
	{
		/* ascending large order (insertion sort of a handful) */
		for (int a = 1; a < nf; a++)
		{
			int v = found[a], b = a - 1;
			while (b >= 0 && found[b] > v)
			{
				found[b + 1] = found[b];
				b--;
			}
			found[b + 1] = v;
		}
		for (int a = 0; a < nf; a++)
		{
			if (WRITE)
			{
				int x = bp_handle(all[largeMap[found[a]]]), y = hs;
				if (normalise && x > y)
				{
					int tmp = x;
					x = y;
					y = tmp;
				}
				int at = o + c;
				if (at < maxPairs) pairs[at] = b3hInt4{x, y, -1, -1};
			}
			c++;
		}
	}
	else
	{
		/* more matches than the local list holds: the all-pairs loop */
		for (int j = 0; j < nLarge; j++)
		{
			b3hAabb l = all[largeMap[j]];
			if (bp_overlap(l, s) && bp_same_scene(sceneOf, bp_handle(l), hs))
			{
				if (WRITE)
				{
					int x = bp_handle(l), y = hs;
					if (normalise && x > y)
					{
						int tmp = x;
						x = y;
						y = tmp;
					}
					int at = o + c;
					if (at < maxPairs) pairs[at] = b3hInt4{x, y, -1, -1};
				}
				c++;
			}
		}
	}
	if (!WRITE) counts[i] = c;
}

static inline int nblk(int n) { return (n + 255) / 256; }
/* large x small pairs: all-pairs for a few large proxies, the sorted window
 * for many; prepare() once per step before the counting pass */
static void large_small_prepare(const b3hAabb* all, const int* largeMap, int nLarge, b3hBroadphaseScratch& sc,
								hipStream_t s)
{
	if (nLarge <= B3H_LARGE_BRUTE_MAX) return;
	b3h_fill_i32((int*)sc.lmaxExt, 0, 1, s);
	hipLaunchKernelGGL(k_large_keys, dim3(nblk(nLarge)), dim3(256), 0, s, nLarge, all, largeMap, sc.lkeys, sc.lvals,
					   sc.lmaxExt);
	b3h_radix_sort(sc.lkeys, sc.lvals, sc.ltmpKeys, sc.ltmpVals, sc.lhist, nLarge, 32, s);
	hipLaunchKernelGGL(k_large_minx, dim3(nblk(nLarge)), dim3(256), 0, s, nLarge, all, largeMap, sc.lvals, sc.lminx);
}
/* ---------------- grid ---------------- */
__device__ __forceinline__ void grid_pos(const b3hAabb& a, const b3hGridParams& gp, int g[3])
{
	float px = (a.m_min.x + a.m_max.x) * 0.5f;
	float py = (a.m_min.y + a.m_max.y) * 0.5f;
	float pz = (a.m_min.z + a.m_max.z) * 0.5f;
	g[0] = (int)floorf(px * gp.invCell[0]) & (gp.dims[0] - 1);
	g[1] = (int)floorf(py * gp.invCell[1]) & (gp.dims[1] - 1);
	g[2] = (int)floorf(pz * gp.invCell[2]) & (gp.dims[2] - 1);
}
__device__ __forceinline__ int grid_hash(int x, int y, int z, const b3hGridParams& gp)
{
	x &= gp.dims[0] - 1;
	y &= gp.dims[1] - 1;
	z &= gp.dims[2] - 1;
	return z * gp.dims[1] * gp.dims[0] + y * gp.dims[0] + x;
}

__global__ __launch_bounds__(256) void k_grid_hash(int n, const b3hAabb* __restrict__ all, const int* __restrict__ smallMap,
												   b3hGridParams gp, unsigned* __restrict__ keys, unsigned* __restrict__ vals)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int g[3];
	grid_pos(all[smallMap[i]], gp, g);
	keys[i] = (unsigned)grid_hash(g[0], g[1], g[2], gp);
	vals[i] = (unsigned)i;
}

__global__ __launch_bounds__(256) void k_cell_start(int n, const unsigned* __restrict__ keys, int* __restrict__ cellStart)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	unsigned k = keys[i];
	if (i == 0 || keys[i - 1] != k) cellStart[k] = i;
}

/* AABBs in hash-sorted order, so the neighbour scans read contiguous records */
__global__ __launch_bounds__(256) void k_grid_gather(int n, const unsigned* __restrict__ vals, const b3hAabb* __restrict__ all,
													 const int* __restrict__ smallMap, b3hAabb* __restrict__ sorted)
{
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	sorted[s] = all[smallMap[vals[s]]];
}

/* the neighbour scan of one small proxy (gridBroadphase.cl:3-211): MODE 0
 * counts its pairs and keeps the first B3H_GRID_STASH in its stash slots;
 * MODE 1 writes all of them at its offset (used for proxies with more pairs
 * than the stash holds).  Pairs (ha < hb) come in neighbour-cell, bucket order. */
template <int MODE>
__device__ __forceinline__ int grid_scan_one(int s, int n, const unsigned* __restrict__ keys,
											 const b3hAabb* __restrict__ sorted, const int* __restrict__ cellStart,
											 const b3hGridParams& gp, int2* __restrict__ stash, int o,
											 b3hInt4* __restrict__ pairs, int maxPairs, const int* __restrict__ sceneOf)
{
	b3hAabb a = sorted[s];
	int ha = bp_handle(a);
	int g[3];
	grid_pos(a, gp, g);
	int c = 0;
	for (int z = -1; z <= 1; z++)
		for (int y = -1; y <= 1; y++)
			for (int x = -1; x <= 1; x++)
			{
				int h = grid_hash(g[0] + x, g[1] + y, g[2] + z, gp);
				int st = cellStart[h];
				if (st == -1) continue;
				int end = st + gp.maxBodiesPerCell;
				if (end > n) end = n;
				if (st >= end) continue;
				/* the next entry's key and box are read while this one is
				 * tested (every index read is below end <= n) */
				unsigned kk = keys[st];
				b3hAabb b = sorted[st];
				for (int k = st;; k++)
				{
					if ((int)kk != h) break;
					const bool more = k + 1 < end;
					unsigned kkN = 0u;
					b3hAabb bN = b;
					if (more)
					{
						kkN = keys[k + 1];
						bN = sorted[k + 1];
					}
					if (k != s && bp_overlap(a, b))
					{
						int hb = bp_handle(b);
						if (ha < hb && bp_same_scene(sceneOf, ha, hb))
						{
							if (MODE == 0)
							{
								if (c < B3H_GRID_STASH) stash[(size_t)s * B3H_GRID_STASH + c] = make_int2(ha, hb);
							}
							else
							{
								int at = o + c;
								if (at < maxPairs) pairs[at] = b3hInt4{ha, hb, -1, -1};
							}
							c++;
						}
					}
					if (!more) break;
					kk = kkN;
					b = bN;
				}
			}
	return c;
}
/* one pass finds every pair and stashes the first few per proxy ... */
__global__ __launch_bounds__(256) void k_grid_pairs(int n, const unsigned* __restrict__ keys,
													const b3hAabb* __restrict__ sorted,
													const int* __restrict__ cellStart, b3hGridParams gp,
													int* __restrict__ counts, int2* __restrict__ stash,
													const int* __restrict__ sceneOf)
{
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	counts[s] = grid_scan_one<0>(s, n, keys, sorted, cellStart, gp, stash, 0, nullptr, 0, sceneOf);
}
/* ... and after the scan of the counts each proxy copies its stash to its
 * offset; a proxy with more pairs than its stash holds scans again and writes
 * them all.  The result is the two-pass (count, write) list, in its order. */
__global__ __launch_bounds__(256) void k_grid_emit(int n, const unsigned* __restrict__ keys,
												   const b3hAabb* __restrict__ sorted,
												   const int* __restrict__ cellStart, b3hGridParams gp,
												   const int* __restrict__ counts, const int* __restrict__ offsets,
												   const int2* __restrict__ stash, b3hInt4* __restrict__ pairs,
												   int maxPairs, const int* __restrict__ sceneOf)
{
	int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	int c = counts[s], o = offsets[s];
	if (c <= B3H_GRID_STASH)
	{
		for (int j = 0; j < c; j++)
		{
			int2 v = stash[(size_t)s * B3H_GRID_STASH + j];
			if (o + j < maxPairs) pairs[o + j] = b3hInt4{v.x, v.y, -1, -1};
		}
	}
	else
		grid_scan_one<1>(s, n, keys, sorted, cellStart, gp, nullptr, o, pairs, maxPairs, sceneOf);
}

__global__ void k_pairs_finish(const int* __restrict__ total, int maxPairs, int* __restrict__ pairCount,
							   int* __restrict__ pairTotal)
{
	int t = *total;
	*pairTotal = t;
	*pairCount = t < maxPairs ? t : maxPairs;
}

/* normalise: pairs as (lower, higher) handle (grid, SAP: sap.cl:66-105);
 * otherwise (large, small) as the BVH broadphase writes them
 * (parallelLinearBvh.cl plbvhLargeAabbAabbTest) */
template <bool WRITE>
static void large_small_pass(int nSmall, int nLarge, const b3hAabb* all, const int* smallMap, const int* largeMap,
							 b3hBroadphaseScratch& sc, int* counts, const int* offsets, b3hInt4* pairs, int maxPairs,
							 hipStream_t s, int normalise = 1)
{
	if (nLarge <= B3H_LARGE_BRUTE_MAX)
		hipLaunchKernelGGL(k_large_small<WRITE>, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, nLarge, all, smallMap,
						   largeMap, counts, offsets, pairs, maxPairs, sc.sceneOf, normalise);
	else
		hipLaunchKernelGGL(k_large_small_sorted<WRITE>, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, nLarge, all,
						   smallMap, largeMap, (const unsigned*)sc.lvals, (const float*)sc.lminx,
						   (const unsigned*)sc.lmaxExt, counts, offsets, pairs, maxPairs, sc.sceneOf, normalise);
}

void b3h_grid_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
					const b3hGridParams& gp, int numCells, b3hInt4* pairs, int maxPairs, int* pairCount,
					int* pairTotal, b3hBroadphaseScratch& sc, hipStream_t s)
{
	int* countsLS = sc.counts;
	int* countsG = sc.counts + nSmall;
	if (nSmall > 0)
	{
		large_small_prepare(all, largeMap, nLarge, sc, s);
		large_small_pass<false>(nSmall, nLarge, all, smallMap, largeMap, sc, countsLS, (const int*)nullptr,
								(b3hInt4*)nullptr, maxPairs, s);
		hipLaunchKernelGGL(k_grid_hash, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, all, smallMap, gp, sc.keys, sc.vals);
		int bits = 0;
		while ((1 << bits) < numCells) bits++;
		b3h_radix_sort(sc.keys, sc.vals, sc.tmpKeys, sc.tmpVals, sc.hist, nSmall, bits, s);
		b3h_fill_i32(sc.cellStart, -1, numCells, s);
		hipLaunchKernelGGL(k_cell_start, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, sc.keys, sc.cellStart);
		hipLaunchKernelGGL(k_grid_gather, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, sc.vals, all, smallMap,
						   sc.sortedAabbs);
		hipLaunchKernelGGL(k_grid_pairs, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, sc.keys,
						   (const b3hAabb*)sc.sortedAabbs, sc.cellStart, gp, countsG, sc.stash, sc.sceneOf);
	}
	int nTot = 2 * nSmall;
	b3h_exclusive_scan(sc.counts, sc.offsets, nullptr, nTot, sc.scanTmp, s);
	if (nSmall > 0)
	{
		large_small_pass<true>(nSmall, nLarge, all, smallMap, largeMap, sc, (int*)nullptr, sc.offsets, pairs,
							   maxPairs, s);
		hipLaunchKernelGGL(k_grid_emit, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, sc.keys,
						   (const b3hAabb*)sc.sortedAabbs, sc.cellStart, gp, (const int*)countsG,
						   (const int*)sc.offsets + nSmall, (const int2*)sc.stash, pairs, maxPairs, sc.sceneOf);
	}
	hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, sc.offsets + nTot, maxPairs, pairCount, pairTotal);
}

/* ---------------- SAP ---------------- */
#define SAP_BLOCK 256
int b3h_sap_num_partials(int nSmall) { return (nSmall + SAP_BLOCK - 1) / SAP_BLOCK; }

__global__ __launch_bounds__(SAP_BLOCK) void k_sap_partials(int n, const b3hAabb* __restrict__ all,
															const int* __restrict__ smallMap, double* __restrict__ partials)
{
	__shared__ double red[6][SAP_BLOCK];
	int t = threadIdx.x;
	int i = blockIdx.x * SAP_BLOCK + t;
	double v[6] = {0, 0, 0, 0, 0, 0};
	if (i < n)
	{
		b3hAabb a = all[smallMap[i]];
		float c[3] = {(a.m_max.x + a.m_min.x) * 0.5f, (a.m_max.y + a.m_min.y) * 0.5f, (a.m_max.z + a.m_min.z) * 0.5f};
		for (int k = 0; k < 3; k++)
		{
			v[k] = (double)c[k];
			v[3 + k] = (double)(c[k] * c[k]);
		}
	}
	for (int k = 0; k < 6; k++) red[k][t] = v[k];
	__syncthreads();
	for (int off = SAP_BLOCK / 2; off > 0; off >>= 1)
	{
		if (t < off)
			for (int k = 0; k < 6; k++) red[k][t] += red[k][t + off];
		__syncthreads();
	}
	if (t < 6) partials[blockIdx.x * 6 + t] = red[t][0];
}

__global__ void k_sap_axis(int numPartials, const double* __restrict__ partials, int n, int* __restrict__ axisOut)
{
	double s[6] = {0, 0, 0, 0, 0, 0};
	for (int b = 0; b < numPartials; b++)
		for (int k = 0; k < 6; k++) s[k] += partials[b * 6 + k];
	int axis = 0;
	if (n > 0)
	{
		double v[3];
		for (int k = 0; k < 3; k++) v[k] = s[3 + k] - (s[k] * s[k]) / (double)n;
		if (v[1] > v[0]) axis = 1;
		if (v[2] > v[axis]) axis = 2;
	}
	*axisOut = axis;
}

__device__ __forceinline__ unsigned float_flip(float f)
{
	unsigned u = f2u(f);
	unsigned mask = (unsigned)(-(int)(u >> 31)) | 0x80000000u;
	return u ^ mask;
}
__device__ __forceinline__ float axis_val(F4 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

__global__ __launch_bounds__(256) void k_sap_keys(int n, const b3hAabb* __restrict__ all, const int* __restrict__ smallMap,
												  const int* __restrict__ axisPtr, unsigned* __restrict__ keys,
												  unsigned* __restrict__ vals)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int axis = *axisPtr;
	keys[i] = float_flip(axis_val(all[smallMap[i]].m_min, axis));
	vals[i] = (unsigned)i;
}

__global__ __launch_bounds__(256) void k_sap_scatter(int n, const b3hAabb* __restrict__ all, const int* __restrict__ smallMap,
													 const unsigned* __restrict__ vals, b3hAabb* __restrict__ sorted)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	sorted[i] = all[smallMap[vals[i]]];
}

template <bool WRITE>
__global__ __launch_bounds__(256) void k_sap_sweep(int n, const b3hAabb* __restrict__ sorted, const int* __restrict__ axisPtr,
												   int* __restrict__ counts, const int* __restrict__ offsets,
												   b3hInt4* __restrict__ pairs, int maxPairs,
												   const int* __restrict__ sceneOf)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int axis = *axisPtr;
	b3hAabb a = sorted[i];
	float amax = axis_val(a.m_max, axis);
	int ha = bp_handle(a);
	int c = 0;
	int o = WRITE ? offsets[i] : 0;
	for (int j = i + 1; j < n; j++)
	{
		b3hAabb b = sorted[j];
		if (amax < axis_val(b.m_min, axis)) break;
		if (bp_overlap(a, b) && bp_same_scene(sceneOf, ha, bp_handle(b)))
		{
			if (WRITE)
			{
				int at = o + c;
				if (at < maxPairs) pairs[at] = b3hInt4{ha, bp_handle(b), -1, -1};
			}
			c++;
		}
	}
	if (!WRITE) counts[i] = c;
}

void b3h_sap_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
				   b3hInt4* pairs, int maxPairs, int* pairCount, int* pairTotal, b3hBroadphaseScratch& sc, hipStream_t s)
{
	int* countsLS = sc.counts;
	int* countsS = sc.counts + nSmall;
	if (nSmall > 0)
	{
		int np = b3h_sap_num_partials(nSmall);
		hipLaunchKernelGGL(k_sap_partials, dim3(np), dim3(SAP_BLOCK), 0, s, nSmall, all, smallMap, sc.sapPartials);
		hipLaunchKernelGGL(k_sap_axis, dim3(1), dim3(1), 0, s, np, sc.sapPartials, nSmall, sc.axis);
		large_small_prepare(all, largeMap, nLarge, sc, s);
		large_small_pass<false>(nSmall, nLarge, all, smallMap, largeMap, sc, countsLS, (const int*)nullptr,
								(b3hInt4*)nullptr, maxPairs, s);
		hipLaunchKernelGGL(k_sap_keys, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, all, smallMap, sc.axis, sc.keys, sc.vals);
		b3h_radix_sort(sc.keys, sc.vals, sc.tmpKeys, sc.tmpVals, sc.hist, nSmall, 32, s);
		hipLaunchKernelGGL(k_sap_scatter, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, all, smallMap, sc.vals, sc.sortedAabbs);
		hipLaunchKernelGGL(k_sap_sweep<false>, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, sc.sortedAabbs, sc.axis, countsS,
						   (const int*)nullptr, (b3hInt4*)nullptr, maxPairs, sc.sceneOf);
	}
	int nTot = 2 * nSmall;
	b3h_exclusive_scan(sc.counts, sc.offsets, nullptr, nTot, sc.scanTmp, s);
	if (nSmall > 0)
	{
		large_small_pass<true>(nSmall, nLarge, all, smallMap, largeMap, sc, (int*)nullptr, sc.offsets, pairs,
							   maxPairs, s);
		hipLaunchKernelGGL(k_sap_sweep<true>, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, sc.sortedAabbs, sc.axis,
						   (int*)nullptr, sc.offsets + nSmall, pairs, maxPairs, sc.sceneOf);
	}
	hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, sc.offsets + nTot, maxPairs, pairCount, pairTotal);
}

/* contact manifolds per scene of the last step: pairs with a contact slot,
 * counted under the scene of their first body */
__global__ __launch_bounds__(256) void k_scene_counts(const int* __restrict__ nPairs, const b3hInt4* __restrict__ pairs,
													  const int* __restrict__ sceneOf, int* __restrict__ counts)
{
	int n = *nPairs;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		b3hInt4 pr = pairs[i];
		if (pr.z >= 0) atomicAdd(&counts[sceneOf[pr.x]], 1);
	}
}
void b3h_scene_contact_counts(const int* nPairs, int maxPairs, const b3hInt4* pairs, const int* sceneOf, int* counts,
							  int numScenes, hipStream_t s)
{
	b3h_fill_i32(counts, 0, numScenes, s);
	int b = (maxPairs + 255) / 256;
	if (b > 4096) b = 4096;
	if (b < 1) b = 1;
	hipLaunchKernelGGL(k_scene_counts, dim3(b), dim3(256), 0, s, nPairs, pairs, sceneOf, counts);
}

/* ---------------- parallel linear BVH ----------------
 * b3GpuParallelLinearBvh (src/Bullet3OpenCL/BroadphaseCollision/
 * b3GpuParallelLinearBvh.cpp:120-330, kernels/parallelLinearBvh.cl), rebuilt
 * every call over the small proxies:
 *   merged AABB of all small proxies (min/max reduction: exact, so any order
 *   gives the reference's box) -> Morton code of each AABB centre on a 1024^3
 *   grid over that box -> stable radix sort of (code, small index) -> binary
 *   radix tree over the 64-bit keys (code << 32 | sorted position): internal
 *   node i splits leaves i and i+1, and a node's parent is whichever split at
 *   its two range ends has the longer common prefix -- the same tree the
 *   reference finds with binary searches, built here bottom-up in ONE pass
 *   (each leaf walks up; of the two children of a node the second to arrive
 *   merges the AABBs and continues), instead of the reference's per-depth
 *   launches after a host read of the depth.
 * Pairs: each leaf, in Morton order, walks the tree with a stack in LDS and
 * pairs with every later leaf whose AABB overlaps (pairs (query, later) as in
 * plbvhCalculateOverlappingPairs, parallelLinearBvh.cl:185-260); then every
 * small proxy against the large ones, as (large, small)
 * (plbvhLargeAabbAabbTest, :395-420).  Compaction is the grid's: a counting
 * pass that stashes the first pairs of each leaf, a prefix sum, an emitting
 * pass -- so the list has a fixed order: leaves in Morton order, each leaf's
 * partners in traversal order (right child before left), then the large x
 * small pairs small-major.  The reference appends with an atomic counter; its
 * order is whatever the hardware schedules, its pair set is this one. */
#define PLBVH_TPB 128
#define PLBVH_STACK 64
#define PLBVH_REDUCE_BLOCKS 1024
#define PLBVH_INTERNAL 0x80000000u

__global__ __launch_bounds__(256) void k_plbvh_merge_partial(int n, const b3hAabb* __restrict__ all,
															 const int* __restrict__ smallMap,
															 b3hAabb* __restrict__ partials)
{
	__shared__ float red[6][256];
	const int t = threadIdx.x;
	float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
	for (int i = blockIdx.x * 256 + t; i < n; i += gridDim.x * 256)
	{
		b3hAabb a = all[smallMap[i]];
		v[0] = fminf(v[0], a.m_min.x);
		v[1] = fminf(v[1], a.m_min.y);
		v[2] = fminf(v[2], a.m_min.z);
		v[3] = fmaxf(v[3], a.m_max.x);
		v[4] = fmaxf(v[4], a.m_max.y);
		v[5] = fmaxf(v[5], a.m_max.z);
	}
	for (int k = 0; k < 6; k++) red[k][t] = v[k];
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1)
	{
		if (t < off)
		{
			for (int k = 0; k < 3; k++) red[k][t] = fminf(red[k][t], red[k][t + off]);
			for (int k = 3; k < 6; k++) red[k][t] = fmaxf(red[k][t], red[k][t + off]);
		}
		__syncthreads();
	}
	if (t == 0)
	{
		b3hAabb o;
		o.m_min = f4(red[0][0], red[1][0], red[2][0], 0.f);
		o.m_max = f4(red[3][0], red[4][0], red[5][0], 0.f);
		partials[blockIdx.x] = o;
	}
}

/* interleaveBits / getMortonCode, parallelLinearBvh.cl:43-77 */
__device__ __forceinline__ unsigned plbvh_spread(unsigned x)
{
	x &= 0x000003FFu;
	x = (x ^ (x << 16)) & 0xFF0000FFu;
	x = (x ^ (x << 8)) & 0x0300F00Fu;
	x = (x ^ (x << 4)) & 0x030C30C3u;
	x = (x ^ (x << 2)) & 0x09249249u;
	return x;
}
/* grid coordinate of assignMortonCodesAndAabbIndicies (:105-140): floor
 * towards -inf, clamped to [-512, 511], shifted to [0, 1023].  A coordinate
 * of 0/0 (every centre on one plane: the merged box is flat on that axis) is
 * taken as 0 -- what the conversion gives on AMD hardware; C leaves it open. */
__device__ __forceinline__ unsigned plbvh_coord(float rel, float cell)
{
	float p = rel / cell;
	int d;
	if (p != p)
		d = 0;
	else
	{
		p = fminf(fmaxf(p, -513.f), 512.f);
		d = (p >= 0.f) ? (int)p : (int)floorf(p);
	}
	d = d < -512 ? -512 : (d > 511 ? 511 : d);
	return (unsigned)(d + 512);
}

/* one block: reduce the partial boxes to the merged box, then every small
 * proxy's Morton code; merged[0] keeps the box */
__global__ __launch_bounds__(256) void k_plbvh_merge_final(int nPartials, b3hAabb* __restrict__ partials)
{
	__shared__ float red[6][256];
	const int t = threadIdx.x;
	float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
	for (int i = t; i < nPartials; i += 256)
	{
		b3hAabb a = partials[i];
		v[0] = fminf(v[0], a.m_min.x);
		v[1] = fminf(v[1], a.m_min.y);
		v[2] = fminf(v[2], a.m_min.z);
		v[3] = fmaxf(v[3], a.m_max.x);
		v[4] = fmaxf(v[4], a.m_max.y);
		v[5] = fmaxf(v[5], a.m_max.z);
	}
	for (int k = 0; k < 6; k++) red[k][t] = v[k];
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1)
	{
		if (t < off)
		{
			for (int k = 0; k < 3; k++) red[k][t] = fminf(red[k][t], red[k][t + off]);
			for (int k = 3; k < 6; k++) red[k][t] = fmaxf(red[k][t], red[k][t + off]);
		}
		__syncthreads();
	}
	if (t == 0)
	{
		b3hAabb o;
		o.m_min = f4(red[0][0], red[1][0], red[2][0], 0.f);
		o.m_max = f4(red[3][0], red[4][0], red[5][0], 0.f);
		partials[nPartials] = o;
	}
}

__global__ __launch_bounds__(256) void k_plbvh_morton(int n, const b3hAabb* __restrict__ all,
													  const int* __restrict__ smallMap,
													  const b3hAabb* __restrict__ merged, unsigned* __restrict__ keys,
													  unsigned* __restrict__ vals)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const b3hAabb m = *merged;
	b3hAabb a = all[smallMap[i]];
	float gx = (m.m_min.x + m.m_max.x) * 0.5f, gy = (m.m_min.y + m.m_max.y) * 0.5f,
		  gz = (m.m_min.z + m.m_max.z) * 0.5f;
	float cx = (m.m_max.x - m.m_min.x) / 1024.f, cy = (m.m_max.y - m.m_min.y) / 1024.f,
		  cz = (m.m_max.z - m.m_min.z) / 1024.f;
	float px = (a.m_min.x + a.m_max.x) * 0.5f, py = (a.m_min.y + a.m_max.y) * 0.5f,
		  pz = (a.m_min.z + a.m_max.z) * 0.5f;
	unsigned ux = plbvh_coord(px - gx, cx), uy = plbvh_coord(py - gy, cy), uz = plbvh_coord(pz - gz, cz);
	keys[i] = plbvh_spread(ux) | (plbvh_spread(uy) << 1) | (plbvh_spread(uz) << 2);
	vals[i] = (unsigned)i;
}

/* common prefix length of the keys of sorted leaves i and i+1, each key made
 * unique by its position (computeAdjacentPairCommonPrefix, :440-470) */
__device__ __forceinline__ int plbvh_delta(const unsigned* __restrict__ keys, int i)
{
	unsigned long long a = ((unsigned long long)keys[i] << 32) | (unsigned)i;
	unsigned long long b = ((unsigned long long)keys[i + 1] << 32) | (unsigned)(i + 1);
	return __clzll((long long)(a ^ b));
}

/* Bottom-up build.  Internal node p: child[p] = (left, right) with
 * PLBVH_INTERNAL set on internal children; box[2p] = (min.xyz, highest leaf),
 * box[2p+1] = (max.xyz, lowest leaf).  flag[] must be zero.  meta[0] = root. */
__global__ __launch_bounds__(256) void k_plbvh_build(int n, const unsigned* __restrict__ keys,
													 const b3hAabb* __restrict__ leaves, int2* child,
													 b3hFloat4* box, int* flag, int* __restrict__ meta)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	if (n == 1)
	{
		meta[0] = 0;
		return;
	}
	int l = i, r = i;
	unsigned cur = (unsigned)i;
	b3hAabb a = leaves[i];
	float mn[3] = {a.m_min.x, a.m_min.y, a.m_min.z};
	float mx[3] = {a.m_max.x, a.m_max.y, a.m_max.z};
	while (true)
	{
		if (l == 0 && r == n - 1)
		{
			meta[0] = (int)cur;
			break;
		}
		/* parent: the split at the range end with the longer prefix (there
		 * are no ties among distinct keys); this node is its left child when
		 * the parent is the split at r */
		const bool isLeft = (l == 0) || (r != n - 1 && plbvh_delta(keys, r) > plbvh_delta(keys, l - 1));
		const int p = isLeft ? r : l - 1;
		int* cp = (int*)&child[p];
		int* bp = (int*)&box[2 * p];
		if (isLeft)
		{
			cp[0] = (int)cur;
			bp[7] = l; /* box[2p+1].w: lowest leaf */
		}
		else
		{
			cp[1] = (int)cur;
			bp[3] = r; /* box[2p].w: highest leaf */
		}
		__threadfence();
		if (atomicAdd(&flag[p], 1) == 0) break; /* the sibling finishes p */
		__threadfence();
		const int sib = __hip_atomic_load(&cp[isLeft ? 1 : 0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (isLeft)
			r = __hip_atomic_load(&bp[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		else
			l = __hip_atomic_load(&bp[7], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		float smn[3], smx[3];
		if (sib >= 0)
		{
			b3hAabb b = leaves[sib];
			smn[0] = b.m_min.x, smn[1] = b.m_min.y, smn[2] = b.m_min.z;
			smx[0] = b.m_max.x, smx[1] = b.m_max.y, smx[2] = b.m_max.z;
		}
		else
		{
			const int q = sib & 0x7fffffff;
			const float* sb = (const float*)&box[2 * q];
			for (int k = 0; k < 3; k++)
			{
				smn[k] = __hip_atomic_load(&sb[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				smx[k] = __hip_atomic_load(&sb[4 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			}
		}
		for (int k = 0; k < 3; k++)
		{
			mn[k] = fminf(mn[k], smn[k]);
			mx[k] = fmaxf(mx[k], smx[k]);
		}
		float* pb = (float*)&box[2 * p];
		pb[0] = mn[0], pb[1] = mn[1], pb[2] = mn[2];
		pb[4] = mx[0], pb[5] = mx[1], pb[6] = mx[2];
		cur = (unsigned)p | PLBVH_INTERNAL;
	}
}

/* one leaf's traversal: MODE 0 counts its pairs and stashes the first
 * B3H_GRID_STASH, MODE 1 writes them all at offset o.  The stack holds the
 * pending left siblings of the current path (at most one per level of the
 * tree; with 64-bit keys a path has at most 64 internal nodes). */
template <int MODE>
__device__ __forceinline__ int plbvh_query(int q, int n, const b3hAabb* __restrict__ leaves,
										   const int2* __restrict__ child, const b3hFloat4* __restrict__ box,
										   int root, int* stk, int2* __restrict__ stash, int o,
										   b3hInt4* __restrict__ pairs, int maxPairs, const int* __restrict__ sceneOf,
										   int* __restrict__ overflow)
{
	if (q >= n - 1) return 0; /* no later leaf */
	const b3hAabb a = leaves[q];
	const int ha = bp_handle(a);
	int c = 0, sp = 0;
	int v = root & 0x7fffffff;
	while (true)
	{
		const int2 ch = child[v];
		int next = -1;
#pragma unroll
		for (int k = 0; k < 2; k++)
		{
			const int cc = k == 0 ? ch.y : ch.x;
			if (cc >= 0)
			{
				if (cc > q)
				{
					const b3hAabb b = leaves[cc];
					if (bp_overlap(a, b))
					{
						const int hb = bp_handle(b);
						if (bp_same_scene(sceneOf, ha, hb))
						{
							if (MODE == 0)
							{
								if (c < B3H_GRID_STASH) stash[(size_t)q * B3H_GRID_STASH + c] = make_int2(ha, hb);
							}
							else
							{
								const int at = o + c;
								if (at < maxPairs) pairs[at] = b3hInt4{ha, hb, -1, -1};
							}
							c++;
						}
					}
				}
			}
			else
			{
				const int idx = cc & 0x7fffffff;
				const b3hFloat4 lo = box[2 * idx];
				if (f2i(lo.w) > q)
				{
					const b3hFloat4 hi = box[2 * idx + 1];
					if (a.m_min.x <= hi.x && lo.x <= a.m_max.x && a.m_min.y <= hi.y && lo.y <= a.m_max.y &&
						a.m_min.z <= hi.z && lo.z <= a.m_max.z)
					{
						if (next < 0)
							next = idx;
						else if (sp < PLBVH_STACK)
							stk[(sp++) * PLBVH_TPB] = idx;
						else
							*overflow = 1;
					}
				}
			}
		}
		if (next < 0)
		{
			if (sp == 0) break;
			next = stk[(--sp) * PLBVH_TPB];
		}
		v = next;
	}
	return c;
}

__global__ __launch_bounds__(PLBVH_TPB) void k_plbvh_pairs(int n, const b3hAabb* __restrict__ leaves,
														   const int2* __restrict__ child,
														   const b3hFloat4* __restrict__ box,
														   const int* __restrict__ meta, int* __restrict__ counts,
														   int2* __restrict__ stash, const int* __restrict__ sceneOf,
														   int* __restrict__ overflow)
{
	__shared__ int stk[PLBVH_STACK * PLBVH_TPB];
	int q = blockIdx.x * PLBVH_TPB + threadIdx.x;
	if (q >= n) return;
	counts[q] = plbvh_query<0>(q, n, leaves, child, box, meta[0], stk + threadIdx.x, stash, 0, nullptr, 0, sceneOf,
							   overflow);
}

__global__ __launch_bounds__(PLBVH_TPB) void k_plbvh_emit(int n, const b3hAabb* __restrict__ leaves,
														  const int2* __restrict__ child,
														  const b3hFloat4* __restrict__ box,
														  const int* __restrict__ meta, const int* __restrict__ counts,
														  const int* __restrict__ offsets,
														  const int2* __restrict__ stash, b3hInt4* __restrict__ pairs,
														  int maxPairs, const int* __restrict__ sceneOf,
														  int* __restrict__ overflow)
{
	__shared__ int stk[PLBVH_STACK * PLBVH_TPB];
	int q = blockIdx.x * PLBVH_TPB + threadIdx.x;
	if (q >= n) return;
	const int c = counts[q], o = offsets[q];
	if (c <= B3H_GRID_STASH)
	{
		for (int j = 0; j < c; j++)
		{
			int2 v = stash[(size_t)q * B3H_GRID_STASH + j];
			if (o + j < maxPairs) pairs[o + j] = b3hInt4{v.x, v.y, -1, -1};
		}
	}
	else
		plbvh_query<1>(q, n, leaves, child, box, meta[0], stk + threadIdx.x, nullptr, o, pairs, maxPairs, sceneOf,
					   overflow);
}

int b3h_plbvh_reduce_blocks(int nSmall)
{
	int b = nblk(nSmall);
	return b < 1 ? 1 : (b > PLBVH_REDUCE_BLOCKS ? PLBVH_REDUCE_BLOCKS : b);
}

/* Builds the tree over the small proxies (kept in sc for ray queries) and
 * writes the pair list.  sc.bvhMeta: [root, stack overflow flag] */
void b3h_plbvh_build(const b3hAabb* all, const int* smallMap, int nSmall, b3hBroadphaseScratch& sc, hipStream_t s)
{
	if (nSmall <= 0) return;
	const int rb = b3h_plbvh_reduce_blocks(nSmall);
	hipLaunchKernelGGL(k_plbvh_merge_partial, dim3(rb), dim3(256), 0, s, nSmall, all, smallMap, sc.bvhMerge);
	hipLaunchKernelGGL(k_plbvh_merge_final, dim3(1), dim3(256), 0, s, rb, sc.bvhMerge);
	hipLaunchKernelGGL(k_plbvh_morton, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, all, smallMap,
					   (const b3hAabb*)sc.bvhMerge + rb, sc.keys, sc.vals);
	b3h_radix_sort(sc.keys, sc.vals, sc.tmpKeys, sc.tmpVals, sc.hist, nSmall, 30, s);
	hipLaunchKernelGGL(k_grid_gather, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, sc.vals, all, smallMap,
					   sc.sortedAabbs);
	b3h_fill_i32(sc.bvhFlag, 0, nSmall > 1 ? nSmall - 1 : 1, s);
	b3h_fill_i32(sc.bvhMeta, 0, 2, s);
	hipLaunchKernelGGL(k_plbvh_build, dim3(nblk(nSmall)), dim3(256), 0, s, nSmall, (const unsigned*)sc.keys,
					   (const b3hAabb*)sc.sortedAabbs, sc.bvhChild, sc.bvhBox, sc.bvhFlag, sc.bvhMeta);
}

void b3h_plbvh_pairs(const b3hAabb* all, const int* smallMap, int nSmall, const int* largeMap, int nLarge,
					 b3hInt4* pairs, int maxPairs, int* pairCount, int* pairTotal, b3hBroadphaseScratch& sc,
					 hipStream_t s)
{
	int* countsB = sc.counts;
	int* countsLS = sc.counts + nSmall;
	if (nSmall > 0)
	{
		b3h_plbvh_build(all, smallMap, nSmall, sc, s);
		const int tb = (nSmall + PLBVH_TPB - 1) / PLBVH_TPB;
		hipLaunchKernelGGL(k_plbvh_pairs, dim3(tb), dim3(PLBVH_TPB), 0, s, nSmall, (const b3hAabb*)sc.sortedAabbs,
						   (const int2*)sc.bvhChild, (const b3hFloat4*)sc.bvhBox, (const int*)sc.bvhMeta, countsB,
						   sc.stash, sc.sceneOf, sc.bvhMeta + 1);
		large_small_prepare(all, largeMap, nLarge, sc, s);
		large_small_pass<false>(nSmall, nLarge, all, smallMap, largeMap, sc, countsLS, (const int*)nullptr,
								(b3hInt4*)nullptr, maxPairs, s, 0);
	}
	const int nTot = 2 * nSmall;
	b3h_exclusive_scan(sc.counts, sc.offsets, nullptr, nTot, sc.scanTmp, s);
	if (nSmall > 0)
	{
		const int tb = (nSmall + PLBVH_TPB - 1) / PLBVH_TPB;
		hipLaunchKernelGGL(k_plbvh_emit, dim3(tb), dim3(PLBVH_TPB), 0, s, nSmall, (const b3hAabb*)sc.sortedAabbs,
						   (const int2*)sc.bvhChild, (const b3hFloat4*)sc.bvhBox, (const int*)sc.bvhMeta,
						   (const int*)countsB, (const int*)sc.offsets, (const int2*)sc.stash, pairs, maxPairs,
						   sc.sceneOf, sc.bvhMeta + 1);
		large_small_pass<true>(nSmall, nLarge, all, smallMap, largeMap, sc, (int*)nullptr, sc.offsets + nSmall, pairs,
							   maxPairs, s, 0);
	}
	hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, sc.offsets + nTot, maxPairs, pairCount, pairTotal);
}
