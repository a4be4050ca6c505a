/*
 * k_prims.hip -- parallel primitives on gfx950: stable LSD radix sort of
 * (uint32 key, uint32 value) pairs with 8-bit digits, and a device-wide
 * exclusive prefix sum.
 *
 * Radix pass (replaces b3RadixSort32CL, src/Bullet3OpenCL/ParallelPrimitives/
 * b3RadixSort32CL.cpp:174-518, 4-bit digits there):
 *   hist    one block per tile of TILE keys counts its digits (LDS atomics)
 *   scan    digit-major exclusive scan of the [256][tiles] counts
 *   scatter each block walks its tile in order, 256 keys at a time; a key's
 *           rank among equal digits of its wavefront comes from 8 ballots
 *           (lanes with the same digit) and a popcount below the lane, waves
 *           are prefixed through LDS -> stable, deterministic placement.
 * Only as many passes as the key width needs are run.
 */
#include <mutex>
#include "b3h_internal.h"

#define RS_THREADS 256
#define RS_ITEMS 8
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_BUCKETS 256

__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const unsigned* __restrict__ keys, int n, int shift,
														   unsigned* __restrict__ hist, int numTiles)
{
	__shared__ unsigned cnt[RS_BUCKETS];
	int t = threadIdx.x;
	cnt[t] = 0;
	__syncthreads();
	int base = blockIdx.x * RS_TILE;
	for (int k = 0; k < RS_ITEMS; k++)
	{
		int i = base + k * RS_THREADS + t;
		if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xff], 1u);
	}
	__syncthreads();
	hist[t * numTiles + blockIdx.x] = cnt[t];
}

/* exclusive scan of a small array (<= a few hundred thousand) by one block:
 * rounds of 8192 consecutive entries, 8 consecutive per thread (a wavefront
 * reads 2 KB of contiguous memory per round), a block scan of the thread sums
 * and a running carry between rounds */
__global__ __launch_bounds__(1024) void k_scan_single(unsigned* __restrict__ data, int n)
{
	__shared__ unsigned wsum[16];
	const int t = threadIdx.x, lane = t & 63, w = t >> 6;
	unsigned carry = 0;
	for (int base = 0; base < n; base += 8 * 1024)
	{
		const int i0 = base + 8 * t;
		unsigned v[8];
#pragma unroll
		for (int u = 0; u < 8; u++) v[u] = (i0 + u < n) ? data[i0 + u] : 0u;
		unsigned sum = 0;
#pragma unroll
		for (int u = 0; u < 8; u++) sum += v[u];
		/* inclusive scan of the thread sums: within the wavefront, then over
		 * the 16 wavefront totals */
		unsigned x = sum;
#pragma unroll
		for (int o = 1; o < 64; o <<= 1)
		{
			unsigned y = __shfl_up(x, o);
			if (lane >= o) x += y;
		}
		if (lane == 63) wsum[w] = x;
		__syncthreads();
		if (w == 0)
		{
			unsigned y0 = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
			for (int o = 1; o < 16; o <<= 1)
			{
				unsigned y = __shfl_up(y0, o);
				if (lane >= o) y0 += y;
			}
			if (lane < 16) wsum[lane] = y0;
		}
		__syncthreads();
		unsigned run = carry + (w > 0 ? wsum[w - 1] : 0u) + x - sum;
		const unsigned total = wsum[15];
#pragma unroll
		for (int u = 0; u < 8; u++)
		{
			if (i0 + u < n) data[i0 + u] = run;
			run += v[u];
		}
		carry += total;
		__syncthreads();
	}
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(const unsigned* __restrict__ keysIn,
															  const unsigned* __restrict__ valsIn, int n, int shift,
															  const unsigned* __restrict__ histScanned, int numTiles,
															  unsigned* __restrict__ keysOut,
															  unsigned* __restrict__ valsOut)
{
	__shared__ unsigned base[RS_BUCKETS];
	__shared__ unsigned waveCnt[RS_THREADS / 64][RS_BUCKETS];
	int t = threadIdx.x;
	int lane = t & 63;
	int wave = t >> 6;
	base[t] = histScanned[t * numTiles + blockIdx.x];
	unsigned long long lanesBelow = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
	int tileStart = blockIdx.x * RS_TILE;
	for (int k = 0; k < RS_ITEMS; k++)
	{
		for (int w = 0; w < RS_THREADS / 64; w++) waveCnt[w][t] = 0;
		__syncthreads();
		int i = tileStart + k * RS_THREADS + t;
		bool valid = i < n;
		unsigned key = valid ? keysIn[i] : 0u;
		unsigned val = valid ? valsIn[i] : 0u;
		unsigned d = (key >> shift) & 0xff;
		unsigned long long m = __ballot(valid);
#pragma unroll
		for (int b = 0; b < 8; b++)
		{
			unsigned long long bb = __ballot(valid && ((d >> b) & 1));
			m &= ((d >> b) & 1) ? bb : ~bb;
		}
		int rank = __popcll(m & lanesBelow);
		int total = __popcll(m);
		/* the highest lane of each digit group publishes the group size */
		if (valid && rank == total - 1) waveCnt[wave][d] = (unsigned)total;
		__syncthreads();
		if (valid)
		{
			unsigned off = base[d];
			for (int w = 0; w < wave; w++) off += waveCnt[w][d];
			unsigned pos = off + (unsigned)rank;
			keysOut[pos] = key;
			valsOut[pos] = val;
		}
		__syncthreads();
		unsigned add = 0;
		for (int w = 0; w < RS_THREADS / 64; w++) add += waveCnt[w][t];
		base[t] += add;
		__syncthreads();
	}
}

/* Sorts (keys,vals) of length n by the low `bits` bits of the key, stably.
 * keys/vals receive the result; tmpKeys/tmpVals are scratch of length n,
 * hist is scratch of 256 * ceil(n / RS_TILE) entries. */
void b3h_radix_sort(unsigned* keys, unsigned* vals, unsigned* tmpKeys, unsigned* tmpVals, unsigned* hist, int n, int bits,
					hipStream_t s)
{
	if (n <= 0) return;
	int tiles = (n + RS_TILE - 1) / RS_TILE;
	int passes = (bits + 7) / 8;
	unsigned *ki = keys, *vi = vals, *ko = tmpKeys, *vo = tmpVals;
	for (int p = 0; p < passes; p++)
	{
		int shift = 8 * p;
		hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(RS_THREADS), 0, s, ki, n, shift, hist, tiles);
		hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, hist, RS_BUCKETS * tiles);
		hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(RS_THREADS), 0, s, ki, vi, n, shift, hist, tiles, ko, vo);
		unsigned* t;
		t = ki; ki = ko; ko = t;
		t = vi; vi = vo; vo = t;
	}
	if (ki != keys)
	{
		hipMemcpyAsync(keys, ki, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, s);
		hipMemcpyAsync(vals, vi, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, s);
	}
}

size_t b3h_radix_hist_size(int n) { return (size_t)RS_BUCKETS * ((n + RS_TILE - 1) / RS_TILE + 1); }

/* ---- device-wide exclusive scan of int counts (n read from the host) ----
 * out[i] = sum_{j<i} in[j] for i <= n (out has n+1 entries). */
#define SC_THREADS 256
#define SC_ITEMS 16
#define SC_TILE (SC_THREADS * SC_ITEMS)

__global__ __launch_bounds__(SC_THREADS) void k_scan_tiles(const int* __restrict__ in, const int* __restrict__ nPtr,
														   int nFixed, int* __restrict__ out, unsigned* __restrict__ tileSums)
{
	__shared__ int part[SC_THREADS];
	int n = nPtr ? *nPtr : nFixed;
	int t = threadIdx.x;
	int base = blockIdx.x * SC_TILE + t * SC_ITEMS;
	int v[SC_ITEMS];
	int sum = 0;
#pragma unroll
	for (int k = 0; k < SC_ITEMS; k++)
	{
		int i = base + k;
		v[k] = i < n ? in[i] : 0;
		sum += v[k];
	}
	part[t] = sum;
	__syncthreads();
	for (int off = 1; off < SC_THREADS; off <<= 1)
	{
		int x = t >= off ? part[t - off] : 0;
		__syncthreads();
		part[t] += x;
		__syncthreads();
	}
	int run = part[t] - sum;
#pragma unroll
	for (int k = 0; k < SC_ITEMS; k++)
	{
		int i = base + k;
		if (i < n) out[i] = run;
		run += v[k];
	}
	if (t == SC_THREADS - 1) tileSums[blockIdx.x] = (unsigned)part[t];
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_add(int* __restrict__ out, const int* __restrict__ nPtr, int nFixed,
														 const unsigned* __restrict__ tileSumsScanned, int numTiles)
{
	int n = nPtr ? *nPtr : nFixed;
	int t = threadIdx.x;
	int add = (int)tileSumsScanned[blockIdx.x];
	int base = blockIdx.x * SC_TILE + t * SC_ITEMS;
	for (int k = 0; k < SC_ITEMS; k++)
	{
		int i = base + k;
		if (i < n) out[i] += add;
	}
	if (blockIdx.x == 0 && t == 0) out[n] = (int)tileSumsScanned[numTiles];
}

/* tileSums needs maxN/SC_TILE + 2 entries.  The grid is sized from maxN so the
 * call can be captured in a graph while n itself lives on the device. */
void b3h_exclusive_scan(const int* in, int* out, const int* nDev, int maxN, unsigned* tileSums, hipStream_t s)
{
	int tiles = (maxN + SC_TILE - 1) / SC_TILE;
	if (tiles < 1) tiles = 1;
	hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SC_THREADS), 0, s, in, nDev, maxN, out, tileSums);
	hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, tileSums, tiles + 1);
	hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SC_THREADS), 0, s, out, nDev, maxN, tileSums, tiles);
}
size_t b3h_scan_tmp_size(int maxN) { return (size_t)(maxN + SC_TILE - 1) / SC_TILE + 2; }

int b3h_cu_count()
{
	static int cached[64] = {0};
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
	if (cached[dev] <= 0)
	{
		int n = 0;
		if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
		cached[dev] = n;
	}
	return cached[dev];
}

/* fill helpers (used instead of memset nodes so steps can be graph-captured) */
__global__ void k_fill_i32(int* p, int v, int n)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}
void b3h_fill_i32(int* p, int v, int n, hipStream_t s)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
}

/* ---- allocation-free stable sort ----
 * The scratch of b3h_radix_sort32 (ping-pong keys and values and the digit
 * histograms) lives in one file-scope device array instead of hipMalloc'd
 * buffers, for callers that cannot allocate.  Its device address is fetched
 * once per device with hipGetSymbolAddress (the array's name is only a host
 * stand-in) and kept in noalloc_scratch_ptr[device]. */
#define NOALLOC_SORT_HIST (RS_BUCKETS * (B3H_NOALLOC_SORT_MAX / RS_TILE + 1))
#define NOALLOC_SORT_WORDS (2 * B3H_NOALLOC_SORT_MAX + NOALLOC_SORT_HIST)
static __device__ __align__(256) unsigned scratch[NOALLOC_SORT_WORDS];

#define NOALLOC_MAX_DEVICES 64
static unsigned* noalloc_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex noalloc_mutex;

static int noalloc_scratch(unsigned** out)
{
	int dev = -1;
	B3H_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
	{
		b3h_set_error("device %d beyond the %d devices the allocation-free sort supports", dev, NOALLOC_MAX_DEVICES);
		return B3H_ERR_INVALID;
	}
	std::lock_guard<std::mutex> lock(noalloc_mutex);
	if (!noalloc_scratch_ptr[dev])
	{
		void* p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
		{
			b3h_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
			return B3H_ERR_HIP;
		}
		noalloc_scratch_ptr[dev] = (unsigned*)p;
	}
	*out = noalloc_scratch_ptr[dev];
	return B3H_OK;
}

extern "C" int b3h_radix_sort32_noalloc(unsigned* keys, unsigned* vals, int n, int bits)
{
	if (n < 0 || bits < 0 || bits > 32)
	{
		b3h_set_error("invalid argument: n %d, sortBits %d", n, bits);
		return B3H_ERR_INVALID;
	}
	if (n > B3H_NOALLOC_SORT_MAX)
	{
		b3h_set_error("b3h_radix_sort32_noalloc: %d entries exceed its scratch of %d entries", n, B3H_NOALLOC_SORT_MAX);
		return B3H_ERR_CAPACITY;
	}
	if (n == 0) return B3H_OK;
	if (!keys || !vals)
	{
		b3h_set_error("invalid argument: keys and values must be device pointers");
		return B3H_ERR_INVALID;
	}
	unsigned* base = nullptr;
	int r = noalloc_scratch(&base);
	if (r) return r;
	/* every scratch word the sort reads is written earlier in the same call
	 * (histograms by k_radix_hist, ping-pong arrays by k_radix_scatter), so
	 * the previous call's contents need no clearing */
	b3h_radix_sort(keys, vals, base, base + B3H_NOALLOC_SORT_MAX, base + 2 * B3H_NOALLOC_SORT_MAX, n, bits, 0);
	B3H_CHECK(hipGetLastError());
	B3H_CHECK(hipDeviceSynchronize());
	return B3H_OK;
}
