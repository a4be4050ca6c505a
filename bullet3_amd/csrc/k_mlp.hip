/*
 * k_mlp.hip -- the learned warm start's network on gfx950 matrix cores: a
 * 3-layer MLP (in -> 128 -> 128 -> out, GELU, bf16 weights and activations,
 * fp32 accumulation) with its own forward, backward and Adam kernels.
 *
 * Matrix products use v_mfma_f32_32x32x16_bf16.  Operand layout of that
 * instruction for one wave (lane l, r = l % 32, h = l / 32):
 *   A (32 x 16): lane holds A[r][8h .. 8h+7]
 *   B (16 x 32): lane holds B[8h .. 8h+7][r]
 *   D (32 x 32): lane holds D[8(i/4) + 4h + (i%4)][r] in accumulator i = 0..15
 * Both operands are read with K contiguous (A row-major, B as its transpose
 * row-major), so every fragment is one 16-byte load.
 *
 * Rounding points (the CPU oracle follows the same ones): inputs, weights and
 * the two hidden activations are rounded to bf16 (nearest even); products
 * accumulate in fp32; biases are added in fp32 before GELU.
 */
#include "b3h_internal.h"
#include "b3h_mlp.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

__device__ __forceinline__ unsigned short f2bf(float f)
{
	/* round to nearest even (inputs are finite) */
	unsigned u = __builtin_bit_cast(unsigned, f);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (unsigned short)(u >> 16);
}

/* GELU, tanh form: 0.5 x (1 + tanh(u)), u = c (x + 0.044715 x^3), c = sqrt(2/pi),
 * evaluated through 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u)) with the hardware
 * exponential and a fast reciprocal (a few VALU operations instead of a full
 * tanh).  Derivative: s + 2 x s (1 - s) c (1 + 3 * 0.044715 x^2), s = that
 * sigmoid. */
#define GELU_C 0.7978845608028654f
#define GELU_A 0.044715f
__device__ __forceinline__ float gelu_sig(float x)
{
	float u = GELU_C * (x + GELU_A * x * x * x);
	return __frcp_rn(1.f + __expf(-2.f * u));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sig(x); }
__device__ __forceinline__ float gelu_grad(float x)
{
	float s = gelu_sig(x);
	return s + 2.f * x * s * (1.f - s) * GELU_C * (1.f + 3.f * GELU_A * x * x);
}

__device__ __forceinline__ void wave_lds_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* ---------------- generic product (backward pass) ----------------
 * wave (tile, split) accumulates its K-slice of one 32x32 output tile */
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, int kChunk, int splits,
												   const unsigned short* __restrict__ A, int lda,
												   const unsigned short* __restrict__ BT, int ldb,
												   float* __restrict__ P)
{
	int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
	int tilesN = N / 32, tiles = (M / 32) * tilesN;
	if (wave >= tiles * splits) return;
	int split = wave / tiles, tile = wave - split * tiles;
	int tm = tile / tilesN, tn = tile - tm * tilesN;
	int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
	floatx16 acc = {};
	int k0 = split * kChunk, k1 = k0 + kChunk < K ? k0 + kChunk : K;
	const unsigned short* ap = A + (size_t)(tm * 32 + r) * lda + 8 * h;
	const unsigned short* bp = BT + (size_t)(tn * 32 + r) * ldb + 8 * h;
	for (int k = k0; k < k1; k += 16)
	{
		bf16x8 a = *(const bf16x8*)(ap + k);
		bf16x8 b = *(const bf16x8*)(bp + k);
		acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
	}
	float* out = P + (size_t)split * M * N;
#pragma unroll
	for (int i = 0; i < 16; i++)
	{
		int row = 8 * (i >> 2) + 4 * h + (i & 3);
		out[(size_t)(tm * 32 + row) * N + tn * 32 + r] = acc[i];
	}
}
__global__ __launch_bounds__(256) void k_gemm_reduce(int MN, int splits, const float* __restrict__ P,
													 float* __restrict__ C)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= MN) return;
	float s = P[i];
	for (int k = 1; k < splits; k++) s += P[(size_t)k * MN + i];
	C[i] = s;
}

/* ---------------- fused forward ----------------
 * Each block stages the bf16 weight operands in LDS (rows padded by 8 so the
 * 16-byte fragment reads of the 32 lanes of a half-wave spread over the
 * banks); each wave then takes 32-row tiles.  Activations pass through a
 * per-wave LDS tile to become the next layer's A operand.  With `save`, the
 * bf16 input, pre-activations and activations of every row < Mp are kept for
 * the backward pass (rows >= M hold the network's response to zero input). */
#define LDS_K1 (MLP_K1 + 8)
#define LDS_H (MLP_H + 8)
struct MlpOps
{
	const unsigned short* w1t;
	const unsigned short* w2t;
	const unsigned short* w3t;
	const float* b1;
	const float* b2;
	const float* b3;
	const float* inShift; /* [K1] input normalisation: x' = (x - shift) * scale */
	const float* inScale;
};
#define FWD_WAVES 8
#define FWD_THREADS (64 * FWD_WAVES)
__global__ __launch_bounds__(FWD_THREADS) void k_mlp_forward(int M, int Mp, const int* __restrict__ mDev, int mMul, int inDim,
													 int outDim, const float* __restrict__ X, int ldx, MlpOps net,
													 float* __restrict__ Y, int ldy, int save,
													 unsigned short* __restrict__ Xb, float* __restrict__ Z1,
													 unsigned short* __restrict__ H1, float* __restrict__ Z2,
													 unsigned short* __restrict__ H2)
{
	if (mDev)
	{
		/* row count from the device (mMul rows per counted item) */
		M = *mDev * mMul;
		Mp = (M + 31) & ~31;
	}
	__shared__ __attribute__((aligned(16))) unsigned short sW1[MLP_H * LDS_K1];
	__shared__ __attribute__((aligned(16))) unsigned short sW2[MLP_H * LDS_H];
	__shared__ __attribute__((aligned(16))) unsigned short sW3[MLP_K3 * LDS_H];
	__shared__ __attribute__((aligned(16))) unsigned short sAct[FWD_WAVES * 32 * LDS_H];
	int t = threadIdx.x;
	for (int q = t; q < MLP_H * (MLP_K1 / 8); q += FWD_THREADS)
	{
		int row = q / (MLP_K1 / 8), c8 = q - row * (MLP_K1 / 8);
		*(uint4*)(sW1 + row * LDS_K1 + 8 * c8) = ((const uint4*)net.w1t)[q];
	}
	for (int q = t; q < MLP_H * (MLP_H / 8); q += FWD_THREADS)
	{
		int row = q / (MLP_H / 8), c8 = q - row * (MLP_H / 8);
		*(uint4*)(sW2 + row * LDS_H + 8 * c8) = ((const uint4*)net.w2t)[q];
	}
	for (int q = t; q < MLP_K3 * (MLP_H / 8); q += FWD_THREADS)
	{
		int row = q / (MLP_H / 8), c8 = q - row * (MLP_H / 8);
		*(uint4*)(sW3 + row * LDS_H + 8 * c8) = ((const uint4*)net.w3t)[q];
	}
	__syncthreads();
	int wave = t >> 6, lane = t & 63, r = lane & 31, h = lane >> 5;
	unsigned short* act = sAct + wave * 32 * LDS_H;
	for (int row0 = (blockIdx.x * FWD_WAVES + wave) * 32; row0 < Mp; row0 += gridDim.x * FWD_WAVES * 32)
	{
		/* input tile -> bf16 [32][K1] */
		for (int q = lane; q < 32 * MLP_K1; q += 64)
		{
			int rr = q / MLP_K1, c = q - rr * MLP_K1;
			int row = row0 + rr;
			float v = (row < M && c < inDim) ? (X[(size_t)row * ldx + c] - net.inShift[c]) * net.inScale[c] : 0.f;
			unsigned short b = f2bf(v);
			act[rr * LDS_H + c] = b;
			if (save) Xb[(size_t)row * MLP_K1 + c] = b;
		}
		wave_lds_sync();
		floatx16 z[4];
		/* layer 1: K = 32 */
#pragma unroll
		for (int j = 0; j < 4; j++)
		{
			z[j] = (floatx16){};
#pragma unroll
			for (int k = 0; k < MLP_K1; k += 16)
			{
				bf16x8 a = *(const bf16x8*)(act + r * LDS_H + k + 8 * h);
				bf16x8 b = *(const bf16x8*)(sW1 + (j * 32 + r) * LDS_K1 + k + 8 * h);
				z[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, z[j], 0, 0, 0);
			}
		}
		wave_lds_sync();
#pragma unroll
		for (int j = 0; j < 4; j++)
		{
			int c = j * 32 + r;
			float bias = net.b1[c];
#pragma unroll
			for (int i = 0; i < 16; i++)
			{
				int rr = 8 * (i >> 2) + 4 * h + (i & 3);
				float zz = z[j][i] + bias;
				unsigned short hb = f2bf(gelu(zz));
				act[rr * LDS_H + c] = hb;
				if (save)
				{
					size_t o = (size_t)(row0 + rr) * MLP_H + c;
					Z1[o] = zz;
					H1[o] = hb;
				}
			}
		}
		wave_lds_sync();
		/* layer 2: K = 128 */
#pragma unroll
		for (int j = 0; j < 4; j++)
		{
			z[j] = (floatx16){};
#pragma unroll
			for (int k = 0; k < MLP_H; k += 16)
			{
				bf16x8 a = *(const bf16x8*)(act + r * LDS_H + k + 8 * h);
				bf16x8 b = *(const bf16x8*)(sW2 + (j * 32 + r) * LDS_H + k + 8 * h);
				z[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, z[j], 0, 0, 0);
			}
		}
		wave_lds_sync();
#pragma unroll
		for (int j = 0; j < 4; j++)
		{
			int c = j * 32 + r;
			float bias = net.b2[c];
#pragma unroll
			for (int i = 0; i < 16; i++)
			{
				int rr = 8 * (i >> 2) + 4 * h + (i & 3);
				float zz = z[j][i] + bias;
				unsigned short hb = f2bf(gelu(zz));
				act[rr * LDS_H + c] = hb;
				if (save)
				{
					size_t o = (size_t)(row0 + rr) * MLP_H + c;
					Z2[o] = zz;
					H2[o] = hb;
				}
			}
		}
		wave_lds_sync();
		/* layer 3: K = 128, 32 output columns */
		floatx16 y = {};
#pragma unroll
		for (int k = 0; k < MLP_H; k += 16)
		{
			bf16x8 a = *(const bf16x8*)(act + r * LDS_H + k + 8 * h);
			bf16x8 b = *(const bf16x8*)(sW3 + r * LDS_H + k + 8 * h);
			y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, y, 0, 0, 0);
		}
		if (r < outDim)
		{
			float bias = net.b3[r];
#pragma unroll
			for (int i = 0; i < 16; i++)
			{
				int row = row0 + 8 * (i >> 2) + 4 * h + (i & 3);
				if (row < M) Y[(size_t)row * ldy + r] = y[i] + bias;
			}
		}
		wave_lds_sync();
	}
}

/* ---------------- backward pieces ---------------- */
/* dY = 2 (Y - T) / (M outDim) for rows < M, 0 for padding rows; one block so
 * the loss sum has a fixed order */
__global__ __launch_bounds__(256) void k_mse_grad(int M, int Mp, int outDim, const float* __restrict__ Y,
												  const float* __restrict__ T, int ldt, float* __restrict__ dY,
												  float* __restrict__ lossOut)
{
	__shared__ float red[256];
	int t = threadIdx.x;
	float acc = 0.f;
	float scale = 2.f / (float)((double)M * outDim);
	for (int i = t; i < Mp * outDim; i += 256)
	{
		int row = i / outDim, c = i - row * outDim;
		float d = row < M ? Y[i] - T[(size_t)row * ldt + c] : 0.f;
		dY[i] = scale * d;
		acc += d * d;
	}
	red[t] = acc;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1)
	{
		if (t < o) red[t] += red[t + o];
		__syncthreads();
	}
	if (t == 0) *lossOut = red[0] / (float)((double)M * outDim);
}
/* dY fp32 [Mp][outDim] -> bf16 [Mp][K3] (zero padded) */
__global__ __launch_bounds__(256) void k_pack_dy(int Mp, int outDim, const float* __restrict__ dY,
												 unsigned short* __restrict__ D)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= Mp * MLP_K3) return;
	int row = i / MLP_K3, c = i - row * MLP_K3;
	D[i] = c < outDim ? f2bf(dY[(size_t)row * outDim + c]) : (unsigned short)0;
}
/* dZ = dAct * gelu'(Z): fp32 (bias gradients) and bf16 (products) */
__global__ __launch_bounds__(256) void k_gelu_back(int n, const float* __restrict__ dAct, const float* __restrict__ Z,
												   float* __restrict__ dZ, unsigned short* __restrict__ G)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float g = dAct[i] * gelu_grad(Z[i]);
	dZ[i] = g;
	G[i] = f2bf(g);
}
/* S [rows][cols] -> D [cols][rows], bf16, through a 32x32 LDS tile */
__global__ __launch_bounds__(256) void k_transpose_bf16(int rows, int cols, const unsigned short* __restrict__ S,
														unsigned short* __restrict__ D)
{
	__shared__ unsigned short tile[32][33];
	int bx = blockIdx.x * 32, by = blockIdx.y * 32; /* bx: column block of S, by: row block */
	int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for (int k = ty; k < 32; k += 8)
	{
		int rr = by + k, cc = bx + tx;
		tile[k][tx] = (rr < rows && cc < cols) ? S[(size_t)rr * cols + cc] : (unsigned short)0;
	}
	__syncthreads();
	for (int k = ty; k < 32; k += 8)
	{
		int cc = bx + k, rr = by + tx;
		if (cc < cols && rr < rows) D[(size_t)cc * rows + rr] = tile[tx][k];
	}
}
/* column sums of fp32 A [rows][cols]: partials over 256-row chunks, then the
 * chunks in order (a fixed summation order) */
__global__ __launch_bounds__(128) void k_colsum_part(int rows, int cols, const float* __restrict__ A,
													 float* __restrict__ part)
{
	int c = blockIdx.y * 128 + threadIdx.x;
	if (c >= cols) return;
	int r0 = blockIdx.x * 256, r1 = r0 + 256 < rows ? r0 + 256 : rows;
	float s = 0.f;
	for (int r = r0; r < r1; r++) s += A[(size_t)r * cols + c];
	part[(size_t)blockIdx.x * cols + c] = s;
}
__global__ __launch_bounds__(128) void k_colsum_final(int chunks, int cols, const float* __restrict__ part,
													  float* __restrict__ out)
{
	int c = blockIdx.x * 128 + threadIdx.x;
	if (c >= cols) return;
	float s = 0.f;
	for (int k = 0; k < chunks; k++) s += part[(size_t)k * cols + c];
	out[c] = s;
}
__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ w, const float* __restrict__ g,
											  float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
											  float eps, float c1, float c2)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float gi = g[i];
	float mi = b1 * m[i] + (1.f - b1) * gi;
	float vi = b2 * v[i] + (1.f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
/* bf16 copies of a fp32 [rows][cols] master: transposed [cols][rows] and/or as is */
__global__ __launch_bounds__(256) void k_pack_weight(int rows, int cols, const float* __restrict__ W,
													 unsigned short* __restrict__ WT, unsigned short* __restrict__ WN)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= rows * cols) return;
	int r = i / cols, c = i - r * cols;
	unsigned short b = f2bf(W[i]);
	if (WT) WT[(size_t)c * rows + r] = b;
	if (WN) WN[i] = b;
}
__global__ __launch_bounds__(256) void k_fill_f32(float* p, float v, int n)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}
/* minibatch: row j takes source row hash(seed, step, j) mod n */
__device__ __forceinline__ unsigned mix32(unsigned x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}
__global__ __launch_bounds__(64) void k_gather_batch(int n, int rows, int inDim, int outDim, const float* __restrict__ X,
													 int ldx, const float* __restrict__ T, int ldt, unsigned seed,
													 int step, float* __restrict__ xb, float* __restrict__ tb)
{
	int j = blockIdx.x;
	if (j >= rows) return;
	unsigned hsh = mix32(seed ^ mix32((unsigned)step * 0x9e3779b9u + (unsigned)j));
	size_t src = (size_t)(((unsigned long long)hsh * (unsigned long long)n) >> 32);
	int c = threadIdx.x;
	if (c < MLP_K1) xb[(size_t)j * MLP_K1 + c] = c < inDim ? X[src * ldx + c] : 0.f;
	if (c < outDim) tb[(size_t)j * outDim + c] = T[src * ldt + c];
}

/* ======================= launchers ======================= */
static inline unsigned nblk(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }
static inline int pad32(int n) { return (n + 31) & ~31; }

void b3h_mlp_fill_f32(float* p, float v, int n, hipStream_t s)
{
	if (n > 0) hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n)), dim3(256), 0, s, p, v, n);
}

void b3h_mlp_gemm(int M, int N, int K, const unsigned short* A, int lda, const unsigned short* BT, int ldb, float* C,
				  float* partials, int splits, hipStream_t s)
{
	int kChunk = ((K / splits + 15) / 16) * 16;
	if (kChunk < 16) kChunk = 16;
	splits = (K + kChunk - 1) / kChunk;
	int waves = (M / 32) * (N / 32) * splits;
	hipLaunchKernelGGL(k_gemm_bf16, dim3((waves + 3) / 4), dim3(256), 0, s, M, N, K, kChunk, splits, A, lda, BT, ldb,
					   splits == 1 ? C : partials);
	if (splits > 1)
		hipLaunchKernelGGL(k_gemm_reduce, dim3(nblk((size_t)M * N)), dim3(256), 0, s, M * N, splits,
						   (const float*)partials, C);
}

void b3h_mlp_pack(b3hMlp& m, hipStream_t s)
{
	const float* P = m.params;
	hipLaunchKernelGGL(k_pack_weight, dim3(nblk(MLP_K1 * MLP_H)), dim3(256), 0, s, MLP_K1, MLP_H, P + MLP_OFF_W1, m.w1t,
					   (unsigned short*)nullptr);
	hipLaunchKernelGGL(k_pack_weight, dim3(nblk(MLP_H * MLP_H)), dim3(256), 0, s, MLP_H, MLP_H, P + MLP_OFF_W2, m.w2t,
					   m.w2n);
	hipLaunchKernelGGL(k_pack_weight, dim3(nblk(MLP_H * MLP_K3)), dim3(256), 0, s, MLP_H, MLP_K3, P + MLP_OFF_W3, m.w3t,
					   m.w3n);
}

static void forward(const b3hMlp& m, int rows, const float* X, int ldx, float* Y, int ldy, bool save, hipStream_t s)
{
	int Mp = pad32(rows);
	if (Mp == 0) return;
	MlpOps net = {m.w1t,	m.w2t,	  m.w3t,	  m.params + MLP_OFF_B1, m.params + MLP_OFF_B2, m.params + MLP_OFF_B3,
				  m.inShift, m.inScale};
	int tiles = Mp / 32;
	int blocks = (tiles + FWD_WAVES - 1) / FWD_WAVES;
	if (blocks > 512) blocks = 512;
	const b3hMlpWork& w = m.wk;
	hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(FWD_THREADS), 0, s, rows, Mp, (const int*)nullptr, 0, m.inDim,
					   m.outDim, X, ldx, net, Y, ldy, save ? 1 : 0, save ? w.xb : nullptr, save ? w.z1 : nullptr, save ? w.h1 : nullptr,
					   save ? w.z2 : nullptr, save ? w.h2 : nullptr);
}

void b3h_mlp_run_forward(const b3hMlp& m, int rows, const float* X, int ldx, float* Y, int ldy, hipStream_t s)
{
	forward(m, rows, X, ldx, Y, ldy, false, s);
}

void b3h_mlp_run_forward_dev(const b3hMlp& m, const int* countDev, int rowsPerItem, int maxRows, const float* X,
							 int ldx, float* Y, int ldy, hipStream_t s)
{
	MlpOps net = {m.w1t,	m.w2t,	  m.w3t,	  m.params + MLP_OFF_B1, m.params + MLP_OFF_B2, m.params + MLP_OFF_B3,
				  m.inShift, m.inScale};
	int tiles = (maxRows + 31) / 32;
	int blocks = (tiles + FWD_WAVES - 1) / FWD_WAVES;
	if (blocks > 512) blocks = 512;
	if (blocks < 1) blocks = 1;
	hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(FWD_THREADS), 0, s, 0, 0, countDev, rowsPerItem, m.inDim, m.outDim, X,
					   ldx, net, Y, ldy, 0, (unsigned short*)nullptr, (float*)nullptr, (unsigned short*)nullptr,
					   (float*)nullptr, (unsigned short*)nullptr);
}

static void transpose(int rows, int cols, const unsigned short* S, unsigned short* D, hipStream_t s)
{
	hipLaunchKernelGGL(k_transpose_bf16, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, s, rows, cols, S, D);
}
static void colsum(int rows, int cols, const float* A, float* part, float* out, hipStream_t s)
{
	int chunks = (rows + 255) / 256;
	hipLaunchKernelGGL(k_colsum_part, dim3(chunks, (cols + 127) / 128), dim3(128), 0, s, rows, cols, A, part);
	hipLaunchKernelGGL(k_colsum_final, dim3((cols + 127) / 128), dim3(128), 0, s, chunks, cols, (const float*)part, out);
}

/* Gradients of the mean squared error, chain rule over the saved forward:
 *   dY  = 2 (Y - T) / (M out)            db3 = colsum dY      dW3 = H2^T dY
 *   dZ2 = (dY W3^T) * gelu'(Z2)          db2 = colsum dZ2     dW2 = H1^T dZ2
 *   dZ1 = (dZ2 W2^T) * gelu'(Z1)         db1 = colsum dZ1     dW1 = X^T dZ1
 * with the bf16 operands of each product as stored by the forward pass. */
void b3h_mlp_run_backward(b3hMlp& m, int rows, const float* X, int ldx, const float* T, int ldt, float* lossOut,
						  hipStream_t s)
{
	b3hMlpWork& w = m.wk;
	int Mp = pad32(rows);
	int out = m.outDim;
	float* G = m.grads;
	forward(m, rows, X, ldx, w.y, out, true, s);
	hipLaunchKernelGGL(k_mse_grad, dim3(1), dim3(256), 0, s, rows, Mp, out, (const float*)w.y, T, ldt, w.dy, lossOut);
	hipLaunchKernelGGL(k_pack_dy, dim3(nblk((size_t)Mp * MLP_K3)), dim3(256), 0, s, Mp, out, (const float*)w.dy, w.dyb);
	int splits = Mp >= 2048 ? Mp / 1024 : 1;
	if (splits > 32) splits = 32;
	/* layer 3 */
	transpose(Mp, MLP_H, w.h2, w.h2t, s);
	transpose(Mp, MLP_K3, w.dyb, w.dybt, s);
	b3h_mlp_gemm(MLP_H, MLP_K3, Mp, w.h2t, Mp, w.dybt, Mp, G + MLP_OFF_W3, w.part, splits, s);
	b3h_mlp_fill_f32(G + MLP_OFF_B3, 0.f, MLP_K3, s);
	colsum(Mp, out, w.dy, w.colPart, G + MLP_OFF_B3, s);
	b3h_mlp_gemm(Mp, MLP_H, MLP_K3, w.dyb, MLP_K3, m.w3n, MLP_K3, w.dact, w.part, 1, s);
	hipLaunchKernelGGL(k_gelu_back, dim3(nblk((size_t)Mp * MLP_H)), dim3(256), 0, s, Mp * MLP_H, (const float*)w.dact,
					   (const float*)w.z2, w.dz, w.gz);
	/* layer 2 */
	transpose(Mp, MLP_H, w.h1, w.h1t, s);
	transpose(Mp, MLP_H, w.gz, w.gzt, s);
	b3h_mlp_gemm(MLP_H, MLP_H, Mp, w.h1t, Mp, w.gzt, Mp, G + MLP_OFF_W2, w.part, splits, s);
	colsum(Mp, MLP_H, w.dz, w.colPart, G + MLP_OFF_B2, s);
	b3h_mlp_gemm(Mp, MLP_H, MLP_H, w.gz, MLP_H, m.w2n, MLP_H, w.dact, w.part, 1, s);
	hipLaunchKernelGGL(k_gelu_back, dim3(nblk((size_t)Mp * MLP_H)), dim3(256), 0, s, Mp * MLP_H, (const float*)w.dact,
					   (const float*)w.z1, w.dz, w.gz);
	/* layer 1 */
	transpose(Mp, MLP_K1, w.xb, w.xbt, s);
	transpose(Mp, MLP_H, w.gz, w.gzt, s);
	b3h_mlp_gemm(MLP_K1, MLP_H, Mp, w.xbt, Mp, w.gzt, Mp, G + MLP_OFF_W1, w.part, splits, s);
	colsum(Mp, MLP_H, w.dz, w.colPart, G + MLP_OFF_B1, s);
}

void b3h_mlp_run_adam(b3hMlp& m, float lr, hipStream_t s)
{
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	m.adamStep++;
	double c1 = 1.0 - __builtin_pow(0.9, m.adamStep), c2 = 1.0 - __builtin_pow(0.999, m.adamStep);
	hipLaunchKernelGGL(k_adam, dim3(nblk(MLP_NPARAM)), dim3(256), 0, s, MLP_NPARAM, m.params, (const float*)m.grads,
					   m.adamM, m.adamV, lr, b1, b2, eps, (float)c1, (float)c2);
	b3h_mlp_pack(m, s);
}

void b3h_mlp_gather_batch(b3hMlp& m, int n, const float* X, int ldx, const float* T, int ldt, int rows,
						  unsigned seed, int step, hipStream_t s)
{
	hipLaunchKernelGGL(k_gather_batch, dim3(rows), dim3(64), 0, s, n, rows, m.inDim, m.outDim, X, ldx, T, ldt, seed,
					   step, m.wk.xBatch, m.wk.tBatch);
}
