/*
 * b3h_mlp.h -- the learned warm start's network: a 3-layer MLP
 * (in -> 128 -> 128 -> out, GELU, bf16 weights and activations, fp32
 * accumulation) run and trained on the matrix cores.  Internal to the library.
 */
#ifndef B3H_MLP_H
#define B3H_MLP_H

#include <hip/hip_runtime.h>

#define MLP_H 128  /* hidden width */
#define MLP_K1 32  /* input features, zero-padded */
#define MLP_K3 32  /* outputs, zero-padded */
/* fp32 parameter block, padded layout: w1 [K1][H] | b1 [H] | w2 [H][H] | b2 [H] | w3 [H][K3] | b3 [K3] */
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + MLP_K1 * MLP_H)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_H * MLP_K3)
#define MLP_NPARAM (MLP_OFF_B3 + MLP_K3)

struct b3hMlpWork
{
	int capRows;               /* rows the buffers hold (multiple of 32) */
	unsigned short* xb;        /* [rows][K1] bf16 input */
	unsigned short* xbt;       /* [K1][rows] */
	float* z1;                 /* [rows][H] pre-activations */
	float* z2;
	unsigned short* h1;        /* [rows][H] bf16 activations */
	unsigned short* h2;
	unsigned short* h1t;       /* [H][rows] */
	unsigned short* h2t;
	float* y;                  /* [rows][out] */
	float* dy;                 /* [rows][out] */
	unsigned short* dyb;       /* [rows][K3] */
	unsigned short* dybt;      /* [K3][rows] */
	float* dact;               /* [rows][H] gradient w.r.t. an activation */
	float* dz;                 /* [rows][H] gradient w.r.t. a pre-activation (fp32) */
	unsigned short* gz;        /* [rows][H] the same in bf16 */
	unsigned short* gzt;       /* [H][rows] */
	float* part;               /* split-K partial products */
	float* colPart;            /* column-sum partials */
	float* loss;               /* per training step */
	float* xBatch;             /* gathered minibatch inputs [rows][K1] */
	float* tBatch;             /* gathered minibatch targets [rows][out] */
};

struct b3hMlp
{
	int inDim, outDim;
	hipStream_t stream;
	float* params;             /* MLP_NPARAM fp32 masters */
	float* inShift;            /* [K1] input normalisation x' = (x - shift) * scale (part of the model) */
	float* inScale;            /* [K1] */
	float* grads;              /* MLP_NPARAM */
	float* adamM;
	float* adamV;
	int adamStep;
	unsigned short* w1t;       /* [H][K1]  W1 transposed (forward operand) */
	unsigned short* w2t;       /* [H][H]   W2 transposed */
	unsigned short* w3t;       /* [K3][H]  W3 transposed */
	unsigned short* w2n;       /* [H][H]   W2 as is (backward operand) */
	unsigned short* w3n;       /* [H][K3]  W3 as is */
	b3hMlpWork wk;
	int maxTrainRows;
};

/* Y[rows][outDim] (ldy) = net(X[rows][inDim] (ldx)); all device pointers */
void b3h_mlp_run_forward(const b3hMlp& m, int rows, const float* X, int ldx, float* Y, int ldy, hipStream_t s);
/* the same with the row count read on the device: rows = *countDev * rowsPerItem
 * (at most maxRows, which sizes the grid) -- for graph-captured steps */
void b3h_mlp_run_forward_dev(const b3hMlp& m, const int* countDev, int rowsPerItem, int maxRows, const float* X,
							 int ldx, float* Y, int ldy, hipStream_t s);
/* loss and gradients into m.grads for one batch (rows <= m.maxTrainRows) */
void b3h_mlp_run_backward(b3hMlp& m, int rows, const float* X, int ldx, const float* T, int ldt, float* lossOut,
						  hipStream_t s);
/* Adam update of the masters from m.grads, then the bf16 copies */
void b3h_mlp_run_adam(b3hMlp& m, float lr, hipStream_t s);
/* the bf16 operand copies from the fp32 masters */
void b3h_mlp_pack(b3hMlp& m, hipStream_t s);
/* gather rows idx = hash(seed, step, j) mod n into the batch buffers */
void b3h_mlp_gather_batch(b3hMlp& m, int n, const float* X, int ldx, const float* T, int ldt, int rows,
						  unsigned seed, int step, hipStream_t s);
/* C[M x N] = A[M x K] * B, A row-major, B given as BT[N x K] row-major; M, N
 * multiples of 32, K a multiple of 16; splits > 1 sums K-slices in order */
void b3h_mlp_gemm(int M, int N, int K, const unsigned short* A, int lda, const unsigned short* BT, int ldb, float* C,
				  float* partials, int splits, hipStream_t s);
void b3h_mlp_fill_f32(float* p, float v, int n, hipStream_t s);

#endif
