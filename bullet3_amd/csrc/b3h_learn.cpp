/*
 * b3h_learn.cpp -- C ABI of the learned warm start's network (b3hMlp): create,
 * parameters in and out, inference, loss gradients, minibatch Adam training.
 * Everything runs on the device through the kernels of k_mlp.hip; the host
 * only sequences launches.  No counterpart in the reference, whose solver
 * always starts from zero impulses (b3ConvertConstraint4.h:73,107,132).
 */
#include <cmath>
#include <cstring>
#include <vector>
#include "b3h_internal.h"
#include "b3h_mlp.h"

#define CHECK_ARG(c)                                   \
	do                                                 \
	{                                                  \
		if (!(c))                                      \
		{                                              \
			b3h_set_error("invalid argument: %s", #c); \
			return B3H_ERR_INVALID;                    \
		}                                              \
	} while (0)

template <class T>
static int mlp_alloc(T** p, size_t n)
{
	*p = nullptr;
	if (n == 0) n = 1;
	hipError_t e = hipMalloc((void**)p, n * sizeof(T));
	if (e != hipSuccess)
	{
		b3h_set_error("hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
		*p = nullptr;
		return B3H_ERR_HIP;
	}
	return B3H_OK;
}
template <class T>
static void mlp_free(T*& p)
{
	if (p) hipFree((void*)p);
	p = nullptr;
}

static void free_work(b3hMlpWork& w)
{
	mlp_free(w.xb), mlp_free(w.xbt), mlp_free(w.z1), mlp_free(w.z2), mlp_free(w.h1), mlp_free(w.h2);
	mlp_free(w.h1t), mlp_free(w.h2t), mlp_free(w.y), mlp_free(w.dy), mlp_free(w.dyb), mlp_free(w.dybt);
	mlp_free(w.dact), mlp_free(w.dz), mlp_free(w.gz), mlp_free(w.gzt), mlp_free(w.part), mlp_free(w.colPart);
	mlp_free(w.loss), mlp_free(w.xBatch), mlp_free(w.tBatch);
	w.capRows = 0;
}

static int alloc_work(b3hMlpWork& w, int rows, int outDim, int maxSteps)
{
	size_t R = (size_t)rows;
	size_t H = MLP_H;
	int r = 0;
	(r = mlp_alloc(&w.xb, R * MLP_K1)) || (r = mlp_alloc(&w.xbt, R * MLP_K1)) || (r = mlp_alloc(&w.z1, R * H)) ||
		(r = mlp_alloc(&w.z2, R * H)) || (r = mlp_alloc(&w.h1, R * H)) || (r = mlp_alloc(&w.h2, R * H)) ||
		(r = mlp_alloc(&w.h1t, R * H)) || (r = mlp_alloc(&w.h2t, R * H)) || (r = mlp_alloc(&w.y, R * outDim)) ||
		(r = mlp_alloc(&w.dy, R * outDim)) || (r = mlp_alloc(&w.dyb, R * MLP_K3)) ||
		(r = mlp_alloc(&w.dybt, R * MLP_K3)) || (r = mlp_alloc(&w.dact, R * H)) || (r = mlp_alloc(&w.dz, R * H)) ||
		(r = mlp_alloc(&w.gz, R * H)) || (r = mlp_alloc(&w.gzt, R * H)) ||
		(r = mlp_alloc(&w.part, (size_t)32 * MLP_H * MLP_H)) || (r = mlp_alloc(&w.colPart, (R / 256 + 1) * H)) ||
		(r = mlp_alloc(&w.loss, (size_t)maxSteps)) || (r = mlp_alloc(&w.xBatch, R * MLP_K1)) ||
		(r = mlp_alloc(&w.tBatch, R * outDim));
	if (r)
	{
		free_work(w);
		return r;
	}
	w.capRows = rows;
	return B3H_OK;
}

#define MLP_MAX_STEPS_PER_CALL 65536

/* fp32 master layout (padded) <-> packed logical layout
 * w1[in][H] b1[H] w2[H][H] b2[H] w3[H][out] b3[out] */
static int packed_count(int in, int out) { return in * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * out + out; }

static void packed_to_padded(const float* src, int in, int out, float* P)
{
	memset(P, 0, sizeof(float) * MLP_NPARAM);
	const float* s = src;
	for (int i = 0; i < in; i++)
		for (int j = 0; j < MLP_H; j++) P[MLP_OFF_W1 + i * MLP_H + j] = *s++;
	for (int j = 0; j < MLP_H; j++) P[MLP_OFF_B1 + j] = *s++;
	for (int i = 0; i < MLP_H * MLP_H; i++) P[MLP_OFF_W2 + i] = *s++;
	for (int j = 0; j < MLP_H; j++) P[MLP_OFF_B2 + j] = *s++;
	for (int i = 0; i < MLP_H; i++)
		for (int j = 0; j < out; j++) P[MLP_OFF_W3 + i * MLP_K3 + j] = *s++;
	for (int j = 0; j < out; j++) P[MLP_OFF_B3 + j] = *s++;
}
static void padded_to_packed(const float* P, int in, int out, float* dst)
{
	float* d = dst;
	for (int i = 0; i < in; i++)
		for (int j = 0; j < MLP_H; j++) *d++ = P[MLP_OFF_W1 + i * MLP_H + j];
	for (int j = 0; j < MLP_H; j++) *d++ = P[MLP_OFF_B1 + j];
	for (int i = 0; i < MLP_H * MLP_H; i++) *d++ = P[MLP_OFF_W2 + i];
	for (int j = 0; j < MLP_H; j++) *d++ = P[MLP_OFF_B2 + j];
	for (int i = 0; i < MLP_H; i++)
		for (int j = 0; j < out; j++) *d++ = P[MLP_OFF_W3 + i * MLP_K3 + j];
	for (int j = 0; j < out; j++) *d++ = P[MLP_OFF_B3 + j];
}

/* default initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for w1, w2,
 * biases zero, and the output layer scaled by 0.1 so an untrained network
 * predicts small corrections */
static void init_params(int in, int out, unsigned seed, std::vector<float>& packed)
{
	packed.assign(packed_count(in, out), 0.f);
	unsigned long long st = 0x9e3779b97f4a7c15ull ^ seed;
	auto uni = [&]() {
		st ^= st << 13;
		st ^= st >> 7;
		st ^= st << 17;
		return (float)((st >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
	};
	float* p = packed.data();
	float s1 = 1.f / std::sqrt((float)in), s2 = 1.f / std::sqrt((float)MLP_H);
	for (int i = 0; i < in * MLP_H; i++) *p++ = s1 * uni();
	p += MLP_H;
	for (int i = 0; i < MLP_H * MLP_H; i++) *p++ = s2 * uni();
	p += MLP_H;
	for (int i = 0; i < MLP_H * out; i++) *p++ = 0.1f * s2 * uni();
}

extern "C" int b3h_mlp_create(int inDim, int outDim, int maxTrainRows, unsigned seed, void* hipStream, b3hMlp** out)
{
	CHECK_ARG(out && inDim >= 1 && inDim <= MLP_K1 && outDim >= 1 && outDim <= MLP_K3 && maxTrainRows >= 1);
	*out = nullptr;
	b3hMlp* m = new b3hMlp();
	memset(m, 0, sizeof(*m));
	m->inDim = inDim;
	m->outDim = outDim;
	m->stream = (hipStream_t)hipStream;
	m->maxTrainRows = (maxTrainRows + 31) & ~31;
	int r = 0;
	(r = mlp_alloc(&m->params, MLP_NPARAM)) || (r = mlp_alloc(&m->grads, MLP_NPARAM)) ||
		(r = mlp_alloc(&m->inShift, MLP_K1)) || (r = mlp_alloc(&m->inScale, MLP_K1)) ||
		(r = mlp_alloc(&m->adamM, MLP_NPARAM)) || (r = mlp_alloc(&m->adamV, MLP_NPARAM)) ||
		(r = mlp_alloc(&m->w1t, (size_t)MLP_H * MLP_K1)) || (r = mlp_alloc(&m->w2t, (size_t)MLP_H * MLP_H)) ||
		(r = mlp_alloc(&m->w3t, (size_t)MLP_K3 * MLP_H)) || (r = mlp_alloc(&m->w2n, (size_t)MLP_H * MLP_H)) ||
		(r = mlp_alloc(&m->w3n, (size_t)MLP_H * MLP_K3)) ||
		(r = alloc_work(m->wk, m->maxTrainRows, outDim, MLP_MAX_STEPS_PER_CALL));
	if (r)
	{
		b3h_mlp_destroy(m);
		return r;
	}
	{
		std::vector<float> zero(MLP_K1, 0.f), one(MLP_K1, 1.f);
		hipError_t e1 = hipMemcpy(m->inShift, zero.data(), sizeof(float) * MLP_K1, hipMemcpyHostToDevice);
		hipError_t e2 = hipMemcpy(m->inScale, one.data(), sizeof(float) * MLP_K1, hipMemcpyHostToDevice);
		if (e1 != hipSuccess || e2 != hipSuccess)
		{
			b3h_set_error("hipMemcpy: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
			b3h_mlp_destroy(m);
			return B3H_ERR_HIP;
		}
	}
	std::vector<float> packed;
	init_params(inDim, outDim, seed, packed);
	return (r = b3h_mlp_set_params(m, packed.data())) ? (b3h_mlp_destroy(m), r) : (*out = m, B3H_OK);
}

extern "C" int b3h_mlp_destroy(b3hMlp* m)
{
	if (!m) return B3H_OK;
	if (m->stream) hipStreamSynchronize(m->stream);
	else hipDeviceSynchronize();
	free_work(m->wk);
	mlp_free(m->params), mlp_free(m->grads), mlp_free(m->adamM), mlp_free(m->adamV);
	mlp_free(m->inShift), mlp_free(m->inScale);
	mlp_free(m->w1t), mlp_free(m->w2t), mlp_free(m->w3t), mlp_free(m->w2n), mlp_free(m->w3n);
	delete m;
	return B3H_OK;
}

extern "C" int b3h_mlp_num_params(const b3hMlp* m, int* n)
{
	CHECK_ARG(m && n);
	*n = packed_count(m->inDim, m->outDim);
	return B3H_OK;
}

extern "C" int b3h_mlp_set_params(b3hMlp* m, const float* host)
{
	CHECK_ARG(m && host);
	std::vector<float> P(MLP_NPARAM);
	packed_to_padded(host, m->inDim, m->outDim, P.data());
	B3H_CHECK(hipMemcpyAsync(m->params, P.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream));
	b3h_mlp_fill_f32(m->adamM, 0.f, MLP_NPARAM, m->stream);
	b3h_mlp_fill_f32(m->adamV, 0.f, MLP_NPARAM, m->stream);
	m->adamStep = 0;
	b3h_mlp_pack(*m, m->stream);
	B3H_CHECK(hipStreamSynchronize(m->stream));
	return B3H_OK;
}

extern "C" int b3h_mlp_get_params(b3hMlp* m, float* host)
{
	CHECK_ARG(m && host);
	std::vector<float> P(MLP_NPARAM);
	B3H_CHECK(hipMemcpyAsync(P.data(), m->params, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
	B3H_CHECK(hipStreamSynchronize(m->stream));
	padded_to_packed(P.data(), m->inDim, m->outDim, host);
	return B3H_OK;
}

extern "C" int b3h_mlp_forward(b3hMlp* m, const float* xDev, int ldx, int rows, float* yDev, int ldy)
{
	CHECK_ARG(m && rows >= 0 && (rows == 0 || (xDev && yDev)) && ldx >= m->inDim && ldy >= m->outDim);
	b3h_mlp_run_forward(*m, rows, xDev, ldx, yDev, ldy, m->stream);
	B3H_CHECK(hipGetLastError());
	return B3H_OK;
}

extern "C" int b3h_mlp_loss_gradients(b3hMlp* m, const float* xDev, int ldx, const float* tDev, int ldt, int rows,
									  float* gradsHost, float* lossOut)
{
	CHECK_ARG(m && xDev && tDev && rows >= 1 && rows <= m->maxTrainRows && ldx >= m->inDim && ldt >= m->outDim);
	b3h_mlp_run_backward(*m, rows, xDev, ldx, tDev, ldt, m->wk.loss, m->stream);
	B3H_CHECK(hipGetLastError());
	std::vector<float> G(MLP_NPARAM);
	float loss = 0.f;
	B3H_CHECK(hipMemcpyAsync(G.data(), m->grads, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
	B3H_CHECK(hipMemcpyAsync(&loss, m->wk.loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
	B3H_CHECK(hipStreamSynchronize(m->stream));
	if (gradsHost) padded_to_packed(G.data(), m->inDim, m->outDim, gradsHost);
	if (lossOut) *lossOut = loss;
	return B3H_OK;
}

extern "C" int b3h_mlp_train(b3hMlp* m, const float* xDev, int ldx, const float* tDev, int ldt, int rows,
							 int batchRows, int steps, float learningRate, unsigned seed, float* lossHistory)
{
	CHECK_ARG(m && xDev && tDev && rows >= 1 && batchRows >= 1 && batchRows <= m->maxTrainRows && steps >= 0 &&
			  steps <= MLP_MAX_STEPS_PER_CALL && ldx >= m->inDim && ldt >= m->outDim && learningRate > 0.f);
	hipStream_t s = m->stream;
	for (int k = 0; k < steps; k++)
	{
		b3h_mlp_gather_batch(*m, rows, xDev, ldx, tDev, ldt, batchRows, seed, m->adamStep, s);
		b3h_mlp_run_backward(*m, batchRows, m->wk.xBatch, MLP_K1, m->wk.tBatch, m->outDim, m->wk.loss + k, s);
		b3h_mlp_run_adam(*m, learningRate, s);
	}
	B3H_CHECK(hipGetLastError());
	if (lossHistory && steps > 0)
		B3H_CHECK(hipMemcpyAsync(lossHistory, m->wk.loss, sizeof(float) * steps, hipMemcpyDeviceToHost, s));
	B3H_CHECK(hipStreamSynchronize(s));
	return B3H_OK;
}

extern "C" int b3h_mlp_set_input_normalization(b3hMlp* m, const float* shift, const float* scale)
{
	CHECK_ARG(m && shift && scale);
	std::vector<float> sh(MLP_K1, 0.f), sc(MLP_K1, 1.f);
	for (int i = 0; i < m->inDim; i++)
	{
		sh[i] = shift[i];
		sc[i] = scale[i];
	}
	B3H_CHECK(hipMemcpyAsync(m->inShift, sh.data(), sizeof(float) * MLP_K1, hipMemcpyHostToDevice, m->stream));
	B3H_CHECK(hipMemcpyAsync(m->inScale, sc.data(), sizeof(float) * MLP_K1, hipMemcpyHostToDevice, m->stream));
	B3H_CHECK(hipStreamSynchronize(m->stream));
	return B3H_OK;
}
extern "C" int b3h_mlp_get_input_normalization(b3hMlp* m, float* shift, float* scale)
{
	CHECK_ARG(m && shift && scale);
	float sh[MLP_K1], sc[MLP_K1];
	B3H_CHECK(hipMemcpyAsync(sh, m->inShift, sizeof(float) * MLP_K1, hipMemcpyDeviceToHost, m->stream));
	B3H_CHECK(hipMemcpyAsync(sc, m->inScale, sizeof(float) * MLP_K1, hipMemcpyDeviceToHost, m->stream));
	B3H_CHECK(hipStreamSynchronize(m->stream));
	for (int i = 0; i < m->inDim; i++)
	{
		shift[i] = sh[i];
		scale[i] = sc[i];
	}
	return B3H_OK;
}
extern "C" int b3h_mlp_get_dims(const b3hMlp* m, int* inDim, int* hidden, int* outDim)
{
	CHECK_ARG(m);
	if (inDim) *inDim = m->inDim;
	if (hidden) *hidden = MLP_H;
	if (outDim) *outDim = m->outDim;
	return B3H_OK;
}
