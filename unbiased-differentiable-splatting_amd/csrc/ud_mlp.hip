// ud_mlp.hip -- per-tile rate network for Poisson-rate sampling (gfx950).
//
// A 32 -> 64 -> 64 -> 1 perceptron (ReLU hidden layers) maps 32 features of a
// 16x16 pixel tile to the logarithm of its sampling-rate multiplier.  Weights
// and activations enter the matrix cores as bfloat16 and accumulate in FP32
// (v_mfma_f32_16x16x16_bf16); biases, the output layer's dot product, the
// loss and the optimiser run in FP32.  Forward, backward and Adam are all
// kernels here, so the network trains on the GPU with no library GEMM.
//
// Parameter block (FP32, 6337 floats): W1 [32][64], b1 [64], W2 [64][64],
// b2 [64], W3 [64], b3 [1] (row-major, input index first).
#include "ud_device.h"
#include "../../include/udsplat.h"

using namespace ud;

namespace {

constexpr int kIn = 32, kHid = 64;
constexpr int kOffW1 = 0, kOffB1 = kIn * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid, kParams = kOffB3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> bfloat16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__global__ void k_to_bf16(long n, const float* __restrict__ src, uint16_t* __restrict__ dst)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = f2bf(src[t]);
}

// C (M x N) = A (M x K) * B (K x N) on the matrix cores: one wavefront per
// 16 x 16 output tile, K in steps of 16.  A(i, k) = A[i * sai + k * sak],
// B(k, j) = B[k * sbk + j * sbj] (bf16), so transposed operands are strides.
// Lane l holds A(i0 + l % 16, k0 + 4 (l / 16) + v), B(k0 + 4 (l / 16) + v,
// j0 + l % 16) and gets D(i0 + 4 (l / 16) + v, j0 + l % 16), v = 0..3.
// Epilogue: + bias[j] (optional), ReLU (optional), FP32 and/or bf16 store.
__global__ __launch_bounds__(64) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, long sai,
                                                  long sak, const uint16_t* __restrict__ B, long sbk, long sbj,
                                                  const float* __restrict__ bias, int relu, float* __restrict__ C,
                                                  uint16_t* __restrict__ Cb, long ldc)
{
    const int lane = threadIdx.x;
    const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
    const int r = lane & 15, g = lane >> 4;
    const int ia = i0 + r, jb = j0 + r;
    f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            int k = k0 + 4 * g + v;
            a[v] = (ia < M && k < K) ? (short)A[(long)ia * sai + (long)k * sak] : (short)0;
            b[v] = (jb < N && k < K) ? (short)B[(long)k * sbk + (long)jb * sbj] : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        int i = i0 + 4 * g + v, j = j0 + r;
        if (i >= M || j >= N) continue;
        float x = acc[v];
        if (bias) x = x + bias[j];
        if (relu) x = fmaxf(x, 0.0f);
        if (C) C[(long)i * ldc + j] = x;
        if (Cb) Cb[(long)i * ldc + j] = f2bf(x);
    }
}

// output layer: y[t] = sum_k H2[t][k] W3[k] (bf16 products, FP32 sum in k order) + b3
__global__ __launch_bounds__(256) void k_out_layer(int T, const uint16_t* __restrict__ H2b,
                                                   const uint16_t* __restrict__ W3b, const float* __restrict__ b3,
                                                   float* __restrict__ y)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    float s = 0.0f;
    for (int k = 0; k < kHid; k++) s = s + bf2f(H2b[(long)t * kHid + k]) * bf2f(W3b[k]);
    y[t] = s + b3[0];
}

// weighted mean-square loss and its gradient dy = 2 w (y - target) / sum(w);
// one workgroup, fixed-order tree reduction (deterministic)
__global__ __launch_bounds__(1024) void k_loss(int T, const float* __restrict__ y, const float* __restrict__ target,
                                               const float* __restrict__ weight, float* __restrict__ dy,
                                               float* __restrict__ lossOut)
{
    __shared__ float sw[1024], sl[1024];
    float w = 0.0f, l = 0.0f;
    for (int t = threadIdx.x; t < T; t += 1024) {
        float wt = weight ? weight[t] : 1.0f;
        float e = y[t] - target[t];
        w = w + wt;
        l = l + wt * (e * e);
    }
    sw[threadIdx.x] = w;
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sw[threadIdx.x] = sw[threadIdx.x] + sw[threadIdx.x + s];
            sl[threadIdx.x] = sl[threadIdx.x] + sl[threadIdx.x + s];
        }
        __syncthreads();
    }
    const float W = sw[0];
    const float inv = W > 0.0f ? 1.0f / W : 0.0f;
    if (threadIdx.x == 0 && lossOut) lossOut[0] = sl[0] * inv;
    for (int t = threadIdx.x; t < T; t += 1024) {
        float wt = weight ? weight[t] : 1.0f;
        dy[t] = (2.0f * wt) * (y[t] - target[t]) * inv;
    }
}

// dZ2[t][k] = dy[t] W3[k] where H2[t][k] > 0 (ReLU), FP32 and bf16
__global__ __launch_bounds__(256) void k_back_out(int T, const float* __restrict__ dy, const uint16_t* __restrict__ H2b,
                                                  const uint16_t* __restrict__ W3b, float* __restrict__ dZ,
                                                  uint16_t* __restrict__ dZb)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)T * kHid) return;
    int k = (int)(e % kHid);
    float x = bf2f(H2b[e]) > 0.0f ? dy[e / kHid] * bf2f(W3b[k]) : 0.0f;
    dZ[e] = x;
    dZb[e] = f2bf(x);
}

// apply the ReLU mask of H (bf16) to dH in place; bf16 copy
__global__ __launch_bounds__(256) void k_relu_back(long n, const uint16_t* __restrict__ Hb, float* __restrict__ dH,
                                                   uint16_t* __restrict__ dHb)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float x = bf2f(Hb[e]) > 0.0f ? dH[e] : 0.0f;
    dH[e] = x;
    dHb[e] = f2bf(x);
}

// out[c] = sum_t M[t][c] (or, with V, sum_t M[t] V[t][c]) in fixed order: one
// workgroup per column, strided partial sums then a tree
__global__ __launch_bounds__(256) void k_colsum(int T, int C, const float* __restrict__ Mx, const uint16_t* __restrict__ Vb,
                                                float* __restrict__ out)
{
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int t = threadIdx.x; t < T; t += 256) {
        float x = Vb ? Mx[t] * bf2f(Vb[(long)t * C + c]) : Mx[(long)t * C + c];
        s = s + x;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// Adam (Kingma & Ba 2015) with bias correction, FP32 master parameters
__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float gt = g[t];
    float mt = b1 * m[t] + (1.0f - b1) * gt;
    float vt = b2 * v[t] + (1.0f - b2) * (gt * gt);
    m[t] = mt;
    v[t] = vt;
    p[t] = p[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

struct MlpWs {
    uint16_t *Xb, *Pb, *H1b, *H2b, *dZ2b, *dZ1b;
    float *y, *dy, *dZ2, *dZ1;
};

inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t mlp_carve(void* ws, int T, MlpWs* w)
{
    size_t off = 0;
    char* base = (char*)ws;
    auto take = [&](size_t bytes) { size_t o = off; off = a256(off + bytes); return base ? (void*)(base + o) : nullptr; };
    MlpWs m;
    long t = T > 0 ? T : 1;
    m.Xb = (uint16_t*)take(2 * t * kIn);
    m.Pb = (uint16_t*)take(2 * kParams);
    m.H1b = (uint16_t*)take(2 * t * kHid);
    m.H2b = (uint16_t*)take(2 * t * kHid);
    m.dZ2b = (uint16_t*)take(2 * t * kHid);
    m.dZ1b = (uint16_t*)take(2 * t * kHid);
    m.y = (float*)take(4 * t);
    m.dy = (float*)take(4 * t);
    m.dZ2 = (float*)take(4 * t * kHid);
    m.dZ1 = (float*)take(4 * t * kHid);
    if (w) *w = m;
    return off;
}

inline unsigned g256(long n) { return (unsigned)((n + 255) / 256); }

void mlp_forward(int T, const float* X, const float* P, const MlpWs& w, float* y, hipStream_t st)
{
    k_to_bf16<<<g256((long)T * kIn), 256, 0, st>>>((long)T * kIn, X, w.Xb);
    k_to_bf16<<<g256(kParams), 256, 0, st>>>(kParams, P, w.Pb);
    dim3 grid(kHid / 16, (T + 15) / 16);
    // H1 = relu(X W1 + b1), H2 = relu(H1 W2 + b2)
    k_gemm_bf16<<<grid, 64, 0, st>>>(T, kHid, kIn, w.Xb, kIn, 1, w.Pb + kOffW1, kHid, 1, P + kOffB1, 1, nullptr,
                                     w.H1b, kHid);
    k_gemm_bf16<<<grid, 64, 0, st>>>(T, kHid, kHid, w.H1b, kHid, 1, w.Pb + kOffW2, kHid, 1, P + kOffB2, 1, nullptr,
                                     w.H2b, kHid);
    k_out_layer<<<g256(T), 256, 0, st>>>(T, w.H2b, w.Pb + kOffW3, P + kOffB3, y);
}

} // namespace

extern "C" {

int ude_mlp_param_count(void) { return kParams; }

size_t ude_mlp_workspace_bytes(int T) { return mlp_carve(nullptr, T, nullptr) + 256; }

int ude_mlp_forward(int T, const float* X, const float* params, float* y, void* ws, size_t wsBytes, hipStream_t st)
{
    if (T <= 0) return UDE_OK;
    if (wsBytes < ude_mlp_workspace_bytes(T)) return UDE_ERR_WORKSPACE;
    MlpWs w;
    mlp_carve(ws, T, &w);
    mlp_forward(T, X, params, w, y, st);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

int ude_mlp_loss_grad(int T, const float* X, const float* params, const float* target, const float* weight,
                      float* grad, float* loss, void* ws, size_t wsBytes, hipStream_t st)
{
    if (T <= 0) return UDE_ERR_ARG;
    if (wsBytes < ude_mlp_workspace_bytes(T)) return UDE_ERR_WORKSPACE;
    MlpWs w;
    mlp_carve(ws, T, &w);
    mlp_forward(T, X, params, w, w.y, st);
    k_loss<<<1, 1024, 0, st>>>(T, w.y, target, weight, w.dy, loss);
    // output layer: dW3 = H2^T dy, db3 = sum dy, dZ2 = dy W3^T masked by ReLU
    k_colsum<<<kHid, 256, 0, st>>>(T, kHid, w.dy, w.H2b, grad + kOffW3);
    k_colsum<<<1, 256, 0, st>>>(T, 1, w.dy, nullptr, grad + kOffB3);
    k_back_out<<<g256((long)T * kHid), 256, 0, st>>>(T, w.dy, w.H2b, w.Pb + kOffW3, w.dZ2, w.dZ2b);
    // layer 2: dW2 = H1^T dZ2 (64 x T x 64), db2 = colsum dZ2, dH1 = dZ2 W2^T
    dim3 gw(kHid / 16, kHid / 16);
    k_gemm_bf16<<<gw, 64, 0, st>>>(kHid, kHid, T, w.H1b, 1, kHid, w.dZ2b, kHid, 1, nullptr, 0, grad + kOffW2, nullptr,
                                   kHid);
    k_colsum<<<kHid, 256, 0, st>>>(T, kHid, w.dZ2, nullptr, grad + kOffB2);
    dim3 gh(kHid / 16, (T + 15) / 16);
    k_gemm_bf16<<<gh, 64, 0, st>>>(T, kHid, kHid, w.dZ2b, kHid, 1, w.Pb + kOffW2, 1, kHid, nullptr, 0, w.dZ1, nullptr,
                                   kHid);
    k_relu_back<<<g256((long)T * kHid), 256, 0, st>>>((long)T * kHid, w.H1b, w.dZ1, w.dZ1b);
    // layer 1: dW1 = X^T dZ1 (32 x T x 64), db1 = colsum dZ1
    dim3 g1(kHid / 16, kIn / 16);
    k_gemm_bf16<<<g1, 64, 0, st>>>(kIn, kHid, T, w.Xb, 1, kIn, w.dZ1b, kHid, 1, nullptr, 0, grad + kOffW1, nullptr,
                                   kHid);
    k_colsum<<<kHid, 256, 0, st>>>(T, kHid, w.dZ1, nullptr, grad + kOffB1);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

int ude_adam_step(int n, float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                  float eps, int step, hipStream_t st)
{
    if (n <= 0 || step <= 0) return UDE_ERR_ARG;
    float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    k_adam<<<g256(n), 256, 0, st>>>(n, params, grad, m, v, lr, beta1, beta2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? UDE_OK : UDE_ERR_LAUNCH;
}

} // extern "C"
