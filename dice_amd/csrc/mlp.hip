// Candidate-ranking MLP: bf16 MFMA GEMM, GELU, loss, backward, Adam.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

namespace shci {

namespace {

constexpr int kThreads = 256;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short toBf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// 64 x 64 output tile per workgroup of four wavefronts (2 x 2 of 32 x 32);
// K advances 16 at a time through LDS; each wavefront issues 2 x 2
// v_mfma_f32_16x16x16bf16 per step.  Operand layout of that instruction:
// lane l holds A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16],
// i < 4, and receives D[4 (l / 16) + i][l % 16].
__global__ void __launch_bounds__(kThreads) kGemmBf16(int M, int N, int K, const float* __restrict__ A, int64_t sAm,
                                                      int64_t sAk, const float* __restrict__ B, int64_t sBk,
                                                      int64_t sBn, float* __restrict__ C, int64_t ldc, float beta) {
  __shared__ unsigned short sA[64][16 + 4];
  __shared__ unsigned short sB[16][64 + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid * 4 + q;
      const int am = e >> 4, ak = e & 15;
      const int gm = m0 + am, gk = k0 + ak;
      sA[am][ak] = toBf16((gm < M && gk < K) ? A[(int64_t)gm * sAm + (int64_t)gk * sAk] : 0.f);
      const int bk = e >> 6, bn = e & 63;
      const int gk2 = k0 + bk, gn = n0 + bn;
      sB[bk][bn] = toBf16((gk2 < K && gn < N) ? B[(int64_t)gk2 * sBk + (int64_t)gn * sBn] : 0.f);
    }
    __syncthreads();
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; i++) {
      bf16x4 a;
#pragma unroll
      for (int t = 0; t < 4; t++) a[t] = (short)sA[wr + i * 16 + r][kq + t];
#pragma unroll
      for (int j = 0; j < 2; j++) {
        bf16x4 b;
#pragma unroll
        for (int t = 0; t < 4; t++) b[t] = (short)sB[kq + t][wc + j * 16 + r];
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int gm = m0 + wr + i * 16 + rq + t, gn = n0 + wc + j * 16 + col;
        if (gm < M && gn < N) {
          float* c = C + (int64_t)gm * ldc + gn;
          *c = beta == 0.f ? acc[i][j][t] : acc[i][j][t] + beta * *c;
        }
      }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float geluGrad(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Z = C + b (in place), H = gelu(Z)
__global__ void kBiasGelu(float* Z, const float* b, int64_t n, int h, float* H) {
  const int64_t total = n * h;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const float z = Z[q] + b[q % h];
    Z[q] = z;
    H[q] = gelu(z);
  }
}

// y[m] = H[m,:] . w + b (one wavefront per row, fixed-order lane sums)
__global__ void kOutLayer(const float* H, const float* w, const float* b, int64_t n, int h, float* y) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t m = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave; m < n; m += (int64_t)gridDim.x * (blockDim.x >> 6)) {
    float acc = 0.f;
    for (int j = lane; j < h; j += 64) acc = fmaf(H[m * h + j], w[j], acc);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) y[m] = acc + b[0];
  }
}

// dy = (y - t) / n ; red[blockIdx] = partial sums of (y - t)^2
__global__ void kLossGrad(const float* y, const float* t, int64_t n, float* dy, float* red) {
  __shared__ float sred[kThreads / 64];
  float acc = 0.f;
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    const float d = y[m] - t[m];
    dy[m] = d / (float)n;
    acc = fmaf(d, d, acc);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < kThreads / 64; w++) s += sred[w];
    red[blockIdx.x] = s;
  }
}

__global__ void kLossFinal(const float* red, int nb, int64_t n, float* loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int b = 0; b < nb; b++) s += red[b];
    loss[0] = 0.5f * s / (float)n;
  }
}

// out[j] = sum_m a[m] * M[m*h + j] (a == nullptr: plain column sums), one
// block per column, fixed-order reduction
__global__ void kColReduce(const float* a, const float* Mx, int64_t n, int h, float* out) {
  __shared__ float sred[kThreads / 64];
  const int j = blockIdx.x;
  float acc = 0.f;
  for (int64_t m = threadIdx.x; m < n; m += blockDim.x) acc += (a ? a[m] : 1.f) * Mx[m * h + j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < kThreads / 64; w++) s += sred[w];
    out[j] = s;
  }
}

// dZ2[m,j] = dy[m] * w3[j] * gelu'(Z2[m,j])
__global__ void kOuterGelu(const float* dy, const float* w3, const float* Z, int64_t n, int h, float* dZ) {
  const int64_t total = n * h;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x)
    dZ[q] = dy[q / h] * w3[q % h] * geluGrad(Z[q]);
}

// dZ *= gelu'(Z)
__global__ void kGeluBack(float* dZ, const float* Z, int64_t total) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x)
    dZ[q] *= geluGrad(Z[q]);
}

__global__ void kAdam(float* p, const float* g, float* m, float* v, int64_t P, float lr, float b1, float b2, float eps,
                      float c1, float c2) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < P; q += (int64_t)gridDim.x * blockDim.x) {
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  }
}

// out rows = X rows idx[r] (and t[idx[r]])
__global__ void kGatherRows(const float* X, const float* t, const int64_t* idx, int64_t m, int din, float* bx,
                            float* bt) {
  const int64_t total = m * din;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = q / din;
    const int f = (int)(q - r * din);
    bx[q] = X[idx[r] * din + f];
    if (f == 0) bt[r] = t[idx[r]];
  }
}

unsigned gridOf(int64_t total, int64_t cap = 4096) {
  int64_t g = (total + kThreads - 1) / kThreads;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

}  // namespace

MlpLayout mlpLayout(int din, int hidden) {
  MlpLayout L;
  L.din = din;
  L.hidden = hidden;
  size_t o = 0;
  L.w1 = o;
  o += (size_t)hidden * din;
  L.b1 = o;
  o += hidden;
  L.w2 = o;
  o += (size_t)hidden * hidden;
  L.b2 = o;
  o += hidden;
  L.w3 = o;
  o += hidden;
  L.b3 = o;
  o += 1;
  L.total = o;
  return L;
}

void launchGemmBf16(int M, int N, int K, const float* A, int64_t sAm, int64_t sAk, const float* B, int64_t sBk,
                    int64_t sBn, float* C, int64_t ldc, float beta, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  kGemmBf16<<<grid, kThreads, 0, s>>>(M, N, K, A, sAm, sAk, B, sBk, sBn, C, ldc, beta);
  HIP_CHECK(hipGetLastError());
}

void mlpForward(const float* P, const MlpLayout& L, const float* X, int64_t n, MlpActs& A, hipStream_t s) {
  const int h = L.hidden, din = L.din;
  if (n > (int64_t)1 << 30) throw HipError("mlpForward: batch too large");
  const size_t act = (size_t)std::max<int64_t>(n, 1) * h * sizeof(float);
  A.z1.reserve(act);
  A.h1.reserve(act);
  A.z2.reserve(act);
  A.h2.reserve(act);
  A.y.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(float));
  // Z1 = X W1^T ; B(k, j) = W1[j * din + k]
  launchGemmBf16((int)n, h, din, X, din, 1, P + L.w1, 1, din, A.z1.as<float>(), h, 0.f, s);
  kBiasGelu<<<gridOf(n * h), kThreads, 0, s>>>(A.z1.as<float>(), P + L.b1, n, h, A.h1.as<float>());
  launchGemmBf16((int)n, h, h, A.h1.as<float>(), h, 1, P + L.w2, 1, h, A.z2.as<float>(), h, 0.f, s);
  kBiasGelu<<<gridOf(n * h), kThreads, 0, s>>>(A.z2.as<float>(), P + L.b2, n, h, A.h2.as<float>());
  kOutLayer<<<gridOf(n * 64), kThreads, 0, s>>>(A.h2.as<float>(), P + L.w3, P + L.b3, n, h, A.y.as<float>());
  HIP_CHECK(hipGetLastError());
}

void mlpBackward(const float* P, const MlpLayout& L, const float* X, const float* t, int64_t n, MlpActs& A,
                 float* G, float* lossOut, hipStream_t s) {
  const int h = L.hidden, din = L.din;
  const size_t act = (size_t)std::max<int64_t>(n, 1) * h * sizeof(float);
  A.dz1.reserve(act);
  A.dz2.reserve(act);
  A.dy.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(float));
  const unsigned nb = gridOf(n, 1024);
  A.red.reserve((size_t)nb * sizeof(float));
  kLossGrad<<<nb, kThreads, 0, s>>>(A.y.as<float>(), t, n, A.dy.as<float>(), A.red.as<float>());
  kLossFinal<<<1, 64, 0, s>>>(A.red.as<float>(), (int)nb, n, lossOut);
  // output layer
  kColReduce<<<h, kThreads, 0, s>>>(A.dy.as<float>(), A.h2.as<float>(), n, h, G + L.w3);
  kColReduce<<<1, kThreads, 0, s>>>(nullptr, A.dy.as<float>(), n, 1, G + L.b3);
  // second hidden layer
  kOuterGelu<<<gridOf(n * h), kThreads, 0, s>>>(A.dy.as<float>(), P + L.w3, A.z2.as<float>(), n, h,
                                                 A.dz2.as<float>());
  // dW2[o][i] = sum_m dZ2[m][o] H1[m][i]
  launchGemmBf16(h, h, (int)n, A.dz2.as<float>(), 1, h, A.h1.as<float>(), h, 1, G + L.w2, h, 0.f, s);
  kColReduce<<<h, kThreads, 0, s>>>(nullptr, A.dz2.as<float>(), n, h, G + L.b2);
  // first hidden layer: dH1 = dZ2 W2, dZ1 = dH1 * gelu'(Z1)
  launchGemmBf16((int)n, h, h, A.dz2.as<float>(), h, 1, P + L.w2, h, 1, A.dz1.as<float>(), h, 0.f, s);
  kGeluBack<<<gridOf(n * h), kThreads, 0, s>>>(A.dz1.as<float>(), A.z1.as<float>(), n * h);
  // dW1[o][k] = sum_m dZ1[m][o] X[m][k]
  launchGemmBf16(h, din, (int)n, A.dz1.as<float>(), 1, h, X, din, 1, G + L.w1, din, 0.f, s);
  kColReduce<<<h, kThreads, 0, s>>>(nullptr, A.dz1.as<float>(), n, h, G + L.b1);
  HIP_CHECK(hipGetLastError());
}

void launchAdam(float* params, const float* grad, float* m, float* v, size_t P, float lr, float beta1, float beta2,
                float eps, int t, hipStream_t s) {
  const float c1 = 1.f - std::pow(beta1, (float)t), c2 = 1.f - std::pow(beta2, (float)t);
  kAdam<<<gridOf((int64_t)P), kThreads, 0, s>>>(params, grad, m, v, (int64_t)P, lr, beta1, beta2, eps, c1, c2);
  HIP_CHECK(hipGetLastError());
}

void MlpModel::init(int din, int hidden, uint64_t seed, hipStream_t s) {
  L = mlpLayout(din, hidden);
  std::vector<float> h(L.total, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](size_t off, int fanOut, int fanIn) {
    const float a = std::sqrt(6.f / (float)(fanIn + fanOut));
    std::uniform_real_distribution<float> u(-a, a);
    for (size_t q = 0; q < (size_t)fanOut * fanIn; q++) h[off + q] = u(rng);
  };
  fill(L.w1, hidden, din);
  fill(L.w2, hidden, hidden);
  fill(L.w3, 1, hidden);
  P.reserve(L.total * sizeof(float));
  G.reserve(L.total * sizeof(float));
  M.reserve(L.total * sizeof(float));
  V.reserve(L.total * sizeof(float));
  loss.reserve(64);
  HIP_CHECK(hipMemcpyAsync(P.p, h.data(), L.total * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemsetAsync(M.p, 0, L.total * sizeof(float), s));
  HIP_CHECK(hipMemsetAsync(V.p, 0, L.total * sizeof(float), s));
  HIP_CHECK(hipStreamSynchronize(s));
  steps = 0;
}

void MlpModel::setParams(const float* host, hipStream_t s) {
  HIP_CHECK(hipMemcpyAsync(P.p, host, L.total * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

void MlpModel::getParams(float* host, hipStream_t s) const {
  HIP_CHECK(hipMemcpyAsync(host, P.p, L.total * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

void MlpModel::predict(const float* X, int64_t n, float* y, hipStream_t s) {
  mlpForward(P.as<float>(), L, X, n, A, s);
  if (n > 0) HIP_CHECK(hipMemcpyAsync(y, A.y.p, n * sizeof(float), hipMemcpyDeviceToDevice, s));
}

float MlpModel::lossGrad(const float* X, const float* t, int64_t n, hipStream_t s) {
  mlpForward(P.as<float>(), L, X, n, A, s);
  mlpBackward(P.as<float>(), L, X, t, n, A, G.as<float>(), loss.as<float>(), s);
  float h = 0.f;
  HIP_CHECK(hipMemcpyAsync(&h, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return h;
}

void MlpModel::train(const float* X, const float* t, int64_t n, int epochs, int batch, float lr, uint64_t seed,
                     float& firstLoss, float& lastLoss, hipStream_t s) {
  if (n <= 0 || batch <= 0) throw HipError("MlpModel::train: empty training set");
  const int64_t m = std::min<int64_t>(batch, n);
  bx.reserve((size_t)m * L.din * sizeof(float));
  bt.reserve((size_t)m * sizeof(float));
  idx.reserve((size_t)n * sizeof(int64_t));
  std::vector<int64_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 rng(seed);
  firstLoss = lastLoss = 0.f;
  bool first = true;
  for (int e = 0; e < epochs; e++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    HIP_CHECK(hipMemcpyAsync(idx.p, perm.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, s));
    for (int64_t b0 = 0; b0 + m <= n; b0 += m) {
      kGatherRows<<<gridOf(m * L.din), kThreads, 0, s>>>(X, t, idx.as<int64_t>() + b0, m, L.din, bx.as<float>(),
                                                          bt.as<float>());
      mlpForward(P.as<float>(), L, bx.as<float>(), m, A, s);
      mlpBackward(P.as<float>(), L, bx.as<float>(), bt.as<float>(), m, A, G.as<float>(), loss.as<float>(), s);
      launchAdam(P.as<float>(), G.as<float>(), M.as<float>(), V.as<float>(), L.total, lr, 0.9f, 0.999f, 1e-8f,
                 ++steps, s);
      if (first) {
        HIP_CHECK(hipMemcpyAsync(&firstLoss, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
        first = false;
      }
    }
    // the host permutation buffer is reused by the next epoch: wait for this one
    HIP_CHECK(hipStreamSynchronize(s));
  }
  HIP_CHECK(hipMemcpyAsync(&lastLoss, loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace shci
