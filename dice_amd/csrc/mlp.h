// Candidate-ranking MLP of the learned screening form: din -> hidden ->
// hidden -> 1, GELU after the two hidden layers, FP32 master parameters,
// matrix products on bf16 MFMA with FP32 accumulation (operands rounded to
// bf16 on their way into LDS).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "prims.h"

namespace shci {

// flat parameter layout: W1 [hidden x din], b1 [hidden], W2 [hidden x hidden],
// b2 [hidden], W3 [hidden], b3 [1]
struct MlpLayout {
  int din = 0, hidden = 0;
  size_t w1 = 0, b1 = 0, w2 = 0, b2 = 0, w3 = 0, b3 = 0, total = 0;
};
MlpLayout mlpLayout(int din, int hidden);

// C[m*ldc + n] = sum_k A(m,k) B(k,n) (+ beta * C) for m < M, n < N with
// A(m,k) = A[m*sAm + k*sAk], B(k,n) = B[k*sBk + n*sBn]
void launchGemmBf16(int M, int N, int K, const float* A, int64_t sAm, int64_t sAk, const float* B, int64_t sBk,
                    int64_t sBn, float* C, int64_t ldc, float beta, hipStream_t s);

// activations kept for the backward pass (each n x hidden)
struct MlpActs {
  DevMem z1, h1, z2, h2, y, dz1, dz2, dy, red;
};

// y = MLP(X) for the n rows of X (n x din, row-major)
void mlpForward(const float* params, const MlpLayout& L, const float* X, int64_t n, MlpActs& A, hipStream_t s);
// after mlpForward on the same X: loss = 0.5 * mean((y - t)^2) into lossOut
// (device, one float) and its gradient into grad (layout of params)
void mlpBackward(const float* params, const MlpLayout& L, const float* X, const float* t, int64_t n, MlpActs& A,
                 float* grad, float* lossOut, hipStream_t s);
// Adam (FP32 moments), step t >= 1
void launchAdam(float* params, const float* grad, float* m, float* v, size_t P, float lr, float beta1, float beta2,
                float eps, int t, hipStream_t s);

// Parameters, Adam state and work buffers of one network on one stream.
struct MlpModel {
  MlpLayout L;
  DevMem P, G, M, V, loss, bx, bt, idx;
  MlpActs A;
  int steps = 0;
  // Xavier-uniform weights, zero biases (host RNG seeded with seed)
  void init(int din, int hidden, uint64_t seed, hipStream_t s);
  void setParams(const float* host, hipStream_t s);
  void getParams(float* host, hipStream_t s) const;
  // scores for n rows of X (device) into y (device)
  void predict(const float* X, int64_t n, float* y, hipStream_t s);
  // loss and gradient (device G) for the n rows of X with targets t; returns the loss
  float lossGrad(const float* X, const float* t, int64_t n, hipStream_t s);
  // Adam over shuffled minibatches; loss of the first and of the last minibatch
  void train(const float* X, const float* t, int64_t n, int epochs, int batch, float lr, uint64_t seed,
             float& firstLoss, float& lastLoss, hipStream_t s);
};

}  // namespace shci
