// CPU oracle of the candidate-ranking MLP (TEST INFRASTRUCTURE ONLY): the
// forward pass, loss 0.5 * mean((y - t)^2) and its gradient, in double
// precision.  With bf16 != 0 every operand of the three matrix products is
// first rounded to bf16 (round to nearest even), as the GPU's MFMA path
// rounds them; with bf16 == 0 the arithmetic is plain double, which the
// finite-difference test of the gradient uses.
//
// Parameter layout (float): W1 [h x din], b1 [h], W2 [h x h], b2 [h],
// W3 [h], b3 [1].  GELU(z) = z/2 (1 + erf(z / sqrt 2)).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double roundBf16(double v, bool on) {
  if (!on) return v;
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

double gelu(double z) { return 0.5 * z * (1.0 + std::erf(z / std::sqrt(2.0))); }
double geluGrad(double z) {
  return 0.5 * (1.0 + std::erf(z / std::sqrt(2.0))) + z * std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
}

struct Fwd {
  std::vector<double> z1, h1, z2, h2, y;
};

void forward(const float* P, int din, int h, const float* X, int64_t n, bool bf, Fwd& F) {
  const float* W1 = P;
  const float* b1 = W1 + (size_t)h * din;
  const float* W2 = b1 + h;
  const float* b2 = W2 + (size_t)h * h;
  const float* W3 = b2 + h;
  const float* b3 = W3 + h;
  F.z1.assign((size_t)n * h, 0.0);
  F.h1.assign((size_t)n * h, 0.0);
  F.z2.assign((size_t)n * h, 0.0);
  F.h2.assign((size_t)n * h, 0.0);
  F.y.assign(n, 0.0);
  for (int64_t m = 0; m < n; m++) {
    for (int j = 0; j < h; j++) {
      double s = 0.0;
      for (int k = 0; k < din; k++) s += roundBf16(X[m * din + k], bf) * roundBf16(W1[(size_t)j * din + k], bf);
      const double z = s + b1[j];
      F.z1[m * h + j] = z;
      F.h1[m * h + j] = gelu(z);
    }
    for (int j = 0; j < h; j++) {
      double s = 0.0;
      for (int k = 0; k < h; k++) s += roundBf16(F.h1[m * h + k], bf) * roundBf16(W2[(size_t)j * h + k], bf);
      const double z = s + b2[j];
      F.z2[m * h + j] = z;
      F.h2[m * h + j] = gelu(z);
    }
    double s = 0.0;
    for (int k = 0; k < h; k++) s += F.h2[m * h + k] * W3[k];
    F.y[m] = s + b3[0];
  }
}

}  // namespace

extern "C" {

void oref_mlp_forward(const float* P, int din, int h, const float* X, int64_t n, int bf16, double* y) {
  Fwd F;
  forward(P, din, h, X, n, bf16 != 0, F);
  for (int64_t m = 0; m < n; m++) y[m] = F.y[m];
}

// returns the loss; grad (double, parameter layout) receives its gradient
double oref_mlp_loss_grad(const float* P, int din, int h, const float* X, const float* t, int64_t n, int bf16,
                          double* grad) {
  const bool bf = bf16 != 0;
  Fwd F;
  forward(P, din, h, X, n, bf, F);
  const float* W2 = P + (size_t)h * din + h;
  const float* W3 = W2 + (size_t)h * h + h;
  double* gW1 = grad;
  double* gb1 = gW1 + (size_t)h * din;
  double* gW2 = gb1 + h;
  double* gb2 = gW2 + (size_t)h * h;
  double* gW3 = gb2 + h;
  double* gb3 = gW3 + h;
  std::memset(grad, 0, sizeof(double) * ((size_t)h * din + h + (size_t)h * h + h + h + 1));
  double loss = 0.0;
  std::vector<double> dz2(h), dz1(h);
  for (int64_t m = 0; m < n; m++) {
    const double d = F.y[m] - t[m];
    loss += d * d;
    const double dy = d / (double)n;
    for (int k = 0; k < h; k++) gW3[k] += dy * F.h2[m * h + k];
    gb3[0] += dy;
    for (int j = 0; j < h; j++) {
      dz2[j] = dy * W3[j] * geluGrad(F.z2[m * h + j]);
      gb2[j] += dz2[j];
    }
    for (int j = 0; j < h; j++)
      for (int k = 0; k < h; k++) gW2[(size_t)j * h + k] += roundBf16(dz2[j], bf) * roundBf16(F.h1[m * h + k], bf);
    for (int k = 0; k < h; k++) {
      double s = 0.0;
      for (int j = 0; j < h; j++) s += roundBf16(dz2[j], bf) * roundBf16(W2[(size_t)j * h + k], bf);
      dz1[k] = s * geluGrad(F.z1[m * h + k]);
      gb1[k] += dz1[k];
    }
    for (int j = 0; j < h; j++)
      for (int k = 0; k < din; k++) gW1[(size_t)j * din + k] += roundBf16(dz1[j], bf) * roundBf16(X[m * din + k], bf);
  }
  return 0.5 * loss / (double)n;
}

}  // extern "C"
