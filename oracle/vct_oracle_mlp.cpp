// vct_oracle_mlp.cpp - CPU oracle of the learned upsampler's network.
//
// TEST INFRASTRUCTURE ONLY (see vct_oracle.cpp).  The network has no
// counterpart in the reference; this file states it in FP64:
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3
//   loss = mean over n samples and 3 channels of (a * (y - t))^2
// with the analytic gradient of the loss w.r.t. every parameter.  With
// `bf16` set it rounds to BF16 exactly where the device kernels do (weights,
// inputs, hidden activations, and the back-propagated dY, dZ2, dZ1 that feed
// matrix products), accumulating in double; the FP32 parameter layout is the
// one of include/vct.h.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

constexpr int kIn = 34, kHid = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kOut * kHid, kParams = kOffB3 + kOut;

double bf(double v, bool on) {
  if (!on) return v;
  float f = (float)v;
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return v;
  const uint32_t lower = x & 0xffffu, upper = x >> 16;
  uint32_t r = upper;
  if (lower > 0x8000u || (lower == 0x8000u && (upper & 1u))) r = upper + 1u;
  x = r << 16;
  memcpy(&f, &x, 4);
  return (double)f;
}

struct Net {
  std::vector<double> W1, b1, W2, b2, W3, b3;
  Net(const float* p, bool q) {
    W1.resize(kHid * kIn);
    b1.resize(kHid);
    W2.resize(kHid * kHid);
    b2.resize(kHid);
    W3.resize(kOut * kHid);
    b3.resize(kOut);
    for (int i = 0; i < kHid * kIn; ++i) W1[i] = bf(p[kOffW1 + i], q);
    for (int i = 0; i < kHid; ++i) b1[i] = p[kOffB1 + i];
    for (int i = 0; i < kHid * kHid; ++i) W2[i] = bf(p[kOffW2 + i], q);
    for (int i = 0; i < kHid; ++i) b2[i] = p[kOffB2 + i];
    for (int i = 0; i < kOut * kHid; ++i) W3[i] = bf(p[kOffW3 + i], q);
    for (int i = 0; i < kOut; ++i) b3[i] = p[kOffB3 + i];
  }
};

}  // namespace

extern "C" {

// y (n x 3) of the network for features x (n x 34).
void orc_mlp_forward(const float* params, const float* x, int n, int bf16, double* y) {
  const bool q = bf16 != 0;
  Net N(params, q);
  for (int s = 0; s < n; ++s) {
    double in[kIn], h1[kHid], h2[kHid];
    for (int k = 0; k < kIn; ++k) in[k] = bf(x[(size_t)s * kIn + k], q);
    for (int j = 0; j < kHid; ++j) {
      double a = 0;
      for (int k = 0; k < kIn; ++k) a += N.W1[j * kIn + k] * in[k];
      a += N.b1[j];
      h1[j] = bf(a > 0 ? a : 0, q);
    }
    for (int j = 0; j < kHid; ++j) {
      double a = 0;
      for (int k = 0; k < kHid; ++k) a += N.W2[j * kHid + k] * h1[k];
      a += N.b2[j];
      h2[j] = bf(a > 0 ? a : 0, q);
    }
    for (int o = 0; o < kOut; ++o) {
      double a = 0;
      for (int k = 0; k < kHid; ++k) a += N.W3[o * kHid + k] * h2[k];
      y[(size_t)s * kOut + o] = a + N.b3[o];
    }
  }
}

// Loss and its gradient (kParams doubles) over n samples.
double orc_mlp_loss_grad(const float* params, const float* x, const float* alb, const float* t, int n, int bf16,
                         double* grad) {
  const bool q = bf16 != 0;
  Net N(params, q);
  memset(grad, 0, sizeof(double) * kParams);
  const double scale = 2.0 / (3.0 * n);
  double loss = 0;
  for (int s = 0; s < n; ++s) {
    double in[kIn], z1[kHid], h1[kHid], z2[kHid], h2[kHid], y[kOut];
    for (int k = 0; k < kIn; ++k) in[k] = bf(x[(size_t)s * kIn + k], q);
    for (int j = 0; j < kHid; ++j) {
      double a = 0;
      for (int k = 0; k < kIn; ++k) a += N.W1[j * kIn + k] * in[k];
      z1[j] = a + N.b1[j];
      h1[j] = bf(z1[j] > 0 ? z1[j] : 0, q);
    }
    for (int j = 0; j < kHid; ++j) {
      double a = 0;
      for (int k = 0; k < kHid; ++k) a += N.W2[j * kHid + k] * h1[k];
      z2[j] = a + N.b2[j];
      h2[j] = bf(z2[j] > 0 ? z2[j] : 0, q);
    }
    double dy[kOut];
    for (int o = 0; o < kOut; ++o) {
      double a = 0;
      for (int k = 0; k < kHid; ++k) a += N.W3[o * kHid + k] * h2[k];
      y[o] = a + N.b3[o];
      const double al = alb[(size_t)s * kOut + o];
      const double e = al * (y[o] - (double)t[(size_t)s * kOut + o]);
      loss += e * e;
      dy[o] = scale * al * e;
    }
    // back-propagation
    double dyq[kOut];
    for (int o = 0; o < kOut; ++o) {
      grad[kOffB3 + o] += dy[o];
      dyq[o] = bf(dy[o], q);
      for (int k = 0; k < kHid; ++k) grad[kOffW3 + o * kHid + k] += dyq[o] * h2[k];
    }
    double dz2[kHid], dz2q[kHid];
    for (int k = 0; k < kHid; ++k) {
      double a = 0;
      for (int o = 0; o < kOut; ++o) a += N.W3[o * kHid + k] * dyq[o];
      dz2[k] = z2[k] > 0 ? a : 0;
      dz2q[k] = bf(dz2[k], q);
      grad[kOffB2 + k] += dz2[k];
    }
    for (int j = 0; j < kHid; ++j)
      for (int k = 0; k < kHid; ++k) grad[kOffW2 + j * kHid + k] += dz2q[j] * h1[k];
    for (int k = 0; k < kHid; ++k) {
      double a = 0;
      for (int j = 0; j < kHid; ++j) a += N.W2[j * kHid + k] * dz2q[j];
      const double dz1 = z1[k] > 0 ? a : 0;
      grad[kOffB1 + k] += dz1;
      const double dz1q = bf(dz1, q);
      for (int i = 0; i < kIn; ++i) grad[kOffW1 + k * kIn + i] += dz1q * in[i];
    }
  }
  return loss / (3.0 * n);
}

}  // extern "C"
