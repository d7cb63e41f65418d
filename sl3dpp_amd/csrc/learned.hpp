// The learned warm start of the step's fixed point (no counterpart in the
// reference): a multilayer perceptron applied to every memory node, trained
// and evaluated with BF16 matrix-core kernels accumulating in FP32
// (learned.hip), and the moves between the spectral iterate (device layout)
// and the network's per-node rows.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace sl3amd {

// Layers of widths w0 -> w1 -> ... -> wL: hidden layers GELU (the exact erf
// form), the last one linear. Weights W_k (w_k x w_{k-1}, row-major) and
// biases b_k are held in FP32 (the optimiser's copy) and in BF16 (the matrix
// cores' operands); products accumulate in FP32.
struct Epilogue;

class Mlp {
 public:
  Mlp(std::vector<int> widths, int maxRows, uint64_t seed);
  ~Mlp();
  Mlp(const Mlp&) = delete;
  Mlp& operator=(const Mlp&) = delete;
  int Layers() const { return static_cast<int>(widths_.size()) - 1; }
  const std::vector<int>& Widths() const { return widths_; }
  int MaxRows() const { return maxRows_; }
  // y (rows x wL, FP32) from x (rows x w0, FP32), any number of rows.
  void Forward(long long rows, const float* x, float* y, hipStream_t s);
  // The gradients of the loss mean((y - target)^2) over rows x wL into the
  // gradient buffer (rows <= MaxRows). Returns the loss when `loss` is set
  // (a host synchronisation).
  void Backward(int rows, const float* x, const float* target, double* loss, hipStream_t s);
  // One Adam update of every weight and bias from the gradient buffer.
  void Adam(double lr, double beta1, double beta2, double eps, hipStream_t s);
  // Host copies: "weight:k", "bias:k", "grad_weight:k", "grad_bias:k"
  // (k = 1..L); Import sets an FP32 weight or bias (and its BF16 copy).
  std::vector<float> Export(const std::string& name, hipStream_t s) const;
  void Import(const std::string& name, const float* src, size_t count, hipStream_t s);
  long long Steps() const { return steps_; }
  // The products' slice depth: 0 picks per shape by timing (the default),
  // 32 or 128 fixes it. Chunks() lists (kind, M, N, K, depth) per tuned
  // shape.
  void SetChunk(int chunk);
  std::vector<float> Chunks() const;

 private:
  template <bool KA, bool KB, int MODE>
  void Product(int kind, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
               int ldb, const Epilogue& e, hipStream_t s);
  void ForwardChunk(int rows, const float* x, float* y, hipStream_t s);
  size_t Offset(const std::string& name, size_t* count) const;
  std::vector<int> widths_;
  int maxRows_;
  std::vector<size_t> wOff_, bOff_;
  size_t paramCount_ = 0;
  DeviceArray<float> params_, grads_, m1_, m2_;
  DeviceArray<uint16_t> paramsBf16_;
  DeviceArray<uint16_t> xb_;
  std::vector<DeviceArray<float>> z_;      // hidden pre-activations
  std::vector<DeviceArray<uint16_t>> h_;   // hidden activations (BF16)
  DeviceArray<float> y_;
  DeviceArray<uint16_t> d0_, d1_;          // output gradients, ping-pong
  DeviceArray<float> partials_;
  std::vector<float> partialsHost_;
  long long steps_ = 0;
  int fixedChunk_ = 0;
  std::map<std::array<int, 4>, int> tiles_;
  hipEvent_t tuneEvents_[2] = {nullptr, nullptr};
};

// Column means and scales (the population standard deviation, floored at
// `floor` times the largest one) of a rows x cols FP32 matrix.
void ColumnMoments(long long rows, int cols, const float* x, float* mean, float* scale,
                   float floor, hipStream_t s);

// x = (x - mean) / scale per column (inverse: x scale + mean).
void Standardise(float* x, long long rows, int cols, const float* mean, const float* scale,
                 bool inverse, hipStream_t s);

// A linear combination of up to four device vectors.
constexpr int kSpectralTermsMax = 8;
struct SpectralTerms {
  int count = 0;
  double c[kSpectralTermsMax] = {};
  const double* src[kSpectralTermsMax] = {};
};
void LaunchCombineTerms(const SpectralTerms& t, double* dst, size_t n, hipStream_t s);

// out[node * ld + col0 + 2 e + c] = ((spec - base)[row e][node].c - mean) /
// scale, c = re, im, for the rows e of `rowOffsets` (the double offset of
// each selected row in the device layout; a row holds `nodes` complex
// values). base, mean, scale may be null.
void LaunchSpectralGather(const double* spec, const double* base, const long long* rowOffsets,
                          int rows, int nodes, const float* mean, const float* scale,
                          float* out, int ld, int col0, hipStream_t s);
// spec[row e][node].c += in[node * ld + 2 e + c] * scale + mean.
void LaunchSpectralScatterAdd(const float* in, int ld, const long long* rowOffsets, int rows,
                              int nodes, const float* mean, const float* scale, double* spec,
                              hipStream_t s);

}  // namespace sl3amd
