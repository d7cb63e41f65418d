// The learned warm start's device code: a multilayer perceptron on the
// BF16 matrix cores (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation) with its
// training kernels (mean-squared loss, backpropagation, Adam), and the moves
// between the step's spectral iterate and the network's per-node rows. There
// is no counterpart in the reference: this is the warm start handed to
// OpenStep::Converge's `start` (src/Viscoelastic3D.cpp:1916-1921).
#include "learned.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>

#include "kernels.hpp"

namespace sl3amd {
// What a product's epilogue writes (Bf16GemmKernel's MODE selects).
struct Epilogue {
  const float* bias = nullptr;   // per column (kBiasGelu, kBias)
  float* outF = nullptr;         // FP32 output (kStoreF32, kBias; kBiasGelu: z)
  uint16_t* outH = nullptr;      // BF16 output (kBiasGelu: gelu(z); kGeluGrad)
  const float* zPrev = nullptr;  // pre-activations whose GELU derivative scales (kGeluGrad)
  int ldc = 0;
};

namespace {

typedef short Bf16x4 __attribute__((ext_vector_type(4)));
typedef float F32x4 __attribute__((ext_vector_type(4)));

// BF16 by round-to-nearest-even on the dropped 16 bits (NaN kept quiet).
__host__ __device__ inline uint16_t ToBf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

__device__ inline float FromBf16(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }

// GELU, the exact form z Phi(z), and its derivative Phi(z) + z phi(z).
__device__ inline float Gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float GeluGrad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C (M x N) = sum_k A(i, k) B(k, j) in 64 x 64 workgroup tiles of four
// waves, CHUNK-deep slices (32 or 128) staged through LDS as [row][k] BF16
// (rows of the tile: i for A, j for B, padded by 8 so that 16-byte stores
// stay aligned). Each wave holds a 32 x 32 quarter as 2 x 2 products of
// 16 x 16 x 16. Both depths sum every element's products in the same order
// (16 deep at a time, k ascending), so they give the same bits; deeper
// slices take fewer load round trips and barriers, shallower ones less LDS.
// Mlp's tuner picks per shape.
// Operand layouts of v_mfma_f32_16x16x16bf16_1k (lane l): A(i = l % 16,
// k = 4 (l / 16) + 0..3), B(k = 4 (l / 16) + 0..3, j = l % 16), and the
// accumulator C(i = 4 (l / 16) + 0..3, j = l % 16).
constexpr int kTile = 64;

enum { kStoreF32 = 0, kBiasGelu = 1, kBias = 2, kGeluGrad = 3 };


// One thread's eight elements of a staged tile (see StageTile).
template <bool KCONTIG, int CHUNK>
__device__ inline void StageEight(const uint16_t* __restrict__ p, int ld, int outer, int K, int o0,
                                  int k0, bool vec, uint16_t* T, int idx) {
  constexpr int kRow = CHUNK + 8;
  if (KCONTIG) {
    const int r = idx / (CHUNK / 8), ks = (idx % (CHUNK / 8)) * 8;
    const int o = o0 + r, k = k0 + ks;
    uint16_t* dst = T + r * kRow + ks;
    if (vec && o < outer && k + 8 <= K) {
      *reinterpret_cast<uint4*>(dst) =
          *reinterpret_cast<const uint4*>(p + static_cast<size_t>(o) * ld + k);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        dst[e] = (o < outer && k + e < K) ? p[static_cast<size_t>(o) * ld + k + e] : 0;
    }
  } else {
    const int kq = idx / (kTile / 8), rs = (idx % (kTile / 8)) * 8;
    const int k = k0 + kq, o = o0 + rs;
    uint16_t v[8];
    if (vec && k < K && o + 8 <= outer) {
      const uint4 q = *reinterpret_cast<const uint4*>(p + static_cast<size_t>(k) * ld + o);
      std::memcpy(v, &q, 16);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        v[e] = (k < K && o + e < outer) ? p[static_cast<size_t>(k) * ld + o + e] : 0;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) T[(rs + e) * kRow + kq] = v[e];
  }
}

// Stages rows o0..o0+63, depth k0..k0+CHUNK-1 of an operand into
// T[row][k], 8 elements per thread and pass (256 threads). KCONTIG: element
// (o, k) at p[o ld + k]; otherwise at p[k ld + o]. vec: ld a multiple of 8
// and p 16-byte aligned (8-element loads). Out-of-range elements are zeros.
template <bool KCONTIG, int CHUNK>
__device__ inline void StageTile(const uint16_t* __restrict__ p, int ld, int outer, int K, int o0,
                                 int k0, bool vec, uint16_t* T, int tid) {
#pragma unroll
  for (int pass = 0; pass < kTile * CHUNK / (8 * 256); ++pass)
    StageEight<KCONTIG, CHUNK>(p, ld, outer, K, o0, k0, vec, T, tid + pass * 256);
}

template <bool KA, bool KB, int MODE, int CHUNK>
__global__ __launch_bounds__(256) void Bf16GemmKernel(int M, int N, int K,
                                                      const uint16_t* __restrict__ A, int lda,
                                                      bool vecA, const uint16_t* __restrict__ B,
                                                      int ldb, bool vecB, Epilogue e) {
  constexpr int kRow = CHUNK + 8;
  __shared__ __attribute__((aligned(16))) uint16_t As[kTile * kRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile * kRow];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  F32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = F32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += CHUNK) {
    StageTile<KA, CHUNK>(A, lda, M, K, i0, k0, vecA, As, tid);
    StageTile<KB, CHUNK>(B, ldb, N, K, j0, k0, vecB, Bs, tid);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < CHUNK; kk += 16) {
      Bf16x4 a[2], b[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[t] = *reinterpret_cast<const Bf16x4*>(
            &As[(wi + 16 * t + (lane & 15)) * kRow + kk + 4 * (lane >> 4)]);
        b[t] = *reinterpret_cast<const Bf16x4*>(
            &Bs[(wj + 16 * t + (lane & 15)) * kRow + kk + 4 * (lane >> 4)]);
      }
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi + 16 * ti + 4 * (lane >> 4) + r;
        const int j = j0 + wj + 16 * tj + (lane & 15);
        if (i >= M || j >= N) continue;
        const size_t o = static_cast<size_t>(i) * e.ldc + j;
        const float v = acc[ti][tj][r];
        if (MODE == kStoreF32) {
          e.outF[o] = v;
        } else if (MODE == kBiasGelu) {
          const float z = v + e.bias[j];
          e.outF[o] = z;
          e.outH[o] = ToBf16(Gelu(z));
        } else if (MODE == kBias) {
          e.outF[o] = v + e.bias[j];
        } else {
          e.outH[o] = ToBf16(v * GeluGrad(e.zPrev[o]));
        }
      }
}

__global__ __launch_bounds__(256) void ToBf16Kernel(const float* __restrict__ x,
                                                    uint16_t* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull)
    y[i] = ToBf16(x[i]);
}

// dy = scale (y - t) in BF16, and per-workgroup sums of (y - t)^2.
constexpr int kReduceBlocks = 512;
__global__ __launch_bounds__(256) void MseKernel(const float* __restrict__ y,
                                                 const float* __restrict__ t, size_t n,
                                                 float scale, uint16_t* __restrict__ dy,
                                                 float* __restrict__ partials) {
  __shared__ float red[256];
  float s = 0.f;
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
    const float d = y[i] - t[i];
    s += d * d;
    dy[i] = ToBf16(scale * d);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Column sums over row chunks (deterministic: fixed chunks, then the chunks
// summed in order). SQ: sums of (x - center[j])^2.
template <class T, bool SQ>
__global__ __launch_bounds__(256) void ColumnPartialKernel(const T* __restrict__ x,
                                                           long long rows, int cols,
                                                           long long chunk,
                                                           const float* __restrict__ center,
                                                           float* __restrict__ partials) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= cols) return;
  const long long r0 = blockIdx.y * chunk;
  const long long r1 = r0 + chunk < rows ? r0 + chunk : rows;
  const float c = SQ ? center[j] : 0.f;
  float s = 0.f;
  for (long long r = r0; r < r1; ++r) {
    float v;
    if constexpr (sizeof(T) == 2) {
      v = FromBf16(x[r * cols + j]);
    } else {
      v = x[r * cols + j];
    }
    if (SQ) {
      v -= c;
      s += v * v;
    } else {
      s += v;
    }
  }
  partials[blockIdx.y * static_cast<long long>(cols) + j] = s;
}

__global__ __launch_bounds__(256) void ColumnFinishKernel(const float* __restrict__ partials,
                                                          int chunks, int cols, float factor,
                                                          bool root, float* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int c = 0; c < chunks; ++c) s += partials[static_cast<long long>(c) * cols + j];
  s *= factor;
  out[j] = root ? static_cast<float>(std::sqrt(s)) : static_cast<float>(s);
}

// Adam (Kingma and Ba, 2015) with bias corrections c1 = 1 - beta1^t, c2 =
// 1 - beta2^t; the BF16 copy refreshed.
__global__ __launch_bounds__(256) void AdamKernel(float* __restrict__ w, float* __restrict__ m,
                                                  float* __restrict__ v,
                                                  const float* __restrict__ g,
                                                  uint16_t* __restrict__ wb, size_t n, float lr,
                                                  float b1, float b2, float eps, float c1,
                                                  float c2) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = ToBf16(wi);
  }
}

__global__ __launch_bounds__(256) void SpectralGatherKernel(
    const double* __restrict__ spec, const double* __restrict__ base,
    const long long* __restrict__ rowOff, int rows, int nodes, const float* __restrict__ mean,
    const float* __restrict__ scale, float* __restrict__ out, int ld, int col0) {
  const long long idx = blockIdx.x * 256ll + threadIdx.x;
  if (idx >= static_cast<long long>(rows) * nodes) return;
  const int e = static_cast<int>(idx / nodes), node = static_cast<int>(idx % nodes);
  const long long off = rowOff[e] + 2ll * node;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    double v = spec[off + c];
    if (base != nullptr) v -= base[off + c];
    const int f = 2 * e + c;
    if (mean != nullptr) v = (v - mean[f]) / scale[f];
    out[static_cast<long long>(node) * ld + col0 + f] = static_cast<float>(v);
  }
}

__global__ __launch_bounds__(256) void SpectralScatterAddKernel(
    const float* __restrict__ in, int ld, const long long* __restrict__ rowOff, int rows,
    int nodes, const float* __restrict__ mean, const float* __restrict__ scale,
    double* __restrict__ spec) {
  const long long idx = blockIdx.x * 256ll + threadIdx.x;
  if (idx >= static_cast<long long>(rows) * nodes) return;
  const int e = static_cast<int>(idx / nodes), node = static_cast<int>(idx % nodes);
  const long long off = rowOff[e] + 2ll * node;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int f = 2 * e + c;
    double v = in[static_cast<long long>(node) * ld + f];
    if (mean != nullptr) v = v * scale[f] + mean[f];
    spec[off + c] += v;
  }
}

// x = (x - mean) / scale per column, or x scale + mean (inverse).
__global__ __launch_bounds__(256) void StandardiseKernel(float* __restrict__ x, long long rows,
                                                         int cols,
                                                         const float* __restrict__ mean,
                                                         const float* __restrict__ scale,
                                                         bool inverse) {
  const long long n = rows * cols;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += gridDim.x * 256ll) {
    const int j = static_cast<int>(i % cols);
    x[i] = inverse ? x[i] * scale[j] + mean[j] : (x[i] - mean[j]) / scale[j];
  }
}

// dst = sum_j c_j src_j (up to eight terms; dst may be one of the sources:
// every element reads its sources before its store).
__global__ __launch_bounds__(256) void CombineTermsKernel(SpectralTerms t, double* dst,
                                                          size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
    double v = 0.0;
    for (int j = 0; j < t.count; ++j) v += t.c[j] * t.src[j][i];
    dst[i] = v;
  }
}

int GridOf(size_t n) {
  return static_cast<int>(std::min<size_t>((n + 255) / 256, 4096));
}

bool Aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <bool KA, bool KB, int MODE>
void Gemm(int chunk, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
          const Epilogue& e, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  const bool vecA = lda % 8 == 0 && Aligned16(A);
  const bool vecB = ldb % 8 == 0 && Aligned16(B);
  const dim3 grid(CeilDiv(N, kTile), CeilDiv(M, kTile));
  if (chunk == 128) {
    hipLaunchKernelGGL((Bf16GemmKernel<KA, KB, MODE, 128>), grid, dim3(256), 0, s, M, N, K, A,
                       lda, vecA, B, ldb, vecB, e);
  } else {
    hipLaunchKernelGGL((Bf16GemmKernel<KA, KB, MODE, 32>), grid, dim3(256), 0, s, M, N, K, A,
                       lda, vecA, B, ldb, vecB, e);
  }
  SL3_HIP(hipGetLastError());
}

// Column sums of a rows x cols matrix (FP32 or BF16) into out (times
// factor; the root of it when root).
template <class T, bool SQ>
void ColumnSums(const T* x, long long rows, int cols, const float* center, float factor,
                bool root, float* partials, size_t partialCapacity, float* out,
                hipStream_t s) {
  long long chunk = 256;
  long long chunks = (rows + chunk - 1) / chunk;
  while (chunks * cols > static_cast<long long>(partialCapacity)) {
    chunk *= 2;
    chunks = (rows + chunk - 1) / chunk;
  }
  if (chunks < 1) chunks = 1;
  hipLaunchKernelGGL((ColumnPartialKernel<T, SQ>), dim3(CeilDiv(cols, 256), chunks), dim3(256),
                     0, s, x, rows, cols, chunk, center, partials);
  SL3_HIP(hipGetLastError());
  hipLaunchKernelGGL(ColumnFinishKernel, dim3(CeilDiv(cols, 256)), dim3(256), 0, s, partials,
                     static_cast<int>(chunks), cols, factor, root, out);
  SL3_HIP(hipGetLastError());
}

size_t Round8(size_t n) { return (n + 7) & ~static_cast<size_t>(7); }

}  // namespace

// The slice depth of a GEMM shape: fixed, or on the shape's first use the
// faster of the two, timed with device events over ten launches each after
// one untimed launch (every launch writes the same outputs from the
// same inputs, so the trial launches change nothing).
template <bool KA, bool KB, int MODE>
void Mlp::Product(int kind, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
                  int ldb, const Epilogue& e, hipStream_t s) {
  int tile = fixedChunk_;
  if (tile == 0) {
    const std::array<int, 4> key{kind, M, N, K};
    auto it = tiles_.find(key);
    if (it == tiles_.end()) {
      float best = 0.0f;
      int choice = 32;
      for (int trial : {32, 128}) {
        Gemm<KA, KB, MODE>(trial, M, N, K, A, lda, B, ldb, e, s);
        SL3_HIP(hipEventRecord(tuneEvents_[0], s));
        for (int r = 0; r < 10; ++r) Gemm<KA, KB, MODE>(trial, M, N, K, A, lda, B, ldb, e, s);
        SL3_HIP(hipEventRecord(tuneEvents_[1], s));
        SL3_HIP(hipEventSynchronize(tuneEvents_[1]));
        float ms = 0.0f;
        SL3_HIP(hipEventElapsedTime(&ms, tuneEvents_[0], tuneEvents_[1]));
        if (trial == 32 || ms < best) {
          best = ms;
          choice = trial;
        }
      }
      it = tiles_.emplace(key, choice).first;
    }
    tile = it->second;
  }
  Gemm<KA, KB, MODE>(tile, M, N, K, A, lda, B, ldb, e, s);
}

Mlp::~Mlp() {
  for (auto& ev : tuneEvents_)
    if (ev != nullptr) (void)hipEventDestroy(ev);
}

void Mlp::SetChunk(int chunk) {
  if (chunk != 0 && chunk != 32 && chunk != 128) throw ArgumentError("chunk must be 0, 32 or 128");
  fixedChunk_ = chunk;
}

std::vector<float> Mlp::Chunks() const {
  std::vector<float> out;
  for (const auto& [key, tile] : tiles_) {
    for (int v : key) out.push_back(static_cast<float>(v));
    out.push_back(static_cast<float>(tile));
  }
  return out;
}

Mlp::Mlp(std::vector<int> widths, int maxRows, uint64_t seed)
    : widths_(std::move(widths)), maxRows_(maxRows) {
  if (widths_.size() < 2) throw ArgumentError("an MLP needs at least one layer");
  for (int w : widths_)
    if (w < 1) throw ArgumentError("MLP widths must be positive");
  if (maxRows_ < 1) throw ArgumentError("max_rows must be positive");
  const int L = Layers();
  wOff_.assign(L + 1, 0);
  bOff_.assign(L + 1, 0);
  size_t cur = 0;
  for (int k = 1; k <= L; ++k) {
    wOff_[k] = cur;
    cur = Round8(cur + static_cast<size_t>(widths_[k]) * widths_[k - 1]);
    bOff_[k] = cur;
    cur = Round8(cur + widths_[k]);
  }
  paramCount_ = cur;
  // torch.nn.Linear's default initialisation: U(-1/sqrt(in), 1/sqrt(in))
  // for weights and biases; padding entries zero
  std::vector<float> host(paramCount_, 0.f);
  std::mt19937_64 rng(seed);
  for (int k = 1; k <= L; ++k) {
    const float bound = 1.0f / std::sqrt(static_cast<float>(widths_[k - 1]));
    std::uniform_real_distribution<float> u(-bound, bound);
    const size_t nw = static_cast<size_t>(widths_[k]) * widths_[k - 1];
    for (size_t i = 0; i < nw; ++i) host[wOff_[k] + i] = u(rng);
    for (int i = 0; i < widths_[k]; ++i) host[bOff_[k] + i] = u(rng);
  }
  hipStream_t s = Stream();
  params_.Upload(host.data(), paramCount_, s);
  grads_.Resize(paramCount_);
  m1_.Resize(paramCount_);
  m2_.Resize(paramCount_);
  SL3_HIP(hipMemsetAsync(grads_.data(), 0, paramCount_ * sizeof(float), s));
  SL3_HIP(hipMemsetAsync(m1_.data(), 0, paramCount_ * sizeof(float), s));
  SL3_HIP(hipMemsetAsync(m2_.data(), 0, paramCount_ * sizeof(float), s));
  paramsBf16_.Resize(paramCount_);
  hipLaunchKernelGGL(ToBf16Kernel, dim3(GridOf(paramCount_)), dim3(256), 0, s, params_.data(),
                     paramsBf16_.data(), paramCount_);
  SL3_HIP(hipGetLastError());
  const size_t R = static_cast<size_t>(maxRows_);
  int maxW = 0;
  for (int k = 1; k <= L; ++k) maxW = std::max(maxW, widths_[k]);
  xb_.Resize(R * widths_[0]);
  z_.resize(L);
  h_.resize(L);
  for (int k = 1; k < L; ++k) {
    z_[k].Resize(R * widths_[k]);
    h_[k].Resize(R * widths_[k]);
  }
  y_.Resize(R * widths_[L]);
  d0_.Resize(R * maxW);
  d1_.Resize(R * maxW);
  const size_t chunks = (R + 255) / 256;
  partials_.Resize(std::max<size_t>(kReduceBlocks, chunks * maxW));
  partialsHost_.resize(kReduceBlocks);
  for (auto& ev : tuneEvents_) SL3_HIP(hipEventCreate(&ev));
  SL3_HIP(hipStreamSynchronize(s));
}

void Mlp::ForwardChunk(int rows, const float* x, float* y, hipStream_t s) {
  const int L = Layers();
  const size_t nx = static_cast<size_t>(rows) * widths_[0];
  hipLaunchKernelGGL(ToBf16Kernel, dim3(GridOf(nx)), dim3(256), 0, s, x, xb_.data(), nx);
  SL3_HIP(hipGetLastError());
  const uint16_t* in = xb_.data();
  for (int k = 1; k <= L; ++k) {
    Epilogue e;
    e.bias = params_.data() + bOff_[k];
    e.ldc = widths_[k];
    const uint16_t* W = paramsBf16_.data() + wOff_[k];
    if (k < L) {
      e.outF = z_[k].data();
      e.outH = h_[k].data();
      Product<true, true, kBiasGelu>(kBiasGelu, rows, widths_[k], widths_[k - 1], in,
                                     widths_[k - 1], W, widths_[k - 1], e, s);
      in = h_[k].data();
    } else {
      e.outF = y;
      Product<true, true, kBias>(kBias, rows, widths_[k], widths_[k - 1], in, widths_[k - 1],
                                 W, widths_[k - 1], e, s);
    }
  }
}

void Mlp::Forward(long long rows, const float* x, float* y, hipStream_t s) {
  ProfileScope prof("mlp_forward", s);
  const int L = Layers();
  for (long long r0 = 0; r0 < rows; r0 += maxRows_) {
    const int n = static_cast<int>(std::min<long long>(maxRows_, rows - r0));
    ForwardChunk(n, x + r0 * widths_[0], y + r0 * widths_[L], s);
  }
}

void Mlp::Backward(int rows, const float* x, const float* target, double* loss, hipStream_t s) {
  if (rows < 1 || rows > maxRows_) throw ArgumentError("backward: rows must be in 1..max_rows");
  ProfileScope prof("mlp_backward", s);
  const int L = Layers();
  ForwardChunk(rows, x, y_.data(), s);
  const size_t n = static_cast<size_t>(rows) * widths_[L];
  const float scale = 2.0f / static_cast<float>(n);
  hipLaunchKernelGGL(MseKernel, dim3(kReduceBlocks), dim3(256), 0, s, y_.data(), target, n,
                     scale, d0_.data(), partials_.data());
  SL3_HIP(hipGetLastError());
  if (loss != nullptr) {
    SL3_HIP(hipMemcpyAsync(partialsHost_.data(), partials_.data(), kReduceBlocks * sizeof(float),
                           hipMemcpyDeviceToHost, s));
    SL3_HIP(hipStreamSynchronize(s));
    double sum = 0.0;
    for (int b = 0; b < kReduceBlocks; ++b) sum += partialsHost_[b];
    *loss = sum / static_cast<double>(n);
  }
  uint16_t* cur = d0_.data();
  uint16_t* other = d1_.data();
  for (int k = L; k >= 1; --k) {
    const uint16_t* in = k == 1 ? xb_.data() : h_[k - 1].data();
    // dW_k = dZ_k^T H_{k-1} (w_k x w_{k-1})
    Epilogue ew;
    ew.outF = grads_.data() + wOff_[k];
    ew.ldc = widths_[k - 1];
    Product<false, false, kStoreF32>(kStoreF32, widths_[k], widths_[k - 1], rows, cur,
                                     widths_[k], in, widths_[k - 1], ew, s);
    // db_k = column sums of dZ_k
    ColumnSums<uint16_t, false>(cur, rows, widths_[k], nullptr, 1.0f, false, partials_.data(),
                                partials_.size(), grads_.data() + bOff_[k], s);
    if (k > 1) {
      // dZ_{k-1} = (dZ_k W_k) * gelu'(Z_{k-1})
      Epilogue ed;
      ed.outH = other;
      ed.zPrev = z_[k - 1].data();
      ed.ldc = widths_[k - 1];
      Product<true, false, kGeluGrad>(kGeluGrad, rows, widths_[k - 1], widths_[k], cur,
                                      widths_[k], paramsBf16_.data() + wOff_[k], widths_[k - 1],
                                      ed, s);
      std::swap(cur, other);
    }
  }
}

void Mlp::Adam(double lr, double beta1, double beta2, double eps, hipStream_t s) {
  if (!(lr > 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) ||
      !(eps > 0.0)) {
    throw ArgumentError("adam: lr > 0, 0 <= beta < 1, eps > 0");
  }
  ProfileScope prof("mlp_adam", s);
  ++steps_;
  const float c1 = static_cast<float>(1.0 - std::pow(beta1, static_cast<double>(steps_)));
  const float c2 = static_cast<float>(1.0 - std::pow(beta2, static_cast<double>(steps_)));
  hipLaunchKernelGGL(AdamKernel, dim3(GridOf(paramCount_)), dim3(256), 0, s, params_.data(),
                     m1_.data(), m2_.data(), grads_.data(), paramsBf16_.data(), paramCount_,
                     static_cast<float>(lr), static_cast<float>(beta1),
                     static_cast<float>(beta2), static_cast<float>(eps), c1, c2);
  SL3_HIP(hipGetLastError());
}

size_t Mlp::Offset(const std::string& name, size_t* count) const {
  const auto colon = name.find(':');
  if (colon == std::string::npos) throw ArgumentError("unknown MLP field: " + name);
  const std::string field = name.substr(0, colon);
  int k = 0;
  try {
    k = std::stoi(name.substr(colon + 1));
  } catch (const std::exception&) {
    throw ArgumentError("unknown MLP field: " + name);
  }
  if (k < 1 || k > Layers()) throw ArgumentError("MLP layer out of range: " + name);
  if (field == "weight" || field == "grad_weight") {
    *count = static_cast<size_t>(widths_[k]) * widths_[k - 1];
    return wOff_[k];
  }
  if (field == "bias" || field == "grad_bias") {
    *count = widths_[k];
    return bOff_[k];
  }
  throw ArgumentError("unknown MLP field: " + name);
}

std::vector<float> Mlp::Export(const std::string& name, hipStream_t s) const {
  size_t count = 0;
  const size_t off = Offset(name, &count);
  const float* src = (name.rfind("grad_", 0) == 0 ? grads_.data() : params_.data()) + off;
  std::vector<float> out(count);
  SL3_HIP(hipMemcpyAsync(out.data(), src, count * sizeof(float), hipMemcpyDeviceToHost, s));
  SL3_HIP(hipStreamSynchronize(s));
  return out;
}

void Mlp::Import(const std::string& name, const float* src, size_t count, hipStream_t s) {
  if (name.rfind("grad_", 0) == 0) throw ArgumentError("gradients are not imported");
  size_t expect = 0;
  const size_t off = Offset(name, &expect);
  if (count != expect) throw ArgumentError("import size mismatch for " + name);
  SL3_HIP(hipMemcpyAsync(params_.data() + off, src, count * sizeof(float),
                         hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(ToBf16Kernel, dim3(GridOf(count)), dim3(256), 0, s, params_.data() + off,
                     paramsBf16_.data() + off, count);
  SL3_HIP(hipGetLastError());
  SL3_HIP(hipStreamSynchronize(s));
}

void ColumnMoments(long long rows, int cols, const float* x, float* mean, float* scale,
                   float floor, hipStream_t s) {
  if (rows < 1 || cols < 1) throw ArgumentError("column moments: empty matrix");
  const size_t cap = static_cast<size_t>(cols) * 1024;
  DeviceArray<float> partials(cap);
  ColumnSums<float, false>(x, rows, cols, nullptr, 1.0f / static_cast<float>(rows), false,
                           partials.data(), cap, mean, s);
  ColumnSums<float, true>(x, rows, cols, mean, 1.0f / static_cast<float>(rows), true,
                          partials.data(), cap, scale, s);
  std::vector<float> host(cols);
  SL3_HIP(hipMemcpyAsync(host.data(), scale, cols * sizeof(float), hipMemcpyDeviceToHost, s));
  SL3_HIP(hipStreamSynchronize(s));
  float top = 0.f;
  for (float v : host) top = std::max(top, v);
  const float least = top > 0.f ? floor * top : 1.0f;
  for (float& v : host) v = std::max(v, least);
  SL3_HIP(hipMemcpyAsync(scale, host.data(), cols * sizeof(float), hipMemcpyHostToDevice, s));
  SL3_HIP(hipStreamSynchronize(s));
}

void Standardise(float* x, long long rows, int cols, const float* mean, const float* scale,
                 bool inverse, hipStream_t s) {
  const size_t n = static_cast<size_t>(rows) * cols;
  if (n == 0) return;
  hipLaunchKernelGGL(StandardiseKernel, dim3(GridOf(n)), dim3(256), 0, s, x, rows, cols, mean,
                     scale, inverse);
  SL3_HIP(hipGetLastError());
}

void LaunchCombineTerms(const SpectralTerms& t, double* dst, size_t n, hipStream_t s) {
  if (t.count < 1 || t.count > kSpectralTermsMax) throw ArgumentError("combine: 1 to 8 terms");
  if (n == 0) return;
  ProfileScope prof("combine", s);
  hipLaunchKernelGGL(CombineTermsKernel, dim3(GridOf(n)), dim3(256), 0, s, t, dst, n);
  SL3_HIP(hipGetLastError());
}

void LaunchSpectralGather(const double* spec, const double* base, const long long* rowOffsets,
                          int rows, int nodes, const float* mean, const float* scale,
                          float* out, int ld, int col0, hipStream_t s) {
  const long long n = static_cast<long long>(rows) * nodes;
  if (n == 0) return;
  ProfileScope prof("spectral_gather", s);
  hipLaunchKernelGGL(SpectralGatherKernel, dim3(CeilDiv(n, 256)), dim3(256), 0, s, spec, base,
                     rowOffsets, rows, nodes, mean, scale, out, ld, col0);
  SL3_HIP(hipGetLastError());
}

void LaunchSpectralScatterAdd(const float* in, int ld, const long long* rowOffsets, int rows,
                              int nodes, const float* mean, const float* scale, double* spec,
                              hipStream_t s) {
  const long long n = static_cast<long long>(rows) * nodes;
  if (n == 0) return;
  ProfileScope prof("spectral_scatter", s);
  hipLaunchKernelGGL(SpectralScatterAddKernel, dim3(CeilDiv(n, 256)), dim3(256), 0, s, in, ld,
                     rowOffsets, rows, nodes, mean, scale, spec);
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd

using namespace sl3amd;

struct sl3_mlp {
  std::unique_ptr<Mlp> mlp;
};

extern "C" {

int sl3_mlp_create(int layers, const int* widths, int max_rows, uint64_t seed, sl3_mlp** out) {
  return Guard([&] {
    if (out == nullptr || widths == nullptr || layers < 1)
      throw ArgumentError("sl3_mlp_create: layers >= 1, widths and out required");
    *out = nullptr;
    auto h = std::make_unique<sl3_mlp>();
    h->mlp = std::make_unique<Mlp>(std::vector<int>(widths, widths + layers + 1), max_rows, seed);
    *out = h.release();
  });
}

int sl3_mlp_destroy(sl3_mlp* mlp) {
  return Guard([&] { delete mlp; });
}

int sl3_mlp_forward(sl3_mlp* mlp, int64_t rows, const float* x_dev, float* y_dev) {
  return Guard([&] {
    if (mlp == nullptr || rows < 0 || (rows > 0 && (x_dev == nullptr || y_dev == nullptr)))
      throw ArgumentError("sl3_mlp_forward: bad arguments");
    mlp->mlp->Forward(rows, x_dev, y_dev, Stream());
  });
}

int sl3_mlp_backward(sl3_mlp* mlp, int rows, const float* x_dev, const float* target_dev,
                     double* loss) {
  return Guard([&] {
    if (mlp == nullptr || x_dev == nullptr || target_dev == nullptr)
      throw ArgumentError("sl3_mlp_backward: bad arguments");
    mlp->mlp->Backward(rows, x_dev, target_dev, loss, Stream());
  });
}

int sl3_mlp_adam(sl3_mlp* mlp, double lr, double beta1, double beta2, double eps) {
  return Guard([&] {
    if (mlp == nullptr) throw ArgumentError("sl3_mlp_adam: null network");
    mlp->mlp->Adam(lr, beta1, beta2, eps, Stream());
  });
}

int sl3_mlp_set_chunk(sl3_mlp* mlp, int chunk) {
  return Guard([&] {
    if (mlp == nullptr) throw ArgumentError("sl3_mlp_set_chunk: null network");
    mlp->mlp->SetChunk(chunk);
  });
}

int sl3_mlp_export(sl3_mlp* mlp, const char* name, float* dst, int64_t* count) {
  return Guard([&] {
    if (mlp == nullptr || name == nullptr || count == nullptr)
      throw ArgumentError("sl3_mlp_export: bad arguments");
    const auto v = std::string(name) == "chunks" ? mlp->mlp->Chunks()
                                                : mlp->mlp->Export(name, Stream());
    if (dst != nullptr) {
      if (*count < static_cast<int64_t>(v.size())) throw ArgumentError("export buffer too small");
      std::copy(v.begin(), v.end(), dst);
    }
    *count = static_cast<int64_t>(v.size());
  });
}

int sl3_mlp_import(sl3_mlp* mlp, const char* name, const float* src, int64_t count) {
  return Guard([&] {
    if (mlp == nullptr || name == nullptr || src == nullptr || count < 0)
      throw ArgumentError("sl3_mlp_import: bad arguments");
    mlp->mlp->Import(name, src, static_cast<size_t>(count), Stream());
  });
}

int sl3_column_moments(int64_t rows, int cols, const float* x_dev, float floor, float* mean_dev,
                       float* scale_dev) {
  return Guard([&] {
    if (x_dev == nullptr || mean_dev == nullptr || scale_dev == nullptr || !(floor >= 0.f))
      throw ArgumentError("sl3_column_moments: bad arguments");
    ColumnMoments(rows, cols, x_dev, mean_dev, scale_dev, floor, Stream());
  });
}

int sl3_column_standardise(int64_t rows, int cols, float* x_dev, const float* mean_dev,
                           const float* scale_dev, int inverse) {
  return Guard([&] {
    if (rows < 0 || cols < 1 || (rows > 0 && x_dev == nullptr) || mean_dev == nullptr ||
        scale_dev == nullptr)
      throw ArgumentError("sl3_column_standardise: bad arguments");
    Standardise(x_dev, rows, cols, mean_dev, scale_dev, inverse != 0, Stream());
  });
}

}  // extern "C"
