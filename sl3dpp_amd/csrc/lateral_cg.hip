// The vector work of the memory-dof conjugate gradients (lateral_cg.hpp):
// the weighted pairing of a response with a force, and the two update
// sweeps of an iteration. The operator applications reuse the step's
// kernels.
#include "lateral_cg.hpp"

#include <algorithm>

#include "common.hpp"

namespace sl3amd {

namespace {

constexpr int kThreads = 256;
constexpr int kRows = 64;  // tile rows (memory dofs)
constexpr int kCols = 16;  // tile columns
constexpr int kLdRow = kRows + 1;

// One tile: the force part (f - g) through LDS transposed, so both operands
// are read along their contiguous direction; w(m) x (f - g) summed per
// workgroup into partials[blockIdx.x].
__global__ void __launch_bounds__(kThreads)
    LateralPairKernel(const DegreeTask* tasks, const PairWork* work, const double* x,
                      const double* f, const double* g, ChannelTerms ch, double* partials) {
  __shared__ double tile[kCols * kLdRow];
  __shared__ double red[kThreads];
  const PairWork w = work[blockIdx.x];
  const DegreeTask t = tasks[w.task];
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kRows * kCols / kThreads; ++k) {
    const int idx = tid + k * kThreads;
    const int row = idx / kCols, col = idx % kCols;
    const int r = w.r0 + row, c = w.c0 + col;
    double v = 0.0;
    if (r < t.nm && c < t.cols) {
      const size_t at = static_cast<size_t>(t.fOff) + static_cast<size_t>(r) * t.colsPad + c;
      v = g ? f[at] - g[at] : f[at];
    }
    tile[col * kLdRow + row] = v;
  }
  __syncthreads();
  const bool channels = ch.amp != nullptr && t.unitOff >= 0;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kRows * kCols / kThreads; ++k) {
    const int idx = tid + k * kThreads;
    const int col = idx / kRows, row = idx % kRows;
    const int r = w.r0 + row, c = w.c0 + col;
    if (r < t.nm && c < t.cols) {
      double xv = x[static_cast<size_t>(t.fOff) + static_cast<size_t>(c) * t.nmPad + r];
      const int m = c % (t.l + 1), part = c / (t.l + 1);
      if (channels) {
        const size_t pair = static_cast<size_t>(t.l) * (t.l + 1) / 2 + m;
        double u = ch.units[t.dmOff + r] * ch.amp[2 * pair + part];
        if (ch.tideAmp) u += ch.tideUnits[t.dmOff + r] * ch.tideAmp[2 * pair + part];
        xv += ch.scale * u;
      }
      acc += (m == 0 ? 1.0 : 2.0) * xv * tile[col * kLdRow + row];
    }
  }
  red[tid] = acc;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) partials[blockIdx.x] = red[0];
}

// The partial sums in a fixed order: thread i takes i, i + 256, ..., then
// a tree over the threads.
__global__ void __launch_bounds__(kThreads) SumPartialsKernel(const double* partials, int n,
                                                              double* out) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) s += partials[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void CgAdvanceKernel(double* phi, double* y, double* rho, const double* pi,
                                const double* Gpi, const double* q, double alpha, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const double p = pi[i];
    phi[i] += alpha * p;
    y[i] += alpha * Gpi[i];
    rho[i] -= alpha * (p - q[i]);  // r -= alpha A p, A p = Pᵀ (pi - q)
  }
}

__global__ void CgDirectionKernel(double* pi, double* Gpi, const double* rho,
                                  const double* Grho, double beta, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    pi[i] = rho[i] + beta * pi[i];
    Gpi[i] = Grho[i] + beta * Gpi[i];
  }
}

int SweepBlocks(size_t n) {
  const size_t b = (n + kThreads - 1) / kThreads;
  return static_cast<int>(std::clamp<size_t>(b, 1, 8192));
}

}  // namespace

std::vector<PairWork> PairWorkList(const std::vector<DegreeTask>& tasks) {
  std::vector<PairWork> out;
  for (size_t ti = 0; ti < tasks.size(); ++ti) {
    const auto& t = tasks[ti];
    for (int r0 = 0; r0 < t.nm; r0 += kRows) {
      for (int c0 = 0; c0 < t.cols; c0 += kCols) {
        out.push_back(PairWork{static_cast<int>(ti), r0, c0, 0});
      }
    }
  }
  return out;
}

void LaunchLateralPair(const DegreeTask* tasks_dev, const PairWork* work_dev, int workCount,
                       const double* x, const double* f, const double* g,
                       const ChannelTerms& channels, double* partials, double* result,
                       hipStream_t stream) {
  ProfileScope prof("cg_pair", stream);
  if (workCount > 0) {
    hipLaunchKernelGGL(LateralPairKernel, dim3(workCount), dim3(kThreads), 0, stream,
                       tasks_dev, work_dev, x, f, g, channels, partials);
    SL3_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(SumPartialsKernel, dim3(1), dim3(kThreads), 0, stream, partials,
                     workCount, result);
  SL3_HIP(hipGetLastError());
}

void LaunchCgAdvance(double* phi, double* y, double* rho, const double* pi,
                     const double* Gpi, const double* q, double alpha, size_t n,
                     hipStream_t stream) {
  ProfileScope prof("cg_update", stream);
  hipLaunchKernelGGL(CgAdvanceKernel, dim3(SweepBlocks(n)), dim3(kThreads), 0, stream, phi,
                     y, rho, pi, Gpi, q, alpha, n);
  SL3_HIP(hipGetLastError());
}

void LaunchCgDirection(double* pi, double* Gpi, const double* rho, const double* Grho,
                       double beta, size_t n, hipStream_t stream) {
  ProfileScope prof("cg_update", stream);
  hipLaunchKernelGGL(CgDirectionKernel, dim3(SweepBlocks(n)), dim3(kThreads), 0, stream, pi,
                     Gpi, rho, Grho, beta, n);
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
