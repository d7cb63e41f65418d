// The memory response of the degree systems on gfx950 (see response.hpp):
// its set-up from unit-column band solves, and the per-pass grouped GEMM
// x = G f on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// Operand layouts of v_mfma_f64_16x16x4f64 (one wave, lane i):
//   A (16 x 4):  A[i % 16][i / 16]
//   B (4 x 16):  B[i / 16][i % 16]
//   C (16 x 16): four doubles, C[4 j + i / 16][i % 16] for j = 0..3
// G is stored tile by tile in exactly the A order, so each wave streams it
// with one coalesced 512-byte load per instruction; f is row-major with
// rows padded to 16 columns, so a B operand is four 128-byte row segments.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "response.hpp"

namespace sl3amd {

namespace {

constexpr int kThreads = 256;
constexpr int kWave = 64;

typedef double Acc4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kThreads)
    UnitColumnsKernel(const DegreeTask* tasks, const int* dm, const double* scale,
                      const long long* zOff, double* z) {
  const DegreeTask t = tasks[blockIdx.y];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= t.nm) return;
  const int d = dm[t.dmOff + j];
  z[zOff[blockIdx.y] + static_cast<long long>(d) * t.nm + j] = scale[t.scaleOff + d];
}

__global__ void __launch_bounds__(kThreads)
    PackResponseKernel(const DegreeTask* tasks, const int* dm, const double* scale,
                       const long long* zOff, const double* z, double* g) {
  const DegreeTask t = tasks[blockIdx.y];
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long total = static_cast<long long>(t.nmPad) * t.nmPad;
  if (idx >= total) return;
  const int k4 = t.nmPad / 4;
  const long long tile = idx / kWave;
  const int lane = static_cast<int>(idx % kWave);
  const int rb = static_cast<int>(tile / k4), kb = static_cast<int>(tile % k4);
  const int row = 16 * rb + lane % 16, col = 4 * kb + lane / 16;
  double v = 0.0;
  if (row < t.nm && col < t.nm) {
    const int d = dm[t.dmOff + row];
    v = scale[t.scaleOff + d] * z[zOff[blockIdx.y] + static_cast<long long>(d) * t.nm + col];
  }
  g[t.gOff + idx] = v;
}

// One wave: a 32 x 64 block of one task's x (2 x 4 matrix-core tiles),
// accumulated over the whole contraction; blocks past the padded edge are
// skipped (wave-uniform branches). x is written column by column (order
// major), the layout strain recovery reads with neighbouring nodes in
// neighbouring lanes.
__global__ void __launch_bounds__(kThreads)
    ResponseKernel(const DegreeTask* tasks, const ResponseWork* work, int count,
                   const double* g, const double* f, double* x) {
  const int wave = blockIdx.x * (kThreads / kWave) + threadIdx.x / kWave;
  if (wave >= count) return;
  const int lane = threadIdx.x % kWave;
  const ResponseWork w = work[wave];
  const DegreeTask t = tasks[w.task];
  const int rbCount = t.nmPad / 16, k4 = t.nmPad / 4, cbCount = t.colsPad / 16;
  const bool row1 = w.rb + 1 < rbCount;
  const int cols = min(4, cbCount - w.cb);
  const double* a0 = g + t.gOff + static_cast<long long>(w.rb) * k4 * kWave + lane;
  const double* a1 = a0 + static_cast<long long>(k4) * kWave;
  const long long ld = t.colsPad;
  const double* b = f + t.fOff + (lane / 16) * ld + 16 * w.cb + lane % 16;
  Acc4 acc[2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = Acc4{0.0, 0.0, 0.0, 0.0};
  }
#pragma unroll 2
  for (int kb = 0; kb < k4; ++kb) {
    const double av0 = a0[static_cast<long long>(kb) * kWave];
    const double av1 = row1 ? a1[static_cast<long long>(kb) * kWave] : 0.0;
    double bv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bv[c] = c < cols ? b[4 * kb * ld + 16 * c] : 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c < cols) {
        acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv[c], acc[0][c], 0, 0, 0);
        if (row1) {
          acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv[c], acc[1][c], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    if (r == 1 && !row1) break;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c >= cols) break;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long row = 16 * (w.rb + r) + 4 * j + lane / 16;
        const long long col = 16 * (w.cb + c) + lane % 16;
        x[t.fOff + col * t.nmPad + row] = acc[r][c][j];  // column-major
      }
    }
  }
}

__global__ void __launch_bounds__(kThreads)
    ExpandForceKernel(const DegreeTask* tasks, const int* dm, const double* scale,
                      const double* f, double* rhs) {
  const DegreeTask t = tasks[blockIdx.y];
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(t.nm) * t.cols) return;
  const int i = static_cast<int>(idx / t.cols), c = static_cast<int>(idx % t.cols);
  const int d = dm[t.dmOff + i];
  rhs[t.rhsOff + static_cast<long long>(d) * t.cols + c] =
      scale[t.scaleOff + d] * f[t.fOff + static_cast<long long>(i) * t.colsPad + c];
}

}  // namespace

std::vector<ResponseWork> ResponseWorkList(const std::vector<DegreeTask>& tasks) {
  std::vector<ResponseWork> work;
  for (size_t ti = 0; ti < tasks.size(); ++ti) {
    const auto& t = tasks[ti];
    for (int rb = 0; rb < t.nmPad / 16; rb += 2) {
      for (int cb = 0; cb < t.colsPad / 16; cb += 4) {
        work.push_back({static_cast<int>(ti), rb, cb, 0});
      }
    }
  }
  // the longest contractions first, so the short ones fill the tail
  std::stable_sort(work.begin(), work.end(), [&](const ResponseWork& a, const ResponseWork& b) {
    return tasks[a.task].nmPad > tasks[b.task].nmPad;
  });
  return work;
}

void LaunchUnitColumns(const DegreeTask* tasks_dev, int taskCount, int nmMax,
                       const int* dm, const double* scale, const long long* zOff,
                       double* z, hipStream_t stream) {
  if (taskCount == 0 || nmMax == 0) return;
  hipLaunchKernelGGL(UnitColumnsKernel, dim3(CeilDiv(nmMax, kThreads), taskCount),
                     dim3(kThreads), 0, stream, tasks_dev, dm, scale, zOff, z);
  SL3_HIP(hipGetLastError());
}

void LaunchPackResponse(const DegreeTask* tasks_dev, int taskCount, int nmPadMax,
                        const int* dm, const double* scale, const long long* zOff,
                        const double* z, double* g, hipStream_t stream) {
  if (taskCount == 0 || nmPadMax == 0) return;
  const long long total = static_cast<long long>(nmPadMax) * nmPadMax;
  hipLaunchKernelGGL(PackResponseKernel, dim3(CeilDiv(total, kThreads), taskCount),
                     dim3(kThreads), 0, stream, tasks_dev, dm, scale, zOff, z, g);
  SL3_HIP(hipGetLastError());
}

void LaunchResponse(const DegreeTask* tasks_dev, const ResponseWork* work_dev,
                    int workCount, const double* g, const double* f, double* x,
                    hipStream_t stream) {
  if (workCount == 0) return;
  ProfileScope prof("response", stream);
  hipLaunchKernelGGL(ResponseKernel, dim3(CeilDiv(workCount, kThreads / kWave)),
                     dim3(kThreads), 0, stream, tasks_dev, work_dev, workCount, g, f, x);
  SL3_HIP(hipGetLastError());
}

void LaunchExpandForce(const DegreeTask* tasks_dev, int taskCount, int nmMax,
                       int colsMax, const int* dm, const double* scale,
                       const double* f, double* rhs, hipStream_t stream) {
  if (taskCount == 0 || nmMax == 0) return;
  const long long total = static_cast<long long>(nmMax) * colsMax;
  hipLaunchKernelGGL(ExpandForceKernel, dim3(CeilDiv(total, kThreads), taskCount),
                     dim3(kThreads), 0, stream, tasks_dev, dm, scale, f, rhs);
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
