// The memory response of the degree systems on gfx950 (see response.hpp):
// its set-up from unit-column band solves, and the per-pass grouped GEMM
// x = G f on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// Operand layouts of v_mfma_f64_16x16x4f64 (one wave, lane i):
//   A (16 x 4):  A[i % 16][i / 16]
//   B (4 x 16):  B[i / 16][i % 16]
//   C (16 x 16): four doubles, C[4 j + i / 16][i % 16] for j = 0..3
// G is stored tile by tile in exactly the A order, so a workgroup stages a
// chunk of it with coalesced 512-byte runs and each lane reads its operand
// from LDS without conflicts; f is row-major with rows padded to 16 columns.
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

// One workgroup: an RT x CT block of one task's 16 x 16 tiles of x (row
// tiles rb .. rb + RT - 1, column tiles cb .. cb + CT - 1), each wave 2 x 4
// matrix-core tiles: the waves 2 x 2 for a wide task's 64 x 128 blocks (RT
// 4, CT 8), 4 x 1 for a narrow task's 128 x 64 (RT 8, CT 4; a task of at
// most four column tiles then keeps all four waves on products). The
// contraction runs in chunks of 16: each chunk's G tiles (already in
// operand order) and f rows are staged in LDS and shared by the four
// waves, the next chunk's loads in flight under the current products.
// Tiles past the task's padded edge are staged as zeros; x is written
// column by column (order major), the layout strain recovery reads with
// neighbouring nodes in neighbouring lanes. Every x element sums its
// contraction in the same order whatever the tiling.
constexpr int kChunk = 4;  // contraction steps of 4 per staged chunk

template <int RT, int CT>
__device__ inline void ResponseBody(const DegreeTask* tasks, const ResponseWork& w,
                                    const double* g, const double* f, double* x,
                                    const int* live, int realisations, double* ldsA,
                                    double* ldsB) {
  constexpr int kRowTiles = RT, kColTiles = CT;
  constexpr int kWavesCol = 4 / (RT / 2);  // waves along the columns
  constexpr int kAPer = kRowTiles * kChunk * kWave / kThreads;   // staged A per thread
  constexpr int kBPer = 4 * kChunk * 16 * kColTiles / kThreads;  // staged B per thread
  constexpr int kBRow = 16 * kColTiles + 1;
  // As[tile][lane] and Bs[k][column] in the kernel's LDS
  auto As = [&](int i, int j) -> double& { return ldsA[i * kWave + j]; };
  auto Bs = [&](int i, int j) -> double& { return ldsB[i * kBRow + j]; };
  const DegreeTask t = tasks[w.task];
  if (live) {  // skip a tile of converged realisations' columns only
    const int c0 = 16 * w.cb, c1 = min(c0 + 16 * kColTiles, realisations * t.cols) - 1;
    bool any = false;
    for (int r = c0 / t.cols; r <= c1 / t.cols; ++r) any = any || live[r] != 0;
    if (!any) return;
  }
  const int rbCount = t.nmPad / 16, k4 = t.nmPad / 4, cbCount = t.colsPad / 16;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  // the wave's 2 row tiles and 4 column tiles
  const int wr = wave / kWavesCol, wc = wave % kWavesCol;
  const long long ld = t.colsPad;

  double pa[kAPer], pb[kBPer];
  auto load = [&](int kc) {  // contraction steps 4 kc .. 4 kc + 3
#pragma unroll
    for (int q = 0; q < kAPer; ++q) {
      const int e = threadIdx.x + kThreads * q;  // (tile, lane), tile = r kChunk + s
      const int tile = e / kWave, ln = e % kWave;
      const int r = tile / kChunk, s = tile % kChunk;
      pa[q] = 0.0;
      if (w.rb + r < rbCount) {
        pa[q] = g[t.gOff + (static_cast<long long>(w.rb + r) * k4 + kChunk * kc + s) * kWave + ln];
      }
    }
#pragma unroll
    for (int q = 0; q < kBPer; ++q) {
      const int e = threadIdx.x + kThreads * q;  // (k, column)
      const int k = e / (16 * kColTiles), c = e % (16 * kColTiles);
      pb[q] = 0.0;
      if (16 * w.cb + c < t.colsPad) {
        pb[q] = f[t.fOff + (4 * kChunk * kc + k) * ld + 16 * w.cb + c];
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < kAPer; ++q) {
      const int e = threadIdx.x + kThreads * q;
      As(e / kWave, e % kWave) = pa[q];
    }
#pragma unroll
    for (int q = 0; q < kBPer; ++q) {
      const int e = threadIdx.x + kThreads * q;
      Bs(e / (16 * kColTiles), e % (16 * kColTiles)) = pb[q];
    }
  };

  Acc4 acc[2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = Acc4{0.0, 0.0, 0.0, 0.0};
  }
  const int chunks = k4 / kChunk;  // nmPad is a multiple of 16
  load(0);
  for (int kc = 0; kc < chunks; ++kc) {
    store();
    __syncthreads();
    if (kc + 1 < chunks) load(kc + 1);
#pragma unroll
    for (int s = 0; s < kChunk; ++s) {
      const double a0 = As((2 * wr) * kChunk + s, lane);
      const double a1 = As((2 * wr + 1) * kChunk + s, lane);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        // tiles past the task's padded edge are skipped (wave-uniform)
        if (w.cb + 4 * wc + c < cbCount) {
          const double b = Bs(4 * s + lane / 16, (4 * wc + c) * 16 + lane % 16);
          if (w.rb + 2 * wr < rbCount) {
            acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b, acc[0][c], 0, 0, 0);
          }
          if (w.rb + 2 * wr + 1 < rbCount) {
            acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b, acc[1][c], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();
  }
  // guards, not early exits: the accumulators keep static indices
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int rt = w.rb + 2 * wr + r, ct = w.cb + 4 * wc + c;
      if (rt >= rbCount || ct >= cbCount) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long row = 16 * rt + 4 * j + lane / 16;
        const long long col = 16 * ct + lane % 16;
        x[t.fOff + col * t.nmPad + row] = acc[r][c][j];  // column-major
      }
    }
  }
}

// Both tilings in one launch (each item's pad says which), so the narrow
// tasks' workgroups fill in beside the wide ones'.
__global__ void __launch_bounds__(kThreads)
    ResponseKernel(const DegreeTask* tasks, const ResponseWork* work, const double* g,
                   const double* f, double* x, const int* live, int realisations) {
  // the larger of the two tilings' staging: A 8 x 4 tiles, B 16 x (16 x 8 + 1)
  __shared__ double ldsA[8 * kChunk * kWave];
  __shared__ double ldsB[4 * kChunk * (16 * 8 + 1)];
  const ResponseWork w = work[blockIdx.x];
  if (w.pad == 1) {
    ResponseBody<8, 4>(tasks, w, g, f, x, live, realisations, ldsA, ldsB);
  } else {
    ResponseBody<4, 8>(tasks, w, g, f, x, live, realisations, ldsA, ldsB);
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

// Memory-dof force rows (row-major nmPad x colsPad at fOff; Re of
// m = 0..l, then Im) ->
// the reference's solution layout (ndof x (l + 1) complex, column-major, at
// solOff[task]); the other dofs are left as they are.
__global__ void __launch_bounds__(kThreads)
    RowsToSolutionsKernel(const DegreeTask* tasks, const int* dm, const long long* solOff,
                          const double* rows, double* sol) {
  const DegreeTask t = tasks[blockIdx.y];
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int orders = t.l + 1;
  if (idx >= static_cast<long long>(t.nm) * orders) return;
  const int i = static_cast<int>(idx / orders), m = static_cast<int>(idx % orders);
  const int d = dm[t.dmOff + i];
  const double* row = rows + t.fOff + static_cast<long long>(i) * t.colsPad;
  double* z = sol + solOff[blockIdx.y] + 2 * (static_cast<long long>(m) * t.ndof + d);
  z[0] = row[m];
  z[1] = row[orders + m];
}

// The inverse gather into the response layout (column-major: column c of
// a task at fOff + c nmPad, Re of m = 0..l then Im; padding untouched).
__global__ void __launch_bounds__(kThreads)
    SolutionsToRowsKernel(const DegreeTask* tasks, const int* dm, const long long* solOff,
                          const double* sol, double* rows) {
  const DegreeTask t = tasks[blockIdx.y];
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int orders = t.l + 1;
  if (idx >= static_cast<long long>(t.nm) * orders) return;
  const int i = static_cast<int>(idx / orders), m = static_cast<int>(idx % orders);
  const int d = dm[t.dmOff + i];
  const double* z = sol + solOff[blockIdx.y] + 2 * (static_cast<long long>(m) * t.ndof + d);
  rows[t.fOff + static_cast<long long>(m) * t.nmPad + i] = z[0];
  rows[t.fOff + static_cast<long long>(orders + m) * t.nmPad + i] = z[1];
}

}  // namespace

void LaunchRowsToSolutions(const DegreeTask* tasks_dev, int taskCount, int nmMax, int lmax,
                           const int* dm, const long long* solOff, const double* rows,
                           double* solutions, hipStream_t stream) {
  if (taskCount == 0 || nmMax == 0) return;
  const long long total = static_cast<long long>(nmMax) * (lmax + 1);
  hipLaunchKernelGGL(RowsToSolutionsKernel, dim3(CeilDiv(total, kThreads), taskCount),
                     dim3(kThreads), 0, stream, tasks_dev, dm, solOff, rows, solutions);
  SL3_HIP(hipGetLastError());
}

void LaunchSolutionsToRows(const DegreeTask* tasks_dev, int taskCount, int nmMax, int lmax,
                           const int* dm, const long long* solOff, const double* solutions,
                           double* rows, hipStream_t stream) {
  if (taskCount == 0 || nmMax == 0) return;
  const long long total = static_cast<long long>(nmMax) * (lmax + 1);
  hipLaunchKernelGGL(SolutionsToRowsKernel, dim3(CeilDiv(total, kThreads), taskCount),
                     dim3(kThreads), 0, stream, tasks_dev, dm, solOff, solutions, rows);
  SL3_HIP(hipGetLastError());
}

std::vector<ResponseWork> ResponseWorkList(const std::vector<DegreeTask>& tasks) {
  std::vector<ResponseWork> narrow, wide;
  for (size_t ti = 0; ti < tasks.size(); ++ti) {
    const auto& t = tasks[ti];
    const bool thin = t.colsPad / 16 <= kNarrowColTiles;
    const int rStep = thin ? 8 : 4, cStep = thin ? 4 : 8;
    for (int rb = 0; rb < t.nmPad / 16; rb += rStep) {
      for (int cb = 0; cb < t.colsPad / 16; cb += cStep) {
        (thin ? narrow : wide).push_back({static_cast<int>(ti), rb, cb, thin ? 1 : 0});
      }
    }
  }
  // the longest contractions first, so the short ones fill the tail
  auto longest = [&](const ResponseWork& a, const ResponseWork& b) {
    return tasks[a.task].nmPad > tasks[b.task].nmPad;
  };
  std::stable_sort(narrow.begin(), narrow.end(), longest);
  std::stable_sort(wide.begin(), wide.end(), longest);
  narrow.insert(narrow.end(), wide.begin(), wide.end());
  return narrow;
}

int ResponseNarrowCount(const std::vector<ResponseWork>& work) {
  int n = 0;
  while (n < static_cast<int>(work.size()) && work[n].pad == 1) ++n;
  return n;
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
                    int workCount, int narrowCount, const double* g, const double* f, double* x,
                    hipStream_t stream, const int* live, int realisations) {
  if (workCount == 0) return;
  ProfileScope prof("response", stream);
  (void)narrowCount;  // the items say their tiling
  hipLaunchKernelGGL(ResponseKernel, dim3(workCount), dim3(kThreads), 0, stream, tasks_dev,
                     work_dev, g, f, x, live, realisations);
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
