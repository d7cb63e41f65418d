// Batched banded Cholesky and LU factorisations and solves for gfx950,
// FP64, LAPACK band storage: the per-degree systems of the 3D step
// (DegreeOperator / ToroidalOperator::Factorise, BeginSolve;
// src/BandSolvers.cpp:23-139 is the algorithm, LAPACK's unblocked one).
//
// Factorisations: one wavefront per system. A band factorisation is serial
// along the diagonal; what one step j touches is the window of columns
// j..j+kd (Cholesky) or j..j+kl+ku (LU), which is staged in LDS as a ring of
// columns: column j is written back to HBM once final and column j+W read
// in its place, so every band element crosses HBM once each way. Within a
// step the lanes split the trailing update by column; the pivot search is
// the reference's sequential first-largest rule.
//
// Solves: one lane per right-hand-side column, 64 columns of one system per
// workgroup. The band column a step reads is the same for every lane (a
// broadcast load); each lane keeps the kd (or kl + ku) values its
// substitution still updates in a private LDS ring, so the sweep reads and
// writes its column once per direction. Right-hand sides may be laid out
// row-major (columns fastest), which makes the lanes' accesses coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace sl3amd {

namespace {

constexpr int kWave = 64;

__global__ void __launch_bounds__(kWave)
    CholeskyFactorKernel(const BandSystem* systems, double* ab, int* info) {
  extern __shared__ double lds[];
  const BandSystem sys = systems[blockIdx.x];
  const int n = sys.n, kd = sys.kl;
  const int W = kd + 1, ld = kd + 1;
  double* win = lds;           // W columns of ld
  double* row = lds + W * ld;  // row j of U right of the diagonal
  double* a = ab + sys.ab_off;
  const int lane = threadIdx.x;
  for (int c = 0; c < W && c < n; ++c) {
    for (int i = lane; i < ld; i += kWave) {
      win[c * ld + i] = a[static_cast<size_t>(c) * ld + i];
    }
  }
  __syncthreads();
  int result = 0;
  int j = 0;
  for (; j < n; ++j) {
    const int sj = j % W;
    const double d = win[sj * ld + kd];
    if (!(d > 0.0)) {
      result = j + 1;
      break;
    }
    const double ujj = sqrt(d);
    const double inv = 1.0 / ujj;
    const int kn = min(kd, n - 1 - j);
    __syncthreads();
    if (lane == 0) win[sj * ld + kd] = ujj;
    for (int c = lane + 1; c <= kn; c += kWave) {
      const int s = (j + c) % W;
      const double v = win[s * ld + kd - c] * inv;
      win[s * ld + kd - c] = v;
      row[c] = v;
    }
    __syncthreads();
    for (int c = lane + 1; c <= kn; c += kWave) {
      const int s = (j + c) % W;
      const double uc = row[c];
      for (int r = 1; r <= c; ++r) win[s * ld + kd + r - c] -= row[r] * uc;
    }
    __syncthreads();
    for (int i = lane; i < ld; i += kWave) {
      a[static_cast<size_t>(j) * ld + i] = win[sj * ld + i];
      if (j + W < n) win[sj * ld + i] = a[static_cast<size_t>(j + W) * ld + i];
    }
    __syncthreads();
  }
  if (result != 0) {
    // As the reference: the band partly overwritten, the window's columns
    // as they stand.
    for (int c = j; c < min(j + W, n); ++c) {
      for (int i = lane; i < ld; i += kWave) {
        a[static_cast<size_t>(c) * ld + i] = win[(c % W) * ld + i];
      }
    }
  }
  if (lane == 0) info[blockIdx.x] = result;
}

__global__ void __launch_bounds__(kWave)
    LuFactorKernel(const BandSystem* systems, double* ab, int* pivots,
                   int* info) {
  extern __shared__ double lds[];
  __shared__ int pivotShared;
  const BandSystem sys = systems[blockIdx.x];
  const int n = sys.n, kl = sys.kl, ku = sys.ku;
  const int kv = kl + ku;
  const int W = kv + 1, ld = 2 * kl + ku + 1;
  double* win = lds;
  double* a = ab + sys.ab_off;
  int* piv = pivots + sys.piv_off;
  const int lane = threadIdx.x;
  auto at = [&](int i, int c) -> double& {
    return win[(c % W) * ld + (kv + i - c)];
  };
  for (int c = 0; c < W && c < n; ++c) {
    for (int i = lane; i < ld; i += kWave) {
      double v = a[static_cast<size_t>(c) * ld + i];
      // dgbtf2's zeroed fill-in places of the first kv columns
      if (c >= ku + 1 && c < min(kv, n) && i < kv && i >= kv - c &&
          (i - kv + c) < c - ku) {
        v = 0.0;
      }
      win[c * ld + i] = v;
    }
  }
  __syncthreads();
  int result = 0;
  int ju = 0;
  for (int j = 0; j < n; ++j) {
    if (j + kv < n) {
      for (int i = lane; i < kl; i += kWave) at(j + i, j + kv) = 0.0;
    }
    __syncthreads();
    const int km = min(kl, n - 1 - j);
    if (lane == 0) {
      int p = j;
      for (int i = j + 1; i <= j + km; ++i) {
        if (fabs(at(i, j)) > fabs(at(p, j))) p = i;
      }
      pivotShared = p;
      piv[j] = p;
    }
    __syncthreads();
    const int p = pivotShared;
    if (at(p, j) == 0.0) {
      if (result == 0) result = j + 1;
    } else {
      ju = max(ju, min(p + ku, n - 1));
      if (p != j) {
        for (int c = j + lane; c <= ju; c += kWave) {
          const double t = at(p, c);
          at(p, c) = at(j, c);
          at(j, c) = t;
        }
        __syncthreads();
      }
      if (km > 0) {
        const double inv = 1.0 / at(j, j);
        __syncthreads();
        for (int i = j + 1 + lane; i <= j + km; i += kWave) at(i, j) *= inv;
        __syncthreads();
        for (int c = j + 1 + lane; c <= ju; c += kWave) {
          const double ujc = at(j, c);
          for (int i = j + 1; i <= j + km; ++i) at(i, c) -= at(i, j) * ujc;
        }
        __syncthreads();
      }
    }
    const int sj = j % W;
    for (int i = lane; i < ld; i += kWave) {
      a[static_cast<size_t>(j) * ld + i] = win[sj * ld + i];
      if (j + W < n) win[sj * ld + i] = a[static_cast<size_t>(j + W) * ld + i];
    }
    __syncthreads();
  }
  if (lane == 0) info[blockIdx.x] = result;
}

// Solves: block b handles columns tasks[b].y .. +63 of system tasks[b].x.
__global__ void __launch_bounds__(kWave)
    CholeskySolveKernel(const BandSystem* systems, const int2* tasks,
                        const double* ab, double* b) {
  extern __shared__ double ring[];  // (kd+1) x 64, lane fastest
  const int2 task = tasks[blockIdx.x];
  const BandSystem sys = systems[task.x];
  const int lane = threadIdx.x;
  const int col = task.y + lane;
  if (col >= sys.nrhs) return;  // whole lanes only; no barriers below
  const int n = sys.n, kd = sys.kl, W = kd + 1, ld = kd + 1;
  const double* u = ab + sys.ab_off;
  double* x = b + sys.b_off + static_cast<size_t>(col) * sys.col_stride;
  const long long rs = sys.row_stride;
  auto slot = [&](int i) -> double& { return ring[(i % W) * kWave + lane]; };
  // U^T y = b, forward, the sum over i ascending as the reference's.
  for (int j = 0; j < n; ++j) {
    const double* uj = u + static_cast<size_t>(j) * ld + kd;  // U(j,j)
    double s = x[j * rs];
    const int t0 = min(kd, j);
    for (int t = t0; t >= 1; --t) s -= uj[-t] * slot(j - t);
    const double v = s / uj[0];
    slot(j) = v;
    x[j * rs] = v;
  }
  // U x = y, backward; the ring holds the pending values x_{j-kd..j}.
  for (int t = 0; t <= kd && t < n; ++t) slot(n - 1 - t) = x[(n - 1 - t) * rs];
  for (int j = n - 1; j >= 0; --j) {
    const double* uj = u + static_cast<size_t>(j) * ld + kd;
    const double v = slot(j) / uj[0];
    x[j * rs] = v;
    const int t1 = min(kd, j);
    for (int t = 1; t <= t1; ++t) slot(j - t) -= uj[-t] * v;
    if (j - W >= 0) slot(j - W) = x[(j - W) * rs];
  }
}

__global__ void __launch_bounds__(kWave)
    LuSolveKernel(const BandSystem* systems, const int2* tasks,
                  const double* ab, const int* pivots, double* b) {
  extern __shared__ double ring[];  // (kl+ku+1) x 64
  const int2 task = tasks[blockIdx.x];
  const BandSystem sys = systems[task.x];
  const int lane = threadIdx.x;
  const int col = task.y + lane;
  if (col >= sys.nrhs) return;
  const int n = sys.n, kl = sys.kl, ku = sys.ku, kv = kl + ku;
  const int ld = 2 * kl + ku + 1;
  const double* lu = ab + sys.ab_off;
  const int* piv = pivots + sys.piv_off;
  double* x = b + sys.b_off + static_cast<size_t>(col) * sys.col_stride;
  const long long rs = sys.row_stride;
  // L y = P^T b: the ring holds x_j..x_{j+kl}.
  {
    const int R = kl + 1;
    auto slot = [&](int i) -> double& { return ring[(i % R) * kWave + lane]; };
    for (int t = 0; t < R && t < n; ++t) slot(t) = x[t * rs];
    for (int j = 0; j + 1 < n; ++j) {
      const int p = piv[j];
      if (p != j) {
        const double t = slot(p);
        slot(p) = slot(j);
        slot(j) = t;
      }
      const double xj = slot(j);
      const int lm = min(kl, n - 1 - j);
      const double* lj = lu + static_cast<size_t>(j) * ld + kv - j;  // L(i,j) = lj[i]
      for (int i = j + 1; i <= j + lm; ++i) slot(i) -= lj[i] * xj;
      x[j * rs] = xj;
      if (j + R < n) slot(j + R) = x[(j + R) * rs];
    }
    x[(n - 1) * rs] = slot(n - 1);
  }
  // U x = y, backward, kv superdiagonals.
  {
    const int W = kv + 1;
    auto slot = [&](int i) -> double& { return ring[(i % W) * kWave + lane]; };
    for (int t = 0; t < W && t < n; ++t) slot(n - 1 - t) = x[(n - 1 - t) * rs];
    for (int j = n - 1; j >= 0; --j) {
      const double* uj = lu + static_cast<size_t>(j) * ld + kv;  // U(j,j)
      const double v = slot(j) / uj[0];
      x[j * rs] = v;
      const int t1 = min(kv, j);
      for (int t = 1; t <= t1; ++t) slot(j - t) -= uj[-t] * v;
      if (j - W >= 0) slot(j - W) = x[(j - W) * rs];
    }
  }
}

}  // namespace

size_t CholeskyFactorLds(int kdMax) {
  return sizeof(double) * ((kdMax + 1) * (kdMax + 1) + kdMax + 1);
}
size_t LuFactorLds(int klMax, int kuMax) {
  return sizeof(double) * (klMax + kuMax + 1) * (2 * klMax + kuMax + 1);
}

void LaunchCholeskyFactor(const BandSystem* systems_dev, int count, int kdMax,
                          double* ab_dev, int* info_dev, hipStream_t stream) {
  if (count == 0) return;
  ProfileScope prof275("band_factor", stream);
  hipLaunchKernelGGL(CholeskyFactorKernel, dim3(count), dim3(kWave),
                     CholeskyFactorLds(kdMax), stream, systems_dev, ab_dev,
                     info_dev);
  SL3_HIP(hipGetLastError());
}

void LaunchLuFactor(const BandSystem* systems_dev, int count, int klMax,
                    int kuMax, double* ab_dev, int* piv_dev, int* info_dev,
                    hipStream_t stream) {
  if (count == 0) return;
  ProfileScope prof286("band_factor", stream);
  hipLaunchKernelGGL(LuFactorKernel, dim3(count), dim3(kWave),
                     LuFactorLds(klMax, kuMax), stream, systems_dev, ab_dev,
                     piv_dev, info_dev);
  SL3_HIP(hipGetLastError());
}

void LaunchCholeskySolve(const BandSystem* systems_dev, const int2* tasks_dev,
                         int tasks, int kdMax, const double* ab_dev,
                         double* b_dev, hipStream_t stream) {
  if (tasks == 0) return;
  ProfileScope prof297("band_solve", stream);
  hipLaunchKernelGGL(CholeskySolveKernel, dim3(tasks), dim3(kWave),
                     sizeof(double) * (kdMax + 1) * kWave, stream, systems_dev,
                     tasks_dev, ab_dev, b_dev);
  SL3_HIP(hipGetLastError());
}

void LaunchLuSolve(const BandSystem* systems_dev, const int2* tasks_dev,
                   int tasks, int kvMax, const double* ab_dev,
                   const int* piv_dev, double* b_dev, hipStream_t stream) {
  if (tasks == 0) return;
  ProfileScope prof308("band_solve", stream);
  hipLaunchKernelGGL(LuSolveKernel, dim3(tasks), dim3(kWave),
                     sizeof(double) * (kvMax + 1) * kWave, stream, systems_dev,
                     tasks_dev, ab_dev, piv_dev, b_dev);
  SL3_HIP(hipGetLastError());
}

std::vector<int2> SolveTasks(const std::vector<BandSystem>& systems) {
  std::vector<int2> tasks;
  for (size_t s = 0; s < systems.size(); ++s) {
    for (int c = 0; c < systems[s].nrhs; c += kWave) {
      tasks.push_back(make_int2(static_cast<int>(s), c));
    }
  }
  return tasks;
}

}  // namespace sl3amd
