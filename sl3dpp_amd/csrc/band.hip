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
// workgroup. The band column a step reads is the same for every lane; each
// lane keeps the kd (or kl + ku) values its substitution still needs in a
// private LDS ring, so the sweep reads and writes its column once per
// direction. The Cholesky sweeps stage 32 rows at a time (their band
// columns and right-hand side values, every load in flight together), so a
// row's serial step waits on LDS, not on HBM. Right-hand sides may be laid
// out row-major (columns fastest), which makes the lanes' accesses
// coalesced.
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

// Rows of a solve staged per block: the band columns and the right-hand
// side values of kSolveRows rows are read with every load in flight at once
// (the loops below are unrolled over fixed counts), then the serial
// substitution runs from LDS.
constexpr int kSolveRows = 32;
constexpr int kStageUnroll = 8;

// The smallest power of two >= w: the solves' rings index by a mask (a
// modulo by a runtime width costs an integer division per access).
__host__ __device__ inline int RingSlots(int w) {
  int p = 1;
  while (p < w) p <<= 1;
  return p;
}

// dst[e] = src[e] for e < count, the workgroup's 64 lanes together, up to
// 8 loads in flight per lane.
__device__ inline void StageBlock(double* dst, const double* __restrict__ src, int count,
                                  int lane) {
  for (int base = 0; base < count; base += kWave * kStageUnroll) {
    double v[kStageUnroll];
#pragma unroll
    for (int k = 0; k < kStageUnroll; ++k) {
      const int e = base + k * kWave + lane;
      v[k] = e < count ? src[e] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kStageUnroll; ++k) {
      const int e = base + k * kWave + lane;
      if (e < count) dst[e] = v[k];
    }
  }
}

// A lane's right-hand side values of rows j0 .. j0 + rows - 1 into
// xs[r * 64 + lane], all loads in flight at once.
__device__ inline void StageColumn(double* xs, const double* x, long long rs, int j0, int rows,
                                   int lane) {
  double v[kSolveRows];
#pragma unroll
  for (int r = 0; r < kSolveRows; ++r) v[r] = r < rows ? x[(j0 + r) * rs] : 0.0;
#pragma unroll
  for (int r = 0; r < kSolveRows; ++r) {
    if (r < rows) xs[r * kWave + lane] = v[r];
  }
}

__device__ inline void StoreColumn(const double* xs, double* x, long long rs, int j0, int rows,
                                   int lane) {
#pragma unroll
  for (int r = 0; r < kSolveRows; ++r) {
    if (r < rows) x[(j0 + r) * rs] = xs[r * kWave + lane];
  }
}

// Both Cholesky sweeps for a band width KD known at compile time, in the
// column (axpy) form: once a row's value is known, the KD rows after it
// (before it, going back) take its term at once. The pending sums of those
// rows live in registers (acc[k]: the row k steps ahead), so a row's
// critical path is its last term and one multiply by the diagonal's
// reciprocal (formed per block, one lane per row, off the path), while its
// KD updates are independent of one another. Every row takes its terms
// oldest first (t = KD..1), the order of the ring path's sums; the
// reciprocal multiply differs from the ring path's division by at most an
// ulp per row. The band columns of R rows (a panel: 4, 8 or 16, picked per
// batch by the tuner; every row's arithmetic is the same whatever R is,
// so the panel changes the staging, not a bit of the result) and the KD
// after them
// are staged in LDS, double-buffered: the next block's columns and
// right-hand side values are loaded into registers while the current block
// is substituted. Rows past either end of the matrix only ever occupy
// pending sums that are never written out. A wavefront issues at most one
// instruction every four cycles and a row costs about 65 (17 FMAs, the
// LDS reads, the block's share of staging), so one long system's sweep is
// issue-bound: 465 cycles a row measured at n = 1281, kd = 17.
constexpr int kDefaultPanelRows = 8;

// One staged block of forward steps r = j0 + q: y_r = acc[0] U(r, r)^-1,
// then U(r, r + k) y_r off the pending rows, and row r + KD enters with its
// right-hand side bin[q]. fac holds columns j0 .. of the block; a step's
// band entries are read from LDS one step ahead (the scheduling barrier
// keeps those reads ahead of the step before), so LDS latency stays off
// the rows' serial chain.
template <int KD, bool FULL, int R>
__device__ inline void ForwardSteps(const double* fac, const double* rinv, const double* bin,
                                    double* acc, double* out, int rows) {
  constexpr int ld = KD + 1;
  double cf[KD + 1], nf[KD + 1];  // U(r, r + k), k = 1..KD; [0]: the reciprocal
  cf[0] = rinv[0];
#pragma unroll
  for (int k = 1; k <= KD; ++k) cf[k] = fac[k * ld + KD - k];  // column k's entry KD - k
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (q + 1 < R) {
      nf[0] = rinv[q + 1];
#pragma unroll
      for (int k = 1; k <= KD; ++k) nf[k] = fac[(q + 1 + k) * ld + KD - k];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (FULL || q < rows) {
      const double y = acc[0] * cf[0];
      out[q] = y;
#pragma unroll
      for (int k = 1; k < KD; ++k) acc[k] -= cf[k] * y;
      const double enter = bin[q] - cf[KD] * y;
#pragma unroll
      for (int k = 0; k + 1 < KD; ++k) acc[k] = acc[k + 1];
      acc[KD - 1] = enter;
    }
#pragma unroll
    for (int k = 0; k <= KD; ++k) cf[k] = nf[k];
  }
}

// One staged block of backward steps j = j0 + q, q descending: x_j =
// acc[0] U(j, j)^-1, then U(j - k, j) x_j off the pending rows, and row
// j - KD enters with its value yin[q]. fac holds columns j0 .. of the
// block; band entries are read one step ahead as in ForwardSteps.
template <int KD, bool FULL, int R>
__device__ inline void BackwardSteps(const double* fac, const double* rinv, const double* yin,
                                     double* acc, double* out, int rows) {
  constexpr int ld = KD + 1;
  double cf[KD + 1], nf[KD + 1];  // U(j - KD + k, j) at [k], k < KD; [KD]: the reciprocal
  constexpr int q0 = R - 1;
#pragma unroll
  for (int k = 0; k < KD; ++k) cf[k] = fac[q0 * ld + k];
  cf[KD] = rinv[q0];
#pragma unroll
  for (int q = R - 1; q >= 0; --q) {
    if (q > 0) {
#pragma unroll
      for (int k = 0; k < KD; ++k) nf[k] = fac[(q - 1) * ld + k];
      nf[KD] = rinv[q - 1];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (FULL || q < rows) {
      const double x = acc[0] * cf[KD];
      out[q] = x;
      // U(j - k, j) is column j's entry KD - k
#pragma unroll
      for (int k = 1; k < KD; ++k) acc[k] -= cf[KD - k] * x;
      const double enter = yin[q] - cf[0] * x;
#pragma unroll
      for (int k = 0; k + 1 < KD; ++k) acc[k] = acc[k + 1];
      acc[KD - 1] = enter;
    }
#pragma unroll
    for (int k = 0; k <= KD; ++k) cf[k] = nf[k];
  }
}

template <int KD, int R>
__device__ void CholeskySweepsFixed(const double* __restrict__ u, double* __restrict__ x,
                                    long long rs, int n, bool active, int lane, double* lds) {
  constexpr int ld = KD + 1;
  constexpr int F = (R + KD) * ld;  // band doubles a block stages, at most
  constexpr int FL = (F + kWave - 1) / kWave;
  double acc[KD], cur[R], nxt[R], out[R], fn[FL];
  double* rinv = lds + 2 * F;  // R reciprocals of the staged diagonal
  auto value = [&](int row) { return (active && row >= 0 && row < n) ? x[row * rs] : 0.0; };
  // U^T y = b. Block j0 stages columns j0 .. j0 + R + KD - 1 (those < n)
  // and the right-hand sides of rows j0 + KD .. j0 + KD + R - 1.
#pragma unroll
  for (int k = 0; k < KD; ++k) acc[k] = value(k);
  {
    int j0 = 0;
    int cnt = (min(n, R + KD)) * ld;
#pragma unroll
    for (int q = 0; q < R; ++q) nxt[q] = value(KD + q);
#pragma unroll
    for (int k = 0; k < FL; ++k) {
      const int e = k * kWave + lane;
      fn[k] = e < cnt ? u[e] : 0.0;
    }
    for (int blk = 0; j0 < n; ++blk) {
      double* fac = lds + (blk & 1) * F;
#pragma unroll
      for (int k = 0; k < FL; ++k) {
        const int e = k * kWave + lane;
        if (e < F) fac[e] = fn[k];
      }
#pragma unroll
      for (int q = 0; q < R; ++q) cur[q] = nxt[q];
      __syncthreads();
      // the block's diagonal reciprocals, one lane per row
      if (lane < R) rinv[lane] = 1.0 / fac[lane * ld + KD];
      __syncthreads();
      const int rows = min(R, n - j0);
      const int nj0 = j0 + R;
      // the previous block's values go out now, a block before the next
      // wait on this lane's memory operations
      if (active && blk > 0) {
#pragma unroll
        for (int q = 0; q < R; ++q) x[(j0 - R + q) * rs] = out[q];
      }
      if (nj0 < n) {
        const int ncnt = (min(n, nj0 + R + KD) - nj0) * ld;
#pragma unroll
        for (int q = 0; q < R; ++q) nxt[q] = value(nj0 + KD + q);
#pragma unroll
        for (int k = 0; k < FL; ++k) {
          const int e = k * kWave + lane;
          fn[k] = e < ncnt ? u[static_cast<size_t>(nj0) * ld + e] : 0.0;
        }
      }
      if (rows == R) {
        ForwardSteps<KD, true, R>(fac, rinv, cur, acc, out, rows);
      } else {
        ForwardSteps<KD, false, R>(fac, rinv, cur, acc, out, rows);
        // the last block: its values go out here
        if (active) {
#pragma unroll
          for (int q = 0; q < R; ++q) {
            if (q < rows) x[(j0 + q) * rs] = out[q];
          }
        }
      }
      j0 = nj0;
    }
    // a full last block's values
    if (active && n % R == 0) {
#pragma unroll
      for (int q = 0; q < R; ++q) x[(n - R + q) * rs] = out[q];
    }
  }
  __syncthreads();
  // U x = y. Block [j0, j1) from the bottom stages its own columns and
  // the values of rows j0 - KD .. j1 - KD - 1.
#pragma unroll
  for (int k = 0; k < KD; ++k) acc[k] = value(n - 1 - k);
  {
    int j1 = n, j0 = max(0, n - R);
    int cnt = (j1 - j0) * ld;
#pragma unroll
    for (int q = 0; q < R; ++q) nxt[q] = value(j0 + q - KD);
#pragma unroll
    for (int k = 0; k < FL; ++k) {
      const int e = k * kWave + lane;
      fn[k] = e < cnt ? u[static_cast<size_t>(j0) * ld + e] : 0.0;
    }
    for (int blk = 0; j1 > 0; ++blk) {
      double* fac = lds + (blk & 1) * F;
#pragma unroll
      for (int k = 0; k < FL; ++k) {
        const int e = k * kWave + lane;
        if (e < cnt) fac[e] = fn[k];
      }
#pragma unroll
      for (int q = 0; q < R; ++q) cur[q] = nxt[q];
      __syncthreads();
      if (lane < R) rinv[lane] = 1.0 / fac[lane * ld + KD];
      __syncthreads();
      const int rows = j1 - j0;
      const int nj1 = j0, nj0 = max(0, j0 - R);
      const int ncnt = (nj1 - nj0) * ld;
      // the previous (full) block [j1, j1 + R) goes out now
      if (active && blk > 0) {
#pragma unroll
        for (int q = 0; q < R; ++q) x[(j1 + q) * rs] = out[q];
      }
      if (nj1 > 0) {
#pragma unroll
        for (int q = 0; q < R; ++q) nxt[q] = value(nj0 + q - KD);
#pragma unroll
        for (int k = 0; k < FL; ++k) {
          const int e = k * kWave + lane;
          fn[k] = e < ncnt ? u[static_cast<size_t>(nj0) * ld + e] : 0.0;
        }
      }
      if (rows == R) {
        BackwardSteps<KD, true, R>(fac, rinv, cur, acc, out, rows);
      } else {
        BackwardSteps<KD, false, R>(fac, rinv, cur, acc, out, rows);
      }
      if (active && j0 == 0) {
        // the last block (the only one that may be partial) goes out here
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (q < rows) x[q * rs] = out[q];
        }
      }
      j1 = nj1;
      j0 = nj0;
      cnt = ncnt;
    }
  }
}

// Solves: block b handles columns tasks[b].y .. +63 of system tasks[b].x.
// LDS: the ring RingSlots(kd + 1) x 64, lane fastest; the staged band
// columns (kSolveRows + kd) x (kd + 1); the staged values kSolveRows x 64.
template <int PANEL>
__global__ void __launch_bounds__(kWave)
    CholeskySolveKernel(const BandSystem* systems, const int2* tasks,
                        const double* ab, double* b) {
  extern __shared__ double lds[];
  const int2 task = tasks[blockIdx.x];
  const BandSystem sys = systems[task.x];
  const int lane = threadIdx.x;
  const int col = task.y + lane;
  // every lane stages and meets the barriers; only active lanes touch x
  const bool active = col < sys.nrhs;
  const int n = sys.n, kd = sys.kl, ld = kd + 1;
  const double* u = ab + sys.ab_off;
  double* x = b + sys.b_off + static_cast<size_t>(active ? col : 0) * sys.col_stride;
  const long long rs = sys.row_stride;
  // the widths of the step's systems: spheroidal 3 ngll - 1 and toroidal
  // ngll - 1 at ngll 5 and 6
  switch (kd) {
    case 4: CholeskySweepsFixed<4, PANEL>(u, x, rs, n, active, lane, lds); return;
    case 5: CholeskySweepsFixed<5, PANEL>(u, x, rs, n, active, lane, lds); return;
    case 14: CholeskySweepsFixed<14, PANEL>(u, x, rs, n, active, lane, lds); return;
    case 17: CholeskySweepsFixed<17, PANEL>(u, x, rs, n, active, lane, lds); return;
    default: break;
  }
  // a power-of-two ring of at least kd + 1 slots: slot i at i & mask
  const int P = RingSlots(kd + 1), mask = P - 1;
  double* ring = lds;
  double* fac = ring + P * kWave;
  double* xs = fac + (kSolveRows + kd) * ld;
  auto slot = [&](int i) -> double& { return ring[(i & mask) * kWave + lane]; };
  // U^T y = b, forward: row j's sum over U(j - t, j) y_{j-t}, t = kd..1
  // (i ascending, as the reference's), the band column j staged.
  for (int j0 = 0; j0 < n; j0 += kSolveRows) {
    const int rows = min(kSolveRows, n - j0);
    StageBlock(fac, u + static_cast<size_t>(j0) * ld, rows * ld, lane);
    if (active) StageColumn(xs, x, rs, j0, rows, lane);
    __syncthreads();
    if (active) {
      for (int r = 0; r < rows; ++r) {
        const int j = j0 + r;
        const double* uj = fac + r * ld + kd;  // U(j, j); U(j - t, j) at uj[-t]
        double s = xs[r * kWave + lane];
        const int t0 = min(kd, j);
        for (int t = t0; t >= 1; --t) s -= uj[-t] * slot(j - t);
        const double v = s / uj[0];
        slot(j) = v;
        xs[r * kWave + lane] = v;
      }
      StoreColumn(xs, x, rs, j0, rows, lane);
    }
    __syncthreads();
  }
  // U x = y, backward: row i's sum over U(i, i + t) x_{i+t}, t = kd..1 (the
  // order in which the column sweep subtracts them), the band columns
  // i .. i + kd staged; the ring holds x_{i+1..i+kd}.
  for (int j1 = n; j1 > 0; j1 -= kSolveRows) {
    const int j0 = max(0, j1 - kSolveRows);
    const int rows = j1 - j0;
    const int colsStaged = min(n, j1 + kd) - j0;
    StageBlock(fac, u + static_cast<size_t>(j0) * ld, colsStaged * ld, lane);
    if (active) StageColumn(xs, x, rs, j0, rows, lane);
    __syncthreads();
    if (active) {
      for (int r = rows - 1; r >= 0; --r) {
        const int i = j0 + r;
        double s = xs[r * kWave + lane];
        const int t1 = min(kd, n - 1 - i);
        // U(i, i + t) is column i + t's entry kd - t
        for (int t = t1; t >= 1; --t) s -= fac[(r + t) * ld + kd - t] * slot(i + t);
        const double v = s / fac[r * ld + kd];
        slot(i) = v;
        xs[r * kWave + lane] = v;
      }
      StoreColumn(xs, x, rs, j0, rows, lane);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kWave)
    LuSolveKernel(const BandSystem* systems, const int2* tasks,
                  const double* ab, const int* pivots, double* b) {
  extern __shared__ double ring[];  // RingSlots(kl + ku + 1) x 64
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
    const int R = kl + 1, mask = RingSlots(R) - 1;
    auto slot = [&](int i) -> double& { return ring[(i & mask) * kWave + lane]; };
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
    const int W = kv + 1, mask = RingSlots(W) - 1;
    auto slot = [&](int i) -> double& { return ring[(i & mask) * kWave + lane]; };
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
                         double* b_dev, hipStream_t stream, int panelRows) {
  if (tasks == 0) return;
  ProfileScope prof297("band_solve", stream);
  const int panel = panelRows > 0 ? panelRows : kDefaultPanelRows;
  const size_t ring = RingSlots(kdMax + 1) * kWave + (kSolveRows + kdMax) * (kdMax + 1) +
                      kSolveRows * kWave;
  const size_t fixed = 2 * static_cast<size_t>(panel + kdMax) * (kdMax + 1) + panel;
  const size_t lds = sizeof(double) * std::max(ring, fixed);
  switch (panel) {
    case 4:
      hipLaunchKernelGGL(CholeskySolveKernel<4>, dim3(tasks), dim3(kWave), lds, stream,
                         systems_dev, tasks_dev, ab_dev, b_dev);
      break;
    case 8:
      hipLaunchKernelGGL(CholeskySolveKernel<8>, dim3(tasks), dim3(kWave), lds, stream,
                         systems_dev, tasks_dev, ab_dev, b_dev);
      break;
    case 16:
      hipLaunchKernelGGL(CholeskySolveKernel<16>, dim3(tasks), dim3(kWave), lds, stream,
                         systems_dev, tasks_dev, ab_dev, b_dev);
      break;
    default:
      throw ArgumentError("band solve: panel of 4, 8 or 16 rows");
  }
  SL3_HIP(hipGetLastError());
}

// Each panel: one untimed solve, then five between two events, in place on
// b_dev (the caller's scratch: its values are solved over and over).
int TuneCholeskySolvePanel(const BandSystem* systems_dev, const int2* tasks_dev, int tasks,
                           int kdMax, const double* ab_dev, double* b_dev, hipStream_t stream,
                           double* ms_out) {
  if (tasks == 0) return kDefaultPanelRows;
  hipEvent_t ev[2];
  SL3_HIP(hipEventCreate(&ev[0]));
  SL3_HIP(hipEventCreate(&ev[1]));
  int choice = kDefaultPanelRows;
  double best = 0.0;
  for (int i = 0; i < kPanelChoices; ++i) {
    const int panel = kPanelRowsChoices[i];
    LaunchCholeskySolve(systems_dev, tasks_dev, tasks, kdMax, ab_dev, b_dev, stream, panel);
    SL3_HIP(hipEventRecord(ev[0], stream));
    for (int r = 0; r < 5; ++r) {
      LaunchCholeskySolve(systems_dev, tasks_dev, tasks, kdMax, ab_dev, b_dev, stream, panel);
    }
    SL3_HIP(hipEventRecord(ev[1], stream));
    SL3_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.0f;
    SL3_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (ms_out != nullptr) ms_out[i] = ms / 5.0;
    if (i == 0 || ms < best) {
      best = ms;
      choice = panel;
    }
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return choice;
}

void LaunchLuSolve(const BandSystem* systems_dev, const int2* tasks_dev,
                   int tasks, int kvMax, const double* ab_dev,
                   const int* piv_dev, double* b_dev, hipStream_t stream) {
  if (tasks == 0) return;
  ProfileScope prof308("band_solve", stream);
  hipLaunchKernelGGL(LuSolveKernel, dim3(tasks), dim3(kWave),
                     sizeof(double) * RingSlots(kvMax + 1) * kWave, stream, systems_dev,
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
