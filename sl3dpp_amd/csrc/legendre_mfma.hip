// The Legendre stage of the grid transforms on the FP64 matrix cores
// (gfx950, v_mfma_f64_16x16x4f64).
//
// Per order m of a spin block the stage is a real GEMM between the Wigner
// table T[(m, l)][theta] and the data, with the complex coefficients of a
// node as two real columns:
//   synthesis  F_m[theta][col] = sum_l T[(m, l)][theta] S[(m, l)][col]
//   analysis   S[(m, l)][col]  = sum_theta T[(m, l)][theta] F_m[theta][col]
// (the quadrature weights are applied where the longitude stage writes F).
//
// A workgroup takes one order and one block of 64 columns (32 nodes, 16
// columns per wave). Synthesis keeps every colatitude tile of the block in
// its accumulators and streams the degrees through LDS in chunks of 16, so
// each coefficient is read once; analysis takes 64 degrees per workgroup
// (four 16-degree tiles per wave, few registers, many resident waves) and
// streams the colatitudes in chunks of 16, the row blocks of a column
// block adjacent so the repeated data reads come from L2. The next chunk's
// loads are in flight while the current one is multiplied.
//
// The degree truncation of the radial systems enters as workgroup-uniform
// bounds: a block's nodes carry degrees up to the reach of its last node,
// so synthesis contracts over l <= reach only and analysis computes only
// the row tiles up to it (the rows above, and the entries of nodes below a
// degree's first node, are exact zeros).
//
// The workgroups are numbered so that every column block of one order runs
// on the same XCD (workgroups go to the XCDs round robin): the order's
// table is fetched from HBM once, into that XCD's L2.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "transform.hpp"

namespace sl3amd {

namespace {

constexpr int kWave = 64;
constexpr int kThreads = 256;
constexpr int kXcds = 8;
constexpr int kCols = 64;        // columns per workgroup, 16 per wave
constexpr int kKa = 16;          // analysis: colatitudes per staged chunk
constexpr int kKs = 16;          // synthesis: degrees per staged chunk
// staged elements per thread: analysis data, synthesis data (the tables'
// counts follow the tile shape)
constexpr int kFa = kCols * kKa / kThreads;
constexpr int kSs = kKs * kCols / kThreads;

typedef double Acc4 __attribute__((ext_vector_type(4)));

struct OrderRef {
  int block;
  int m;
};

struct Args {
  const double* table[3];
  const int* rowOffset[3];  // per m - mlo: the row of (m, lmin)
  const double* spectral[3];
  double* spectralOut[3];
  const double* addend[3];
  double* fourier[3];
  const double* fourierIn[3];
  int mlo[3];
  int spin[3];
  const OrderRef* orders;
  const int* reach;      // per node: the largest degree it carries (or null)
  const int* firstNode;  // non-null: analysis writes zeros where a node's reach is below the degree
  int orderCount, colBlocks, rowBlocks, thetaBlocks;
  int lmax, nodes, ntheta;
  int fstride;  // the Fourier layout's colatitude stride
  // the node block: its first node and the spectral rows' node count; the
  // Fourier layout's first order per block
  int nodeOffset, nodesTotal;
  int fourierMlo[3];
  const int* live;  // per realisation, 0: converged, its column blocks skipped
  int liveNodes;
  int keepInactive;  // truncated analysis: entries below a degree's truncation untouched
};

// The largest degree any node of a column block carries (workgroup-uniform
// reads of the per-node reach).
__device__ inline int ColumnBlockReach(const Args& a, int node0, int nodeEnd) {
  int r = 0;
  for (int n = node0; n < nodeEnd; ++n) r = max(r, a.reach[a.nodeOffset + n]);
  return r;
}

// Whether a column block holds a node of a realisation still converging.
__device__ inline bool AnyLive(const Args& a, int node0, int nodeEnd) {
  if (a.live == nullptr) return true;
  const int r0 = (a.nodeOffset + node0) / a.liveNodes;
  const int r1 = (a.nodeOffset + nodeEnd - 1) / a.liveNodes;
  for (int r = r0; r <= r1; ++r) {
    if (a.live[r] != 0) return true;
  }
  return false;
}

__device__ inline int LMinOf(int m, int spin) {
  const int am = m < 0 ? -m : m, an = spin < 0 ? -spin : spin;
  return am > an ? am : an;
}

// Workgroup p -> (order z, column block cb, row block rb: degrees for
// analysis, colatitudes for synthesis): p % 8 is the XCD, shared by all
// blocks of an order, the row blocks of a column block adjacent (its data
// is read from L2 after the first).
__device__ inline bool DecodeAnalysis(const Args& a, int& z, int& cb, int& rb) {
  const int p = blockIdx.x;
  const int xcd = p % kXcds;
  int q = p / kXcds;
  rb = q % a.rowBlocks;
  q /= a.rowBlocks;
  cb = q % a.colBlocks;
  z = (q / a.colBlocks) * kXcds + xcd;
  return z < a.orderCount;
}

// One workgroup: ROWS degrees x 64 columns of one order (each wave 16
// columns, ROWS / 16 16-degree tiles), the colatitudes streamed in chunks
// of 16 with the next chunk's loads in flight under the products. Every
// output sums its colatitudes in the same order whatever ROWS is: the tile
// shape changes the work per workgroup and the data's reuse, not a bit of
// the result.
template <int ROWS>
__global__ void __launch_bounds__(kThreads) LegendreAnalysisMfmaKernel(Args a) {
  constexpr int kRowBlock = ROWS;
  constexpr int kTa = kRowBlock * kKa / kThreads;
  // padded rows: the staging stores run along the first index
  __shared__ double Ts[kKa][kRowBlock + 1];  // [theta][degree row]
  __shared__ double Fs[kKa][kCols + 1];      // [theta][column]
  int z, cb, rb;
  if (!DecodeAnalysis(a, z, cb, rb)) return;
  const OrderRef ord = a.orders[z];
  const int b = ord.block, m = ord.m;
  const int lmin = LMinOf(m, a.spin[b]);
  const int l0 = lmin + rb * kRowBlock;
  if (l0 > a.lmax) return;
  const int lEnd = min(a.lmax, l0 + kRowBlock - 1);  // the block's last degree
  const int cols = 2 * a.nodes;
  const int c0 = cb * kCols;
  const int node0 = c0 / 2, nodeEnd = min(a.nodes, node0 + kCols / 2);
  if (!AnyLive(a, node0, nodeEnd)) return;
  // degrees that reach a node of the column block (the largest reach among
  // its nodes: nodes of several realisations may share a block)
  const int lhi = min(lEnd, a.reach ? ColumnBlockReach(a, node0, nodeEnd) : a.lmax);
  const int rt = lhi >= l0 ? (lhi - l0 + 16) / 16 : 0;  // live 16-degree tiles
  const int row0 = a.rowOffset[b][m - a.mlo[b]] - lmin;  // row of (m, l) = row0 + l
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const double* T = a.table[b];
  const double* F = a.fourierIn[b];
  const size_t mi = static_cast<size_t>(m - a.fourierMlo[b]);
  // the block's columns of the spectral rows
  const size_t rowCols = 2 * static_cast<size_t>(a.nodesTotal);
  double* out = a.spectralOut[b] + 2 * a.nodeOffset;
  const double* add = a.addend[b] ? a.addend[b] + 2 * a.nodeOffset : nullptr;

  Acc4 acc[kRowBlock / 16];
#pragma unroll
  for (int t = 0; t < kRowBlock / 16; ++t) acc[t] = Acc4{0.0, 0.0, 0.0, 0.0};
  if (rt > 0) {
    // a chunk of colatitudes k0 .. k0 + 15 into registers: the table rows
    // of the live degrees (16 threads per row), the block's nodes ((theta,
    // part) pairs contiguous per node)
    double pt[kTa], pf[kFa];
    auto load = [&](int k0) {
#pragma unroll
      for (int q = 0; q < kTa; ++q) {
        const int e = threadIdx.x + kThreads * q;
        const int r = e / kKa, k = e % kKa;
        const int l = l0 + r, th = k0 + k;
        pt[q] = 0.0;
        if (l <= lhi && th < a.ntheta) pt[q] = T[static_cast<size_t>(row0 + l) * a.ntheta + th];
      }
#pragma unroll
      for (int q = 0; q < kFa; ++q) {
        const int e = threadIdx.x + kThreads * q;
        const int nodeLocal = e / (2 * kKa), rem = e % (2 * kKa);
        const int k = rem >> 1, part = rem & 1;
        const int node = node0 + nodeLocal, th = k0 + k;
        pf[q] = 0.0;
        if (node < a.nodes && th < a.ntheta) {
          pf[q] = F[((mi * a.nodes + node) * a.fstride + th) * 2 + part];
        }
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int q = 0; q < kTa; ++q) {
        const int e = threadIdx.x + kThreads * q;
        Ts[e % kKa][e / kKa] = pt[q];
      }
#pragma unroll
      for (int q = 0; q < kFa; ++q) {
        const int e = threadIdx.x + kThreads * q;
        const int nodeLocal = e / (2 * kKa), rem = e % (2 * kKa);
        Fs[rem >> 1][2 * nodeLocal + (rem & 1)] = pf[q];
      }
    };
    load(0);
    for (int k0 = 0; k0 < a.ntheta; k0 += kKa) {
      store();
      __syncthreads();
      if (k0 + kKa < a.ntheta) load(k0 + kKa);
#pragma unroll
      for (int ks = 0; ks < kKa / 4; ++ks) {
        const int k = 4 * ks + lane / 16;
        const double bv = Fs[k][wave * 16 + lane % 16];
#pragma unroll
        for (int t = 0; t < kRowBlock / 16; ++t) {
          if (t < rt) {
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[k][t * 16 + lane % 16], bv, acc[t],
                                                          0, 0, 0);
          }
        }
      }
      __syncthreads();
    }
  }

  // C[row][col]: lane i, element j -> row 4 j + i / 16 of the tile, column
  // i % 16; guards, not early exits, keep the accumulators' indices static.
  // Degrees above the column block's reach, and nodes below a degree's
  // first node, are zeros (plus the addend).
  const int c = c0 + wave * 16 + lane % 16;
  if (c < cols) {
#pragma unroll
    for (int t = 0; t < kRowBlock / 16; ++t) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int l = l0 + t * 16 + 4 * j + lane / 16;
        if (l > lEnd) continue;
        const size_t row = static_cast<size_t>(row0 + l);
        double v = l <= lhi ? acc[t][j] : 0.0;
        if (a.firstNode && a.reach[a.nodeOffset + (c >> 1)] < l) {
          if (a.keepInactive) continue;
          v = 0.0;
        }
        if (add) v += add[row * rowCols + c];
        out[row * rowCols + c] = v;
      }
    }
  }
}

// One workgroup: THETA colatitudes x 64 columns of one order (each wave 16
// columns, THETA / 16 16-colatitude tiles), the degrees up to the block's
// reach streamed in chunks of 16 with the next chunk's loads in flight
// under the products; the colatitude blocks of a column block adjacent on
// one XCD. Every output sums its degrees in the same order whatever THETA
// is.
template <int THETA>
__global__ void __launch_bounds__(kThreads) LegendreSynthesisMfmaKernel(Args a) {
  constexpr int kThetaBlock = THETA;
  constexpr int kTs = kKs * kThetaBlock / kThreads;
  __shared__ double Ts[kKs][kThetaBlock];  // [degree][theta]
  __shared__ double Ss[kKs][kCols];        // [degree][column]
  int z, cb, tb;
  if (!DecodeAnalysis(a, z, cb, tb)) return;
  const OrderRef ord = a.orders[z];
  const int b = ord.block, m = ord.m;
  const int lmin = LMinOf(m, a.spin[b]);
  if (lmin > a.lmax) return;
  const int th0 = tb * kThetaBlock;
  if (th0 >= a.ntheta) return;
  const int cols = 2 * a.nodes;
  const int c0 = cb * kCols;
  const int node0 = c0 / 2, nodeEnd = min(a.nodes, node0 + kCols / 2);
  if (!AnyLive(a, node0, nodeEnd)) return;
  const int lhi = a.reach ? min(a.lmax, ColumnBlockReach(a, node0, nodeEnd)) : a.lmax;
  const int row0 = a.rowOffset[b][m - a.mlo[b]] - lmin;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const double* T = a.table[b];
  // the block's columns of the spectral rows
  const size_t rowCols = 2 * static_cast<size_t>(a.nodesTotal);
  const double* S = a.spectral[b] + 2 * a.nodeOffset;

  // a chunk of degrees l0 .. l0 + 15 into registers: the table rows over
  // the block's colatitudes and the coefficient rows of its columns
  double pt[kTs], ps[kSs];
  auto load = [&](int l0) {
#pragma unroll
    for (int q = 0; q < kTs; ++q) {
      const int e = threadIdx.x + kThreads * q;
      const int k = e / kThetaBlock, th = th0 + e % kThetaBlock;
      const int l = l0 + k;
      pt[q] = 0.0;
      if (l <= lhi && th < a.ntheta) pt[q] = T[static_cast<size_t>(row0 + l) * a.ntheta + th];
    }
#pragma unroll
    for (int q = 0; q < kSs; ++q) {
      const int e = threadIdx.x + kThreads * q;
      const int k = e / kCols, cc = e % kCols;
      const int l = l0 + k, c = c0 + cc;
      ps[q] = 0.0;
      if (l <= lhi && c < cols) ps[q] = S[static_cast<size_t>(row0 + l) * rowCols + c];
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < kTs; ++q) {
      const int e = threadIdx.x + kThreads * q;
      Ts[e / kThetaBlock][e % kThetaBlock] = pt[q];
    }
#pragma unroll
    for (int q = 0; q < kSs; ++q) {
      const int e = threadIdx.x + kThreads * q;
      Ss[e / kCols][e % kCols] = ps[q];
    }
  };

  Acc4 acc[kThetaBlock / 16];
#pragma unroll
  for (int t = 0; t < kThetaBlock / 16; ++t) acc[t] = Acc4{0.0, 0.0, 0.0, 0.0};
  if (lmin <= lhi) load(lmin);
  for (int l0 = lmin; l0 <= lhi; l0 += kKs) {
    store();
    __syncthreads();
    if (l0 + kKs <= lhi) load(l0 + kKs);
#pragma unroll
    for (int ks = 0; ks < kKs / 4; ++ks) {
      const int k = 4 * ks + lane / 16;
      const double bv = Ss[k][wave * 16 + lane % 16];
#pragma unroll
      for (int t = 0; t < kThetaBlock / 16; ++t) {
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[k][t * 16 + lane % 16], bv, acc[t], 0, 0,
                                                      0);
      }
    }
    __syncthreads();
  }

  // C[theta][col] -> F[((m - mlo) nodes + node) ntheta + theta] (re, im),
  // the block's orders and nodes
  double* Fo = a.fourier[b];
  const size_t mi = static_cast<size_t>(m - a.fourierMlo[b]);
  const int c = c0 + wave * 16 + lane % 16;
  if (c < cols) {
    const int node = c >> 1, part = c & 1;
    double* dst = Fo + ((mi * a.nodes + node) * a.fstride) * 2 + part;
#pragma unroll
    for (int t = 0; t < kThetaBlock / 16; ++t) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int th = th0 + t * 16 + 4 * j + lane / 16;
        if (th < a.ntheta) dst[static_cast<size_t>(th) * 2] = acc[t][j];
      }
    }
  }
}

Args MakeArgs(const LegendrePlan& p) {
  Args a{};
  for (int b = 0; b < 3; ++b) {
    a.table[b] = p.table[b];
    a.rowOffset[b] = p.rowOffset[b];
    a.mlo[b] = p.mlo[b];
    a.spin[b] = p.spin[b];
    a.fourierMlo[b] = p.fourierMlo[b];
  }
  a.nodeOffset = p.nodeOffset;
  a.nodesTotal = p.nodesTotal > 0 ? p.nodesTotal : p.nodes;
  a.orders = reinterpret_cast<const OrderRef*>(p.orders);
  a.orderCount = p.orderCount;
  a.colBlocks = CeilDiv(2 * p.nodes, kCols);
  a.lmax = p.lmax;
  a.nodes = p.nodes;
  a.ntheta = p.ntheta;
  a.fstride = p.fstride;
  a.live = p.live;
  a.liveNodes = p.liveNodes;
  a.keepInactive = p.keepInactive ? 1 : 0;
  return a;
}

dim3 GridOf(const Args& a, int rowBlocks) {
  return dim3(CeilDiv(a.orderCount, kXcds) * kXcds * a.colBlocks * rowBlocks);
}

}  // namespace

void LaunchLegendreSynthesisMfma(const LegendrePlan& p, const double* const spectral[3],
                                 double* const fourier[3], const int* reach,
                                 hipStream_t stream) {
  Args a = MakeArgs(p);
  for (int b = 0; b < 3; ++b) {
    a.spectral[b] = spectral[b];
    a.fourier[b] = fourier[b];
  }
  a.reach = reach;
  if (a.orderCount == 0) return;
  const int theta = p.synthesisTheta > 0 ? p.synthesisTheta : kDefaultSynthesisTheta;
  a.thetaBlocks = CeilDiv(p.ntheta, theta);
  a.rowBlocks = a.thetaBlocks;  // the block index the decode splits off
  const dim3 grid = GridOf(a, a.thetaBlocks);
  switch (theta) {
    case 32: hipLaunchKernelGGL(LegendreSynthesisMfmaKernel<32>, grid, dim3(kThreads), 0, stream, a); break;
    case 64: hipLaunchKernelGGL(LegendreSynthesisMfmaKernel<64>, grid, dim3(kThreads), 0, stream, a); break;
    case 112: hipLaunchKernelGGL(LegendreSynthesisMfmaKernel<112>, grid, dim3(kThreads), 0, stream, a); break;
    case 208: hipLaunchKernelGGL(LegendreSynthesisMfmaKernel<208>, grid, dim3(kThreads), 0, stream, a); break;
    default: throw ArgumentError("Legendre synthesis: colatitude block 32, 64, 112 or 208");
  }
  SL3_HIP(hipGetLastError());
}

void LaunchLegendreAnalysisMfma(const LegendrePlan& p, const double* const fourier[3],
                                double* const spectral[3], const double* const addend[3],
                                const int* reach, const int* firstNode, hipStream_t stream) {
  Args a = MakeArgs(p);
  for (int b = 0; b < 3; ++b) {
    a.fourierIn[b] = fourier[b];
    a.spectralOut[b] = spectral[b];
    a.addend[b] = addend ? addend[b] : nullptr;
  }
  a.reach = reach;
  a.firstNode = firstNode;
  if (a.orderCount == 0) return;
  const int rows = p.analysisRows > 0 ? p.analysisRows : kDefaultAnalysisRows;
  a.rowBlocks = CeilDiv(p.lmax + 1, rows);
  const dim3 grid = GridOf(a, a.rowBlocks);
  switch (rows) {
    case 16: hipLaunchKernelGGL(LegendreAnalysisMfmaKernel<16>, grid, dim3(kThreads), 0, stream, a); break;
    case 32: hipLaunchKernelGGL(LegendreAnalysisMfmaKernel<32>, grid, dim3(kThreads), 0, stream, a); break;
    case 64: hipLaunchKernelGGL(LegendreAnalysisMfmaKernel<64>, grid, dim3(kThreads), 0, stream, a); break;
    case 128: hipLaunchKernelGGL(LegendreAnalysisMfmaKernel<128>, grid, dim3(kThreads), 0, stream, a); break;
    default: throw ArgumentError("Legendre analysis: degree block 16, 32, 64 or 128");
  }
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
