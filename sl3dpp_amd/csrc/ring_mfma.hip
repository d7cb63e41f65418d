// The longitude stage of the grid transforms on the FP64 matrix cores
// (gfx950, v_mfma_f64_16x16x4f64), for grids with nphi = R P, P odd,
// R = 2 or 4.
//
// With m = R m1 + m2 and n = n1 + P n2 (m2, n2 < R; m1, n1 < P),
//   x[n1 + P n2] = sum_m2 e^{2 pi i m2 n2 / R} e^{2 pi i m2 n1 / nphi} Y_m2[n1],
//   Y_m2[n1]     = sum_m1 F_{R m1 + m2} e^{2 pi i m1 n1 / P},
// so synthesis reads each sequence z_m2[m1] = F_{R m1 + m2} straight from
// the staged coefficients (no sums, no twiddles), runs the P-point
// transforms on the matrix cores, and finishes with the twiddle and an
// R-point transform across the R sequences of a ring, which sit in
// neighbouring lanes (register butterflies over lane ^ 4, ^ 2 and the part
// swap ^ 1). Analysis mirrors it: R-point butterflies and conjugate
// twiddles at the grid points, the P-point forward transforms, and each
// output G_m2[m1] is F_{R m1 + m2} itself, scattered to an LDS tile and
// copied out with consecutive rings in consecutive lanes.
//
// Each P-point transform is taken in the symmetric form
//   Y_j = h_0 + sum_k (h_k + h_{P-k}) cos(2 pi jk/P) + i (h_k - h_{P-k}) sin(..)
// (H = (P-1)/2 frequencies, padded to 16): products of the constant cos
// and sin matrices (A operands, staged in LDS) with the sequences (B
// operands). One wave holds 8 complex sequences as the 16 real columns of
// a tile (column 2 s + part, s = ring-in-wave x R + m2), so the inverse
// transform's accumulators are the forward transform's operands and the
// pointwise product between them never leaves the registers. The spin-0
// block packs two real rings into one complex sequence.
//
// The band limit |m| <= L leaves h_k = h_{P-k} = 0 for k > kmax =
// floor((L + R - 1) / R): the inverse contracts over k <= kmax only and the
// forward computes only the row tiles that hold orders it stores. Grids
// whose H fills the 16 JT rows exactly (FULL) drop the per-point padding
// checks.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "ring_complex.hpp"
#include "transform.hpp"

namespace sl3amd {

namespace {

constexpr int kWave = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kWave * kWaves;
constexpr int kSeq = 8;  // complex sequences per wave

typedef double Acc4 __attribute__((ext_vector_type(4)));

__device__ inline double Shfl(double v, int mask) { return __shfl_xor(v, mask, kWave); }

// Lane exchanges inside a quad (DPP quad_perm: xor 1 = [1,0,3,2], xor 2 =
// [2,3,0,1]) and inside 32 lanes (ds_swizzle bit mode, xor 4), without an
// address register or a round trip through LDS memory.
template <int CTRL>
__device__ inline double QuadPerm(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double Xor1(double v) { return QuadPerm<0xB1>(v); }
__device__ inline double Xor2(double v) { return QuadPerm<0x4E>(v); }
__device__ inline double Xor4(double v) {
  constexpr int kPattern = 0x1F | (4 << 10);  // and 0x1f, or 0, xor 4
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), kPattern);
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), kPattern);
  return __hiloint2double(hi, lo);
}

// The sum over the four lanes of a column (lane, lane ^ 16, ^ 32, ^ 48),
// the same value in each of them.
__device__ inline double ColumnSum(double v) {
  v += Shfl(v, 16);
  v += Shfl(v, 32);
  return v;
}

// The staged rows' zero padding beyond +-lmax: the inverse reads orders up
// to R 4 kIn + R - 1 away from order 0.
__host__ __device__ inline int RowPad(int lmax, int R, int P) {
  const int H = (P - 1) / 2;
  const int kmax = H < (lmax + R - 1) / R ? H : (lmax + R - 1) / R;
  const int kIn = (kmax + 3) / 4;
  const int pad = R * 4 * kIn + R - 1 - lmax;
  return pad > 0 ? pad : 0;
}

struct Geometry {
  int R, P, H, nphi, L, mlo, orders;
  int kIn, tOut;  // the band limit: contraction blocks of the inverse, row tiles of the forward
  int ringsPerWave, ringsPerGroup;  // rings of the grid
  // The staged rows (one per sequence slot: a ring, or a spin-0 ring pair):
  // orders -(L + pad)..L + pad at L + pad + m, zero beyond +-L.
  int rows, pad, stride;
};

template <bool REAL>
__device__ inline Geometry MakeGeometry(const RingLaunch& a) {
  Geometry g;
  g.R = a.tables.R;
  g.P = a.tables.P;
  g.H = (g.P - 1) / 2;
  g.nphi = a.nphi;
  g.L = a.lmax;
  g.mlo = REAL ? 0 : -a.lmax;
  g.orders = g.L - g.mlo + 1;
  g.ringsPerWave = (REAL ? 2 * kSeq : kSeq) / g.R;
  g.ringsPerGroup = kWaves * g.ringsPerWave;
  g.rows = kWaves * kSeq / g.R;
  g.pad = RowPad(g.L, g.R, g.P);
  g.stride = 2 * (g.L + g.pad) + 1;
  const int kmax = min(g.H, (g.L + g.R - 1) / g.R);
  g.kIn = (kmax + 3) / 4;
  g.tOut = (kmax + 15) / 16;
  return g;
}

// Per-lane indexing: column col = 2 s + p, sequence s = gl R + m2.
struct Lane {
  int lane, col, quad, s, p, m2, a, b;
  int n2;      // the lane's points are n1 + P n2 after the butterflies
  int slot;    // the sequence's ring (spin 0: ring pair) within the group
  int ring;    // the lane's own grid ring within the group
  bool valid;  // that ring exists
  // signs that turn the butterflies' role selects into fused multiply-adds:
  // sP -1 for the real part; sA, sB +1 for the sum half of the R-point
  // stages, -1 for the difference; cI1, cI2 the x i of the (a, b) = (1, 1)
  // lanes (keep the value, take the swapped part)
  double sP, sA, sB, cI1, cI2;
};

template <bool REAL>
__device__ inline Lane MakeLane(const Geometry& g, int th0, int ntheta) {
  Lane l;
  const int wave = threadIdx.x / kWave;
  l.lane = threadIdx.x % kWave;
  l.col = l.lane % 16;
  l.quad = l.lane / 16;
  l.s = l.col / 2;
  l.p = l.col % 2;
  l.m2 = l.s % g.R;
  l.a = l.m2 >> 1;
  l.b = l.m2 & 1;
  l.n2 = g.R == 4 ? l.a + 2 * l.b : l.b;
  l.slot = wave * (kSeq / g.R) + l.s / g.R;
  l.ring = REAL ? 2 * l.slot + l.p : l.slot;
  l.valid = th0 + l.ring < ntheta;
  l.sP = l.p == 0 ? -1.0 : 1.0;
  l.sA = l.a == 0 ? 1.0 : -1.0;
  l.sB = l.b == 0 ? 1.0 : -1.0;
  const bool rot = g.R == 4 && l.a == 1 && l.b == 1;
  l.cI1 = rot ? 0.0 : 1.0;
  l.cI2 = rot ? l.sP : 0.0;
  return l;
}

// Synthesis at n1: V = e^{2 pi i m2 n1 / nphi} Y, then the R-point inverse
// transform across the lanes of the ring; returns the lane's part of
// x[n1 + P n2].
__device__ inline double InverseButterfly(const Geometry& g, const Lane& ln, double y,
                                          double c, double s) {
  const double v = fma(c, y, (ln.sP * s) * Xor1(y));
  if (g.R == 4) {
    const double e0 = fma(ln.sA, v, Xor4(v));
    const double e = fma(ln.cI2, Xor1(e0), ln.cI1 * e0);  // x i on the (1, 1) lanes
    return fma(ln.sB, e, Xor2(e));
  }
  return fma(ln.sB, v, Xor2(v));
}

// Analysis at n1: the R-point forward transform across the lanes of the
// ring, then the conjugate twiddle; returns the lane's part of U_m2[n1].
__device__ inline double ForwardButterfly(const Geometry& g, const Lane& ln, double y,
                                          double c, double s) {
  double w;
  if (g.R == 4) {
    const double d0 = fma(ln.sB, y, Xor2(y));
    const double d = fma(-ln.cI2, Xor1(d0), ln.cI1 * d0);  // x (-i) on the (1, 1) lanes
    w = fma(ln.sA, d, Xor4(d));
  } else {
    w = fma(ln.sB, y, Xor2(y));
  }
  return fma(c, w, (-ln.sP * s) * Xor1(w));
}

// The P-point inverse transforms of the lane's sequence on the matrix
// cores, from its staged row (the lane's part, order m2 at offset 0): Y at j
// and P - j, and at 0. The sequence's element k is the order m2 + R k, its
// element P - k the order m2 - R k: reads 2 R k doubles up and down the
// row, zero beyond +-L (and so for k past kmax or H).
template <int JT, bool REAL, bool FULL>
__device__ inline void InverseSequences(const Geometry& g, const Lane& ln, const double* row,
                                        const double* tabC, const double* tabS,
                                        double (&yj)[4 * JT], double (&yq)[4 * JT], double& y0) {
  constexpr int KS = 4 * JT;
  double av[KS], bv[KS];
  double sum = 0.0;
  const int R = g.R;
  const double* up = row + 2 * R * (1 + ln.quad);  // k = 1 + 4 kb + quad
  const double* dn = row - 2 * R * (1 + ln.quad);
#pragma unroll
  for (int kb = 0; kb < KS; ++kb) {
    if (kb >= g.kIn) {
      av[kb] = bv[kb] = 0.0;
      continue;
    }
    const double zk = up[8 * R * kb], zq = dn[-8 * R * kb];
    av[kb] = zk + zq;
    bv[kb] = zk - zq;
    sum += av[kb];
  }
  const double h0 = row[0];
  y0 = h0 + ColumnSum(sum);
  Acc4 C[JT], S[JT];
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    C[t] = Acc4{0.0, 0.0, 0.0, 0.0};
    S[t] = Acc4{0.0, 0.0, 0.0, 0.0};
  }
  // 2 JT independent accumulation chains, interleaved
#pragma unroll
  for (int kb = 0; kb < KS; ++kb) {
    if (kb >= g.kIn) break;
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int at = (t * KS + kb) * kWave + ln.lane;
      C[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabC[at], av[kb], C[t], 0, 0, 0);
      S[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabS[at], bv[kb], S[t], 0, 0, 0);
    }
  }
  // Y_j = h0 + C_j + i S_j, Y_{P-j} = h0 + C_j - i S_j
  const double sg = ln.p == 0 ? -1.0 : 1.0;
#pragma unroll
  for (int t = 0; t < JT; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double sx = Xor1(S[t][q]);
      yj[4 * t + q] = h0 + C[t][q] + sg * sx;
      yq[4 * t + q] = h0 + C[t][q] - sg * sx;
    }
  }
}

// The forward transforms of the lane's sequence (u at j, P - j and 0,
// after the forward butterflies) on the matrix cores; each output G(k) is
// the order m2 + R k (G(P - k): m2 - R k), written to the lane's part of its
// staged row (order m2 at offset 0) where |order| <= L.
template <int JT, bool FULL>
__device__ inline void ForwardSequences(const Geometry& g, const Lane& ln, const double* tabC,
                                        const double* tabS, const double (&uj)[4 * JT],
                                        const double (&uq)[4 * JT], double u0, double* row) {
  constexpr int KS = 4 * JT;
  double av[KS], bv[KS];
  double sum = 0.0;
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    av[v] = uj[v] + uq[v];
    bv[v] = uj[v] - uq[v];
    sum += av[v];
  }
  const double g0 = u0 + ColumnSum(sum);
  Acc4 C[JT], S[JT];
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    C[t] = Acc4{0.0, 0.0, 0.0, 0.0};
    S[t] = Acc4{0.0, 0.0, 0.0, 0.0};
  }
#pragma unroll
  for (int kb = 0; kb < KS; ++kb) {
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      if (t >= g.tOut) break;
      const int at = (t * KS + kb) * kWave + ln.lane;
      C[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabC[at], av[kb], C[t], 0, 0, 0);
      S[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabS[at], bv[kb], S[t], 0, 0, 0);
    }
  }
  // G(k) = u0 + C'_k - i S'_k, G(P - k) = u0 + C'_k + i S'_k
  if (ln.quad == 0) row[0] = g0;
  const int R = g.R;
  double* up = row + 2 * R * (1 + ln.quad);
  double* dn = row - 2 * R * (1 + ln.quad);
  const int reach = g.L - R * (1 + ln.quad);  // R (16 t + 4 q) -+ m2 may not pass it
  const double sg = ln.p == 0 ? 1.0 : -1.0;
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    if (t >= g.tOut) break;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int off = R * (16 * t + 4 * q);
      const double sx = sg * Xor1(S[t][q]);
      const double uc = u0 + C[t][q];
      if (off + ln.m2 <= reach) up[2 * off] = uc + sx;
      if (off - ln.m2 <= reach) dn[-2 * off] = uc - sx;
    }
  }
}

// The staged tile to the Fourier layout F[m - mlo][node][theta], scaled by
// (2 pi / nphi) w_theta; consecutive threads take consecutive rings. Spin
// 0: F1_m = (Z_m + conj Z_{-m}) / 2, F2_m = (Z_m - conj Z_{-m}) / 2i.
template <bool REAL>
__device__ inline void CopyOut(const Geometry& g, const RingLaunch& a, const double2* Zs,
                               int node, int th0, double* out) {
  constexpr int kBatch = 9;
  const double scale = 2.0 * 3.14159265358979323846 / g.nphi;
  const int mCount = g.L - g.mlo + 1;
  // element e = thread + 256 i: ring e % rings (fixed per thread), order e / rings
  const int rpg = g.ringsPerGroup;
  const int ring = static_cast<int>(threadIdx.x) % rpg;
  const int moStep = kThreads / rpg;
  const int th = th0 + ring;
  if (th >= a.ntheta) return;
  const double w = scale * a.wtheta[th];
  double2* dst = reinterpret_cast<double2*>(out) + static_cast<size_t>(node) * a.fstride + th;
  const size_t mStride = static_cast<size_t>(a.nodes) * a.fstride;
  for (int mo0 = static_cast<int>(threadIdx.x) / rpg; mo0 < mCount; mo0 += kBatch * moStep) {
    double2 res[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int mo = min(mo0 + u * moStep, mCount - 1);
      if (REAL) {
        const double2* row = Zs + (ring / 2) * g.stride + (g.L + g.pad);
        const double2 zp = row[mo];
        const double2 zm = row[-mo];
        res[u] = ring % 2 == 0 ? make_double2(0.5 * (zp.x + zm.x), 0.5 * (zp.y - zm.y))
                               : make_double2(0.5 * (zp.y + zm.y), -0.5 * (zp.x - zm.x));
      } else {
        res[u] = Zs[ring * g.stride + g.pad + mo];  // order mo - L at L + pad + order
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int mo = mo0 + u * moStep;
      if (mo < mCount) dst[mo * mStride] = make_double2(res[u].x * w, res[u].y * w);
    }
  }
}

// The op classes, one kernel each so that each carries only its own live
// registers: synthesis only (synthesise, finish), one analysis (analyse,
// weighted, node scaled, pass), the step start (two analyses).
enum { kClassSynth = 0, kClassOne = 1, kClassStart = 2 };

// One workgroup: the rings th0 .. th0 + rings - 1 of one node.
template <int JT, bool REAL, int CLASS, bool FULL>
__device__ inline void RingBody(const RingLaunch& a, int node, int th0) {
  constexpr int KS = 4 * JT;
  extern __shared__ double lds[];
  const Geometry g = MakeGeometry<REAL>(a);
  const Lane ln = MakeLane<REAL>(g, th0, a.ntheta);
  // The lane's points: n1 = j = 1 + 16 t + 4 q + quad (v = 4 t + q), P - j
  // and 0, each at n1 + P n2. Where a lane has no point (padding, a ring
  // past the last) the index is its ring's point P n2, always in bounds:
  // loads are unconditional and issue back to back, only stores are
  // predicated. For spin 0 a sequence's two parts are two rings: the lane
  // holds its own ring's value at every point.
  const int th = th0 + ln.ring;
  const size_t ringBase = (static_cast<size_t>(node) * a.ntheta + (ln.valid ? th : 0)) * g.nphi;
  const size_t p0 = ringBase + static_cast<size_t>(g.P) * ln.n2;
  auto jOf = [&](int v) { return 1 + 16 * (v / 4) + 4 * (v % 4) + ln.quad; };
  auto lv = [&](int v) { return ln.valid && (FULL || jOf(v) <= g.H); };
  auto pj = [&](int v) -> size_t { return FULL || lv(v) ? p0 + jOf(v) : p0; };
  auto pq = [&](int v) -> size_t { return FULL || lv(v) ? p0 + (g.P - jOf(v)) : p0; };
  auto element = [&](size_t pt) -> size_t { return REAL ? pt : 2 * pt + ln.p; };
  // the twiddle index m2 n1 at n1 = j (clamped to H) and P - j
  auto twj = [&](int v) { return ln.m2 * (FULL ? jOf(v) : min(jOf(v), g.H)); };
  auto twq = [&](int v) { return ln.m2 * (g.P - (FULL ? jOf(v) : min(jOf(v), g.H))); };
  const bool own0 = ln.valid && ln.quad == 0;  // the lane that stores point P n2

  // the pass's and the weighted analysis' weights first: they stream in
  // under the staging and the inverse transforms
  const bool weighted = CLASS == kClassOne && (a.op == kRingPass || a.op == kRingWeighted);
  double wj[KS], wq[KS], w0 = 0.0;
  if (weighted) {
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      wj[v] = a.weight[pj(v)];
      wq[v] = a.weight[pq(v)];
    }
    w0 = a.weight[p0];
  }
  // LDS: the twiddles cos / sin(2 pi n / nphi), the operand tables, then
  // one tile that holds the staged input coefficients and, after them, the
  // staged output orders.
  constexpr bool kTabLds = true;
  constexpr int kTab = JT * KS * kWave;
  double* cosN = lds;
  double* sinN = lds + g.nphi;
  double* tables = lds + 2 * g.nphi;
  double* stageD = tables + (kTabLds ? 2 * kTab : 0);
  double2* stage = reinterpret_cast<double2*>(stageD);
  // the lane's part of its staged row, order m2 at offset 0
  double* row = stageD + 2 * (ln.slot * g.stride + g.L + g.pad + ln.m2) + ln.p;
  const double* tabC = a.tables.tabC;
  const double* tabS = a.tables.tabS;
  // Staging: every copy issues all of a thread's loads before its stores,
  // so their latencies overlap instead of adding up.
  for (int i0 = 0; i0 < g.nphi; i0 += 2 * kThreads) {
    double c[2], sn[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + threadIdx.x + u * kThreads;
      c[u] = i < g.nphi ? a.tables.cosN[i] : 0.0;
      sn[u] = i < g.nphi ? a.tables.sinN[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + threadIdx.x + u * kThreads;
      if (i < g.nphi) {
        cosN[i] = c[u];
        sinN[i] = sn[u];
      }
    }
  }
  if constexpr (kTabLds) {
    constexpr int kPerThread = kTab / kThreads;  // JT^2: the tables hold 256 JT^2 values
    double tc[kPerThread], ts[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      tc[u] = a.tables.tabC[threadIdx.x + u * kThreads];
      ts[u] = a.tables.tabS[threadIdx.x + u * kThreads];
    }
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      tables[threadIdx.x + u * kThreads] = tc[u];
      tables[kTab + threadIdx.x + u * kThreads] = ts[u];
    }
    tabC = tables;
    tabS = tables + kTab;
  }
  const bool synth = a.op == kRingSynthesise || a.op == kRingPass ||
                     a.op == kRingFinish || a.op == kRingStartFromSpectral;
  if (synth) {
    // element e = thread + 256 i: ring e % rings (the same for all of a
    // thread's elements, rings dividing 256), stored order e / rings
    constexpr int kBatch = 9;
    const int rpg = g.ringsPerGroup;
    const int count = g.orders * rpg;
    const int ring = static_cast<int>(threadIdx.x) % rpg;
    const int moStep = kThreads / rpg;
    const int th = th0 + ring;
    const double2* F = reinterpret_cast<const double2*>(a.fourierIn);
    for (int e0 = 0; e0 < count; e0 += kBatch * kThreads) {
      const int mo0 = (e0 + static_cast<int>(threadIdx.x)) / rpg;
      double2 v[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int mo = mo0 + u * moStep;
        v[u] = make_double2(0.0, 0.0);
        if (mo < g.orders && th < a.ntheta) {
          v[u] = F[(static_cast<size_t>(mo) * a.nodes + node) * a.fstride + th];
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int mo = mo0 + u * moStep;
        if (REAL) {
          // the ring pair's sequence Z_m = F1_m + i F2_m at m and -m
          // (F_{-m} = conj F_m, the imaginary part at m = 0 dropped); the
          // partner ring is the neighbouring lane, active whenever this is
          const double ox = Xor1(v[u].x), oy = Xor1(v[u].y);
          const bool even = ring % 2 == 0;
          const double2 f1 = even ? v[u] : make_double2(ox, oy);
          const double2 f2 = even ? make_double2(ox, oy) : v[u];
          double2* row = stage + (ring / 2) * g.stride + (g.L + g.pad);
          if (mo < g.orders) {
            if (mo == 0) {
              if (even) row[0] = make_double2(f1.x, f2.x);
            } else if (even) {
              row[mo] = make_double2(f1.x - f2.y, f1.y + f2.x);
            } else {
              row[-mo] = make_double2(f1.x + f2.y, f2.x - f1.y);
            }
          }
        } else if (mo < g.orders) {
          stage[ring * g.stride + g.pad + mo] = v[u];  // order mo - L at L + pad + order
        }
      }
    }
    // the rows' padding beyond +-L is zero
    for (int i = threadIdx.x; i < 2 * g.pad * g.rows; i += kThreads) {
      const int slot = i / (2 * g.pad), k = i % (2 * g.pad);
      stage[slot * g.stride + (k < g.pad ? k : 2 * g.L + 1 + k)] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();


  // x at the lane's points
  double xj[KS], xq[KS], x0 = 0.0;
  if (synth) {
    double yj[KS], yq[KS], y0;
    InverseSequences<JT, REAL, FULL>(g, ln, row, tabC, tabS, yj, yq, y0);
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      xj[v] = InverseButterfly(g, ln, yj[v], cosN[twj(v)], sinN[twj(v)]);
      xq[v] = InverseButterfly(g, ln, yq[v], cosN[twq(v)], sinN[twq(v)]);
    }
    x0 = InverseButterfly(g, ln, y0, 1.0, 0.0);
  } else {
    // a stored ring (complex: interleaved parts; real: the lane's own ring)
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      const double vj = a.gridIn[element(pj(v))], vq = a.gridIn[element(pq(v))];
      xj[v] = lv(v) ? vj : 0.0;
      xq[v] = lv(v) ? vq : 0.0;
    }
    const double v0 = a.gridIn[element(p0)];
    x0 = ln.valid ? v0 : 0.0;
  }
  __syncthreads();  // the staged coefficients are read; the tile is reused

  const double2* Zs = stage;
  // y at the lane's points -> forward butterflies -> forward transforms
  auto forward = [&](const double (&yj)[KS], const double (&yq)[KS], double y0, double* out,
                     const double* cosT, const double* sinT) {
    double uj[KS], uq[KS];
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      uj[v] = ForwardButterfly(g, ln, yj[v], cosT[twj(v)], sinT[twj(v)]);
      uq[v] = ForwardButterfly(g, ln, yq[v], cosT[twq(v)], sinT[twq(v)]);
    }
    const double u0 = ForwardButterfly(g, ln, y0, 1.0, 0.0);
    ForwardSequences<JT, FULL>(g, ln, tabC, tabS, uj, uq, u0, row);
    __syncthreads();
    CopyOut<REAL>(g, a, Zs, node, th0, out);
  };

  if constexpr (CLASS == kClassSynth) {
    if (a.op == kRingFinish) {
      // d on the grid and m+ = m~ + beta d in place
      double bj[KS], bq[KS], mj[KS], mq[KS];
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        bj[v] = a.beta[pj(v)];
        bq[v] = a.beta[pq(v)];
        mj[v] = a.memOut[element(pj(v))];
        mq[v] = a.memOut[element(pq(v))];
      }
      const double b0 = a.beta[p0], m0 = a.memOut[element(p0)];
      if (own0) {
        a.gridOut[element(p0)] = x0;
        a.memOut[element(p0)] = m0 + b0 * x0;
      }
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        if (lv(v)) {
          a.gridOut[element(pj(v))] = xj[v];
          a.gridOut[element(pq(v))] = xq[v];
          a.memOut[element(pj(v))] = mj[v] + bj[v] * xj[v];
          a.memOut[element(pq(v))] = mq[v] + bq[v] * xq[v];
        }
      }
    } else {
      if (own0) a.gridOut[element(p0)] = x0;
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        if (lv(v)) {
          a.gridOut[element(pj(v))] = xj[v];
          a.gridOut[element(pq(v))] = xq[v];
        }
      }
    }
  } else if constexpr (CLASS == kClassStart) {
    // src/Viscoelastic3D.cpp:1283-1332, one branch: the weights of
    // h = dt itau(x), m~ = e m + alpha d0, beta, W = mu (beta - beta_ref),
    // the fixed force mu m~; then W d0 (:1361-1365). x holds d0.
    const double mu = a.muNode[node];
    const double href = a.dt * a.itauRefNode[node];
    auto phi1 = [](double h) {
      return fabs(h) < 1e-3 ? 1.0 - h / 2.0 + h * h / 6.0 - h * h * h / 24.0 +
                                  h * h * h * h / 120.0
                            : (1.0 - exp(-h)) / h;
    };
    const double refBeta =
        a.scheme == 1 ? 1.0 - 1.0 / (1.0 + href) : 1.0 - phi1(href);
    // every point's itau and m first (unconditional loads, issued together),
    // after the inverse transforms: held through them they would spill
    __builtin_amdgcn_sched_barrier(0);
    double hj[KS], hq[KS], mInJ[KS], mInQ[KS];
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      hj[v] = a.itau[pj(v)];
      hq[v] = a.itau[pq(v)];
      mInJ[v] = a.memIn[element(pj(v))];
      mInQ[v] = a.memIn[element(pq(v))];
    }
    const double h0 = a.itau[p0], mIn0 = a.memIn[element(p0)];
    // returns W d0 and writes m~, (spin 0) beta and W; mt gets m~
    auto start = [&](size_t pt, double itau, double mIn, double d0, bool store,
                     double& mt) -> double {
      const double h = a.dt * itau;
      double ew, aw, bw;
      if (a.scheme == 1) {
        const double w = 1.0 / (1.0 + h);
        ew = w;
        aw = 0.0;
        bw = 1.0 - w;
      } else {
        const double e = exp(-h);
        const double p = phi1(h);
        ew = e;
        aw = p - e;
        bw = 1.0 - p;
      }
      const double mNew = ew * mIn + aw * d0;
      if (store) {
        a.memOut[element(pt)] = mNew;
        if (REAL) {
          a.betaOut[pt] = bw;
          a.weightOut[pt] = mu * (bw - refBeta);
        }
      }
      mt = mNew;
      return mu * (bw - refBeta) * d0;
    };
    double yj[KS], yq[KS], y0 = 0.0, m0 = 0.0, unused;
    y0 = start(p0, h0, mIn0, x0, own0, m0);
    if (!ln.valid) y0 = m0 = 0.0;
    m0 *= mu;
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      yj[v] = start(pj(v), hj[v], mInJ[v], xj[v], lv(v), unused);
      yq[v] = start(pq(v), hq[v], mInQ[v], xq[v], lv(v), unused);
      if (!lv(v)) yj[v] = yq[v] = 0.0;
    }
    // the analysis of W d0 (:1361-1376) ...
    int fresh1 = 0;  // the twiddles read again, not held from the synthesis
    asm volatile("" : "+v"(fresh1));
    forward(yj, yq, y0, a.fourierOut2, cosN + fresh1, sinN + fresh1);
    __syncthreads();
    // ... and of the fixed memory force mu m~ (:1355-1356): m~ at the
    // lane's points j and P - j read back from the lane's own stores (in the
    // cache) rather than held in registers through the first analysis; at
    // P n2, which one lane of the column stores, every lane kept its value.
    double mj[KS], mq[KS];
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      mj[v] = lv(v) ? mu * a.memOut[element(pj(v))] : 0.0;
      mq[v] = lv(v) ? mu * a.memOut[element(pq(v))] : 0.0;
    }
    // the twiddles read again (an opaque offset), not held from the first
    int fresh = 0;
    asm volatile("" : "+v"(fresh));
    forward(mj, mq, m0, a.fourierOut, cosN + fresh, sinN + fresh);
  } else {
    // y = factor x (pass and weighted: the point's weight; node scaled: mu)
    double yj[KS], yq[KS], y0 = x0;
    if (weighted) {
      y0 = ln.valid ? w0 * x0 : 0.0;
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        yj[v] = lv(v) ? wj[v] * xj[v] : 0.0;
        yq[v] = lv(v) ? wq[v] * xq[v] : 0.0;
      }
    } else if (a.op == kRingNodeScaled) {
      const double mu = a.muNode[node];
      y0 = mu * x0;
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        yj[v] = mu * xj[v];
        yq[v] = mu * xq[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        yj[v] = xj[v];
        yq[v] = xq[v];
      }
    }
    forward(yj, yq, y0, a.fourierOut, cosN, sinN);
  }
}

// Every class runs two waves per SIMD (two workgroups per CU, as the LDS
// tile allows): a register budget of 256 per lane. The step start fits it
// by reading m~ back for its second analysis instead of holding it. Five
// tiles (64 < H <= 80: lmax 192's P = 145) hold a fifth of each per-point
// array more and their tables fill most of the LDS, so one workgroup per CU
// and one wave per SIMD: the budget is 512.
template <int JT>
constexpr int WavesPerEu() { return JT >= 5 ? 1 : 2; }
template <int JT, bool REAL, int CLASS, bool FULL>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WavesPerEu<JT>())))
    RingMfmaKernel(RingLaunch a) {
  if (a.live && a.live[blockIdx.y / a.liveNodes] == 0) return;  // a converged realisation
  RingBody<JT, REAL, CLASS, FULL>(a, blockIdx.y, blockIdx.x * MakeGeometry<REAL>(a).ringsPerGroup);
}

template <int JT, bool REAL, bool FULL>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WavesPerEu<JT>())))
    RingMfmaStartKernel(RingLaunch a) {
  RingBody<JT, REAL, kClassStart, FULL>(a, blockIdx.y,
                                        blockIdx.x * MakeGeometry<REAL>(a).ringsPerGroup);
}

// The pass over all three blocks in one launch. A super-tile is one node's
// 16 rings (R = 4): one workgroup of each complex block with whole complex
// values per lane (ring_complex.hpp) and one 16-ring workgroup of spin 0 on
// the body above, the three numbered onto one XCD (workgroups go to the XCDs
// round robin) so the rings' weights come from its L2 after the first read;
// one launch also leaves a single tail instead of three.
constexpr int kPassGroups = 3;

template <int JT, bool FULL, int KIN>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WavesPerEu<JT>())))
    RingMfmaPassKernel(RingLaunch base, RingBlocks blocks) {
  const int p = blockIdx.x;
  const int st = (p / (8 * kPassGroups)) * 8 + p % 8;  // the super-tile
  const int sub = (p / 8) % kPassGroups;
  const int groups = (base.ntheta + 15) / 16;
  if (st >= groups * base.nodes) return;
  const int node = st / groups, th16 = (st % groups) * 16;
  if (base.live && base.live[node / base.liveNodes] == 0) return;  // a converged realisation
  if (sub < 2) {
    complex_lanes::ComplexPassBody<JT, FULL, KIN>(
        base, reinterpret_cast<const double2*>(blocks.in[sub]),
        reinterpret_cast<double2*>(blocks.out[sub]), node, th16);
  } else {
    RingLaunch a = base;
    a.block = 2;
    a.fourierIn = blocks.in[2];
    a.fourierOut = blocks.out[2];
    RingBody<JT, true, kClassOne, FULL>(a, node, th16);
  }
}

// The step start of a complex block from a stored strain grid on the
// complex-lane body: m~ and the analysis of W d0, then that of mu m~.
template <int JT, bool FULL, bool START>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WavesPerEu<JT>())))
    RingStartComplexKernel(RingLaunch a) {
  complex_lanes::ComplexStartBody<JT, FULL, START>(a, blockIdx.y,
                                                   blockIdx.x * complex_lanes::kRingsPerGroup);
}

template <int JT, int CLASS, bool FULL>
void LaunchClass(const RingLaunch& r, dim3 grid, size_t lds, hipStream_t stream) {
  if constexpr (CLASS == kClassStart) {
    if (r.block != 2 && r.op == kRingStart && r.tables.R == 4 && r.lmax >= 3) {
      // 16 rings of one node per workgroup, as the staged body's grid for a
      // complex block; LDS: the twiddles and the operand tables
      const size_t ldsC = sizeof(double) * (2 * static_cast<size_t>(r.nphi) +
                                            2 * static_cast<size_t>(JT) * 4 * JT * kWave);
      const dim3 gridC(CeilDiv(r.ntheta, complex_lanes::kRingsPerGroup), r.nodes);
      hipLaunchKernelGGL((RingStartComplexKernel<JT, FULL, true>), gridC, dim3(kThreads), ldsC,
                         stream, r);
      SL3_HIP(hipGetLastError());
      hipLaunchKernelGGL((RingStartComplexKernel<JT, FULL, false>), gridC, dim3(kThreads), ldsC,
                         stream, r);
      return;
    }
    if (r.block == 2) {
      hipLaunchKernelGGL((RingMfmaStartKernel<JT, true, FULL>), grid, dim3(kThreads), lds,
                         stream, r);
    } else {
      hipLaunchKernelGGL((RingMfmaStartKernel<JT, false, FULL>), grid, dim3(kThreads), lds,
                         stream, r);
    }
  } else {
    if (r.block == 2) {
      hipLaunchKernelGGL((RingMfmaKernel<JT, true, CLASS, FULL>), grid, dim3(kThreads), lds,
                         stream, r);
    } else {
      hipLaunchKernelGGL((RingMfmaKernel<JT, false, CLASS, FULL>), grid, dim3(kThreads), lds,
                         stream, r);
    }
  }
}

template <int JT, int CLASS>
void LaunchFull(const RingLaunch& r, dim3 grid, size_t lds, hipStream_t stream) {
  // every one of the 16 JT rows is a frequency: no padding checks
  if ((r.tables.P - 1) / 2 == 16 * JT) {
    LaunchClass<JT, CLASS, true>(r, grid, lds, stream);
  } else {
    LaunchClass<JT, CLASS, false>(r, grid, lds, stream);
  }
}

template <int JT>
void LaunchJT(const RingLaunch& r, dim3 grid, size_t lds, hipStream_t stream) {
  switch (r.op) {
    case kRingSynthesise:
    case kRingFinish:
      LaunchFull<JT, kClassSynth>(r, grid, lds, stream);
      break;
    case kRingStart:
    case kRingStartFromSpectral:
      LaunchFull<JT, kClassStart>(r, grid, lds, stream);
      break;
    default:
      LaunchFull<JT, kClassOne>(r, grid, lds, stream);
  }
}

}  // namespace

int RingMfmaTiles(int P) {
  if (P % 2 == 0) return 0;
  const int H = (P - 1) / 2;
  const int jt = std::max(1, CeilDiv(H, 16));
  return jt <= kRingMfmaMaxTiles ? jt : 0;
}

size_t RingMfmaLds(const RingLaunch& r) {
  const int slots = kWaves * kSeq / r.tables.R;
  // the operand tables: 2 x (JT x KS x 64) doubles
  const size_t tables = 2 * sizeof(double) * (4 * r.tables.jt * r.tables.jt) * kWave;
  // the staged rows (input coefficients, then outputs): slots x padded orders
  const int stride = 2 * (r.lmax + RowPad(r.lmax, r.tables.R, r.tables.P)) + 1;
  const size_t rows = 16 * static_cast<size_t>(slots) * stride;
  return sizeof(double) * 2 * r.nphi + tables + rows;
}

bool RingPassFusable(const RingLaunch& r) {
  return r.tables.R == 4 && r.tables.jt > 0 && r.lmax >= 3 && RingMfmaLds(r) <= 160 * 1024;
}

void LaunchRingPassMfma(const RingLaunch& r, const RingBlocks& blocks, hipStream_t stream) {
  if (!RingPassFusable(r)) throw ArgumentError("ring pass: no fused launch for this grid");
  const int superTiles = CeilDiv(r.ntheta, 16) * r.nodes;
  const dim3 grid(CeilDiv(superTiles, 8) * 8 * kPassGroups);
  // the spin-0 workgroups' staging tile is the larger footprint
  const size_t lds = RingMfmaLds(r);
  const bool full = (r.tables.P - 1) / 2 == 16 * r.tables.jt;
  const int H = (r.tables.P - 1) / 2;
  const int kIn = (std::min(H, (r.lmax + 3) / 4) + 3) / 4;
  ProfileScope prof("ring_pass", stream);
#define SL3_RING_PASS(JT, FULL, KIN)                                                    \
  hipLaunchKernelGGL((RingMfmaPassKernel<JT, FULL, KIN>), grid, dim3(kThreads), lds,  \
                     stream, r, blocks)
  // the workload's grid (lmax 128: P = 97, three full tiles, kmax 32) with
  // the complex blocks' contraction fixed; every other grid with a run-time
  // count
  if (r.tables.jt == 3 && full && kIn == 8) {
    SL3_RING_PASS(3, true, 8);
  } else if (r.tables.jt == 5 && !full && kIn == 12) {
    // lmax 192 (P = 145, H = 72, kmax 48): the complex blocks hold 12
    // operand blocks instead of 20
    SL3_RING_PASS(5, false, 12);
  } else {
    switch (r.tables.jt) {
      case 1: if (full) SL3_RING_PASS(1, true, 0); else SL3_RING_PASS(1, false, 0); break;
      case 2: if (full) SL3_RING_PASS(2, true, 0); else SL3_RING_PASS(2, false, 0); break;
      case 3: if (full) SL3_RING_PASS(3, true, 0); else SL3_RING_PASS(3, false, 0); break;
      case 4: if (full) SL3_RING_PASS(4, true, 0); else SL3_RING_PASS(4, false, 0); break;
      case 5: if (full) SL3_RING_PASS(5, true, 0); else SL3_RING_PASS(5, false, 0); break;
      default: throw ArgumentError("ring pass: no matrix-core tiling for this grid");
    }
  }
#undef SL3_RING_PASS
  SL3_HIP(hipGetLastError());
}

void LaunchRingMfma(const RingLaunch& r, hipStream_t stream) {
  const int R = r.tables.R;
  if (R != 2 && R != 4) throw ArgumentError("ring transform: R must be 2 or 4");
  const bool real = r.block == 2;
  const int rpg = kWaves * (real ? 2 * kSeq : kSeq) / R;
  const dim3 grid(CeilDiv(r.ntheta, rpg), r.nodes);
  const size_t lds = RingMfmaLds(r);
  ProfileScope prof("ring", stream);
  switch (r.tables.jt) {
    case 1: LaunchJT<1>(r, grid, lds, stream); break;
    case 2: LaunchJT<2>(r, grid, lds, stream); break;
    case 3: LaunchJT<3>(r, grid, lds, stream); break;
    case 4: LaunchJT<4>(r, grid, lds, stream); break;
    case 5: LaunchJT<5>(r, grid, lds, stream); break;
    default: throw ArgumentError("ring transform: no matrix-core tiling for this grid");
  }
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
