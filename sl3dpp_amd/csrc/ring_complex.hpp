// The pass's longitude stage (Fourier coefficients -> rings x weight ->
// Fourier coefficients) for a complex block, with whole complex values per
// lane.
//
// The transform is the matrix-core form of ring_mfma.hip: nphi = 4 P, P odd,
// sequences z_m2[m1] = F_{4 m1 + m2}, P-point transforms in the symmetric
// cos / sin form on v_mfma_f64_16x16x4f64 with the band limit |m| <= L
// cutting the contraction, twiddles and 4-point transforms across the four
// sequences of a ring. Here a wave holds 16 complex sequences (4 rings x
// the 4 residues m2) as the 16 columns of two tiles, one of real parts and
// one of imaginary parts, so that
//  * each table operand read from LDS feeds two products;
//  * a lane holds a point's real and imaginary part: the twiddle and the
//    x i rotations are in-lane arithmetic, and the two radix-2 stages of the
//    4-point transform exchange with lane ^ 2 and ^ 1 (DPP quad moves);
//  * a lane's operands are the coefficients of its own ring and residue,
//    read straight from the Fourier layout F[m + L][node][theta] (four
//    consecutive rings of an order are 64 contiguous bytes), and its outputs
//    are stored the same way: no staging tile, so LDS holds only the
//    twiddles and the operand tables.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "transform.hpp"

namespace sl3amd {

// Device code only, for the one translation unit that launches it
// (ring_mfma.hip's fused pass).
namespace complex_lanes {

constexpr int kWave = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kWave * kWaves;
constexpr int kRingsPerWave = 4;  // 16 sequences: 4 rings x 4 residues
constexpr int kRingsPerGroup = kWaves * kRingsPerWave;

typedef double Acc4 __attribute__((ext_vector_type(4)));

template <int CTRL>
__device__ inline double QuadPerm(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double Xor1(double v) { return QuadPerm<0xB1>(v); }  // [1,0,3,2]
__device__ inline double Xor2(double v) { return QuadPerm<0x4E>(v); }  // [2,3,0,1]

// The sum over the four lanes of a column (lane, lane ^ 16, ^ 32, ^ 48).
__device__ inline double ColumnSum(double v) {
  v += __shfl_xor(v, 16, kWave);
  v += __shfl_xor(v, 32, kWave);
  return v;
}

struct Lane {
  int lane, quad, m2, ring;
  bool valid;
  double sA, sB;  // +1 for the sum half of the residue's stages, -1 for the difference
  bool rot;       // the (1, 1) lanes: x i between the stages
};

// Synthesis at n1: v = e^{2 pi i m2 n1 / nphi} y, then the 4-point inverse
// transform across the ring's lanes (m2 = 2 a + b: a over lane ^ 2, the
// x i^{bc} twiddle, b over lane ^ 1); the lane ends with x[n1 + P (a + 2 b)].
__device__ inline void InverseButterfly(const Lane& ln, double& yr, double& yi, double c,
                                        double s) {
  const double vr = c * yr - s * yi;
  const double vi = fma(c, yi, s * yr);
  double er = fma(ln.sA, vr, Xor2(vr));
  double ei = fma(ln.sA, vi, Xor2(vi));
  if (ln.rot) {
    const double t = er;
    er = -ei;
    ei = t;
  }
  yr = fma(ln.sB, er, Xor1(er));
  yi = fma(ln.sB, ei, Xor1(ei));
}

// Analysis at n1: the 4-point forward transform across the ring's lanes,
// then the conjugate twiddle; the lane ends with U_m2[n1].
__device__ inline void ForwardButterfly(const Lane& ln, double& yr, double& yi, double c,
                                        double s) {
  double dr = fma(ln.sB, yr, Xor1(yr));
  double di = fma(ln.sB, yi, Xor1(yi));
  if (ln.rot) {  // x (-i)
    const double t = dr;
    dr = di;
    di = -t;
  }
  const double wr = fma(ln.sA, dr, Xor2(dr));
  const double wi = fma(ln.sA, di, Xor2(di));
  yr = fma(c, wr, s * wi);
  yi = c * wi - s * wr;
}

// A barrier to the scheduler only: the phases below keep their own loads,
// so that one phase's prefetches do not hold registers through another.
__device__ inline void PhaseFence() { __builtin_amdgcn_sched_barrier(0); }

// A value the compiler may not carry from one phase into another (it
// recomputes what depends on it instead of holding it in registers).
__device__ inline int Opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// A block's Fourier coefficients at byte offset b from a node's base (the
// base uniform, the offset 32-bit: one register per address).
__device__ inline double2 LoadAt(const double2* base, unsigned b) {
  return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(base) + b);
}
__device__ inline void StoreAt(double2* base, unsigned b, double2 v) {
  *reinterpret_cast<double2*>(reinterpret_cast<char*>(base) + b) = v;
}

// One workgroup: rings th0 .. th0 + 15 of one node of one complex block.
// KIN: the inverse contraction's 4-deep blocks (the band limit's
// ceil(kmax / 4)) when fixed at compile time, 0 for a run-time count.
template <int JT, bool FULL, int KIN>
__device__ inline void ComplexPassBody(const RingLaunch& a, const double2* __restrict__ in,
                                       double2* __restrict__ out, int node, int th0) {
  constexpr int KS = 4 * JT;
  constexpr int KA = KIN > 0 ? KIN : KS;  // operand blocks held
  extern __shared__ double lds[];
  const int P = a.tables.P, H = (P - 1) / 2, L = a.lmax, nphi = a.nphi;
  const int kmax = min(H, (L + 3) / 4);
  const int kIn = KIN > 0 ? KIN : (kmax + 3) / 4;
  const int tOut = KIN > 0 ? (KIN + 3) / 4 : (kmax + 15) / 16;
  Lane ln;
  ln.lane = threadIdx.x % kWave;
  ln.quad = ln.lane / 16;
  const int s = ln.lane % 16;
  ln.m2 = s & 3;
  const int ra = ln.m2 >> 1, rb = ln.m2 & 1;
  ln.ring = (threadIdx.x / kWave) * kRingsPerWave + (s >> 2);
  const int th = th0 + ln.ring;
  ln.valid = th < a.ntheta;
  ln.sA = ra == 0 ? 1.0 : -1.0;
  ln.sB = rb == 0 ? 1.0 : -1.0;
  ln.rot = ra == 1 && rb == 1;
  const int n2 = ra + 2 * rb;

  // LDS: the twiddles cos / sin(2 pi n / nphi), then the operand tables
  constexpr int kTab = JT * KS * kWave;
  double* cosN = lds;
  double* sinN = lds + nphi;
  double* tabC = lds + 2 * nphi;
  double* tabS = tabC + kTab;
  for (int i0 = 0; i0 < nphi; i0 += 2 * kThreads) {
    double c[2], sn[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + threadIdx.x + u * kThreads;
      c[u] = i < nphi ? a.tables.cosN[i] : 0.0;
      sn[u] = i < nphi ? a.tables.sinN[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + threadIdx.x + u * kThreads;
      if (i < nphi) {
        cosN[i] = c[u];
        sinN[i] = sn[u];
      }
    }
  }
  {
    constexpr int kPerThread = kTab / kThreads;
    double tc[kPerThread], ts[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      tc[u] = a.tables.tabC[threadIdx.x + u * kThreads];
      ts[u] = a.tables.tabS[threadIdx.x + u * kThreads];
    }
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      tabC[threadIdx.x + u * kThreads] = tc[u];
      tabS[threadIdx.x + u * kThreads] = ts[u];
    }
  }

  // The lane's sequence: order m2 + 4 k at element k, m2 - 4 k at P - k,
  // zero beyond +-L (the band limit leaves k <= kmax). Loads are
  // unconditional where the order exists (a ring past the last reads the
  // group's first ring), and issue together.
  // byte offsets within the block: (m + L) nodes ntheta + theta, below 4 GiB
  const unsigned mStride = 16u * static_cast<unsigned>(a.nodes) * a.fstride;
  const double2* F = in + static_cast<size_t>(node) * a.fstride;
  const unsigned thIn = 16u * static_cast<unsigned>(ln.valid ? th : th0);
  auto at = [&](int m, unsigned thb) { return static_cast<unsigned>(m + L) * mStride + thb; };
  const double2 h0 = LoadAt(F, at(ln.m2, thIn));
  double avr[KA], avi[KA], bvr[KA], bvi[KA];
  double sumr = 0.0, sumi = 0.0;
#pragma unroll
  for (int kb = 0; kb < KA; ++kb) {
    const int k = 1 + 4 * kb + ln.quad;
    const int mu = ln.m2 + 4 * k, md = ln.m2 - 4 * k;
    double2 up = make_double2(0.0, 0.0), dn = make_double2(0.0, 0.0);
    if (KIN > 0 || kb < kIn) {
      if (mu <= L) up = LoadAt(F, at(mu, thIn));
      if (md >= -L) dn = LoadAt(F, at(md, thIn));
    }
    avr[kb] = up.x + dn.x;
    avi[kb] = up.y + dn.y;
    bvr[kb] = up.x - dn.x;
    bvi[kb] = up.y - dn.y;
    sumr += avr[kb];
    sumi += avi[kb];
  }
  // the weights at the lane's points (j past H clamped), in flight through
  // the inverse transforms
  auto jOf = [&](int v) { return 1 + 16 * (v / 4) + 4 * (v % 4) + ln.quad; };
  auto live = [&](int v) { return FULL || jOf(v) <= H; };
  auto jc = [&](int v) { return FULL ? jOf(v) : min(jOf(v), H); };
  const double* w = a.weight + (static_cast<size_t>(node) * a.ntheta + (ln.valid ? th : th0)) *
                                   nphi + static_cast<size_t>(P) * n2;
  double wj[KS], wq[KS];
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    const int j = jc(v);
    wj[v] = w[j];
    wq[v] = w[P - j];
  }
  const double w0 = w[0];
  __syncthreads();  // the tables are staged

  // Y_j = h0 + C_j + i S_j, Y_{P-j} = h0 + C_j - i S_j: C = cos (h_k + h_{P-k}),
  // S = sin (h_k - h_{P-k}), real and imaginary tiles apart
  double yjr[KS], yji[KS], yqr[KS], yqi[KS];
  {
    // every tile's four chains, the contraction outermost: the operands die
    // as they are used
    Acc4 Cr[JT], Ci[JT], Sr[JT], Si[JT];
#pragma unroll
    for (int t = 0; t < JT; ++t) Cr[t] = Ci[t] = Sr[t] = Si[t] = Acc4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kb = 0; kb < KA; ++kb) {
      if (KIN == 0 && kb >= kIn) break;
#pragma unroll
      for (int t = 0; t < JT; ++t) {
        const int i = (t * KS + kb) * kWave + ln.lane;
        const double c = tabC[i], sn = tabS[i];
        Cr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(c, avr[kb], Cr[t], 0, 0, 0);
        Ci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(c, avi[kb], Ci[t], 0, 0, 0);
        Sr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, bvr[kb], Sr[t], 0, 0, 0);
        Si[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, bvi[kb], Si[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < JT; ++t) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int v = 4 * t + q;
        const double cr = h0.x + Cr[t][q], ci = h0.y + Ci[t][q];
        yjr[v] = cr - Si[t][q];
        yji[v] = ci + Sr[t][q];
        yqr[v] = cr + Si[t][q];
        yqi[v] = ci - Sr[t][q];
      }
    }
  }
  double y0r = h0.x + ColumnSum(sumr), y0i = h0.y + ColumnSum(sumi);
  PhaseFence();

  // The lane's points n1 = j = 1 + 16 t + 4 q + quad, P - j and 0, each at
  // n1 + P n2 on its ring; j past H (padding) is clamped for the twiddle and
  // the weight read, and its value dropped.
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    const int j = jc(v);
    InverseButterfly(ln, yjr[v], yji[v], cosN[ln.m2 * j], sinN[ln.m2 * j]);
    InverseButterfly(ln, yqr[v], yqi[v], cosN[ln.m2 * (P - j)], sinN[ln.m2 * (P - j)]);
    if (v % 2 == 1) PhaseFence();  // two points' twiddles at a time
  }
  InverseButterfly(ln, y0r, y0i, 1.0, 0.0);
  PhaseFence();
  // y = W x at the points (zero on a ring past the last and on padding)
  {
    const double w0v = ln.valid ? w0 : 0.0;
    y0r *= w0v;
    y0i *= w0v;
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      const double a0 = ln.valid && live(v) ? wj[v] : 0.0;
      const double a1 = ln.valid && live(v) ? wq[v] : 0.0;
      yjr[v] *= a0;
      yji[v] *= a0;
      yqr[v] *= a1;
      yqi[v] *= a1;
    }
  }
  PhaseFence();
  // forward butterflies, then the P-point forward transforms; the twiddles
  // are read again (an opaque offset), not held from the inverse
  const double* cosF = cosN + Opaque(0);
  const double* sinF = sinN + Opaque(0);
  const int m2f = Opaque(ln.m2);
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    const int j = jc(v);
    ForwardButterfly(ln, yjr[v], yji[v], cosF[m2f * j], sinF[m2f * j]);
    ForwardButterfly(ln, yqr[v], yqi[v], cosF[m2f * (P - j)], sinF[m2f * (P - j)]);
    if (v % 2 == 1) PhaseFence();
  }
  ForwardButterfly(ln, y0r, y0i, 1.0, 0.0);
  PhaseFence();
  sumr = sumi = 0.0;
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    const double ar = yjr[v] + yqr[v], ai = yji[v] + yqi[v];
    const double br = yjr[v] - yqr[v], bi = yji[v] - yqi[v];
    yjr[v] = ar;
    yji[v] = ai;
    yqr[v] = br;
    yqi[v] = bi;
    sumr += ar;
    sumi += ai;
  }
  const double g0r = y0r + ColumnSum(sumr), g0i = y0i + ColumnSum(sumi);
  // G(k) = u0 + C_k - i S_k at order m2 + 4 k, G(P - k) = u0 + C_k + i S_k at
  // m2 - 4 k, scaled by (2 pi / nphi) w_theta, stored where |order| <= L
  const double scale = ln.valid ? 2.0 * 3.14159265358979323846 / nphi * a.wtheta[th] : 0.0;
  double2* G = out + static_cast<size_t>(node) * a.fstride;
  const int m2 = Opaque(ln.m2), quad = Opaque(ln.quad);
  const unsigned thOut = 16u * static_cast<unsigned>(Opaque(th));
  if (ln.valid && quad == 0) StoreAt(G, at(m2, thOut), make_double2(scale * g0r, scale * g0i));
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    if (t >= tOut) break;
    Acc4 Cr = {0.0, 0.0, 0.0, 0.0}, Ci = Cr, Sr = Cr, Si = Cr;
#pragma unroll
    for (int kb = 0; kb < KS; ++kb) {
      const int i = (t * KS + kb) * kWave + ln.lane;
      const double c = tabC[i], sn = tabS[i];
      Cr = __builtin_amdgcn_mfma_f64_16x16x4f64(c, yjr[kb], Cr, 0, 0, 0);
      Ci = __builtin_amdgcn_mfma_f64_16x16x4f64(c, yji[kb], Ci, 0, 0, 0);
      Sr = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, yqr[kb], Sr, 0, 0, 0);
      Si = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, yqi[kb], Si, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = 1 + 16 * t + 4 * q + quad;
      const double cr = y0r + Cr[q], ci = y0i + Ci[q];
      const int mu = m2 + 4 * k, md = m2 - 4 * k;
      if (ln.valid && mu <= L) {
        StoreAt(G, at(mu, thOut), make_double2(scale * (cr + Si[q]), scale * (ci - Sr[q])));
      }
      if (ln.valid && md >= -L) {
        StoreAt(G, at(md, thOut), make_double2(scale * (cr - Si[q]), scale * (ci + Sr[q])));
      }
    }
    PhaseFence();
  }
}

// The forward half of the pass from the lane's point values (y at n1 = j,
// P - j and 0, each at n1 + P n2 of its ring): the forward butterflies, the
// P-point forward transforms, the stores to the Fourier layout at |order| <=
// L, scaled by (2 pi / nphi) w_theta. The same statements as the pass's.
template <int JT>
__device__ inline void ComplexForwardStore(const RingLaunch& a, const Lane& ln, int th, int node,
                                           const double* cosN, const double* sinN,
                                           const double* tabC, const double* tabS,
                                           double (&yjr)[4 * JT], double (&yji)[4 * JT],
                                           double (&yqr)[4 * JT], double (&yqi)[4 * JT],
                                           double y0r, double y0i, double2* out, bool full) {
  constexpr int KS = 4 * JT;
  const int P = a.tables.P, H = (P - 1) / 2, L = a.lmax, nphi = a.nphi;
  const int kmax = min(H, (L + 3) / 4);
  const int tOut = (kmax + 15) / 16;
  auto jOf = [&](int v) { return 1 + 16 * (v / 4) + 4 * (v % 4) + ln.quad; };
  auto jc = [&](int v) { return full ? jOf(v) : min(jOf(v), H); };
  const double* cosF = cosN + Opaque(0);
  const double* sinF = sinN + Opaque(0);
  const int m2f = Opaque(ln.m2);
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    const int j = jc(v);
    ForwardButterfly(ln, yjr[v], yji[v], cosF[m2f * j], sinF[m2f * j]);
    ForwardButterfly(ln, yqr[v], yqi[v], cosF[m2f * (P - j)], sinF[m2f * (P - j)]);
    if (v % 2 == 1) PhaseFence();
  }
  ForwardButterfly(ln, y0r, y0i, 1.0, 0.0);
  PhaseFence();
  double sumr = 0.0, sumi = 0.0;
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    const double ar = yjr[v] + yqr[v], ai = yji[v] + yqi[v];
    const double br = yjr[v] - yqr[v], bi = yji[v] - yqi[v];
    yjr[v] = ar;
    yji[v] = ai;
    yqr[v] = br;
    yqi[v] = bi;
    sumr += ar;
    sumi += ai;
  }
  const double g0r = y0r + ColumnSum(sumr), g0i = y0i + ColumnSum(sumi);
  const double scale = ln.valid ? 2.0 * 3.14159265358979323846 / nphi * a.wtheta[th] : 0.0;
  double2* G = out + static_cast<size_t>(node) * a.fstride;
  const unsigned mStride = 16u * static_cast<unsigned>(a.nodes) * a.fstride;
  auto at = [&](int m, unsigned thb) { return static_cast<unsigned>(m + L) * mStride + thb; };
  const int m2 = Opaque(ln.m2), quad = Opaque(ln.quad);
  const unsigned thOut = 16u * static_cast<unsigned>(Opaque(th));
  if (ln.valid && quad == 0) StoreAt(G, at(m2, thOut), make_double2(scale * g0r, scale * g0i));
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    if (t >= tOut) break;
    Acc4 Cr = {0.0, 0.0, 0.0, 0.0}, Ci = Cr, Sr = Cr, Si = Cr;
#pragma unroll
    for (int kb = 0; kb < KS; ++kb) {
      const int i = (t * KS + kb) * kWave + ln.lane;
      const double c = tabC[i], sn = tabS[i];
      Cr = __builtin_amdgcn_mfma_f64_16x16x4f64(c, yjr[kb], Cr, 0, 0, 0);
      Ci = __builtin_amdgcn_mfma_f64_16x16x4f64(c, yji[kb], Ci, 0, 0, 0);
      Sr = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, yqr[kb], Sr, 0, 0, 0);
      Si = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, yqi[kb], Si, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = 1 + 16 * t + 4 * q + quad;
      const double cr = y0r + Cr[q], ci = y0i + Ci[q];
      const int mu = m2 + 4 * k, md = m2 - 4 * k;
      if (ln.valid && mu <= L) {
        StoreAt(G, at(mu, thOut), make_double2(scale * (cr + Si[q]), scale * (ci - Sr[q])));
      }
      if (ln.valid && md >= -L) {
        StoreAt(G, at(md, thOut), make_double2(scale * (cr - Si[q]), scale * (ci + Sr[q])));
      }
    }
    PhaseFence();
  }
}

// The step start's analyses of a complex block with whole complex values
// per lane (src/Viscoelastic3D.cpp:1283-1365, one branch), in two launches:
// START: from d0 (gridIn), the memory m (memIn) and itau, the weights of
// h = dt itau, m~ = e m + alpha d0 stored (memOut), and y = W d0 (W =
// mu (beta - beta_ref)) transformed into fourierOut2; otherwise y = mu m~ of
// the stored memOut transformed into fourierOut. One workgroup: rings th0
// .. th0 + 15 of one node; each lane's points are the pass's.
template <int JT, bool FULL, bool START>
__device__ inline void ComplexStartBody(const RingLaunch& a, int node, int th0) {
  constexpr int KS = 4 * JT;
  extern __shared__ double lds[];
  const int P = a.tables.P, H = (P - 1) / 2, nphi = a.nphi;
  Lane ln;
  ln.lane = threadIdx.x % kWave;
  ln.quad = ln.lane / 16;
  const int s = ln.lane % 16;
  ln.m2 = s & 3;
  const int ra = ln.m2 >> 1, rb = ln.m2 & 1;
  ln.ring = (threadIdx.x / kWave) * kRingsPerWave + (s >> 2);
  const int th = th0 + ln.ring;
  ln.valid = th < a.ntheta;
  ln.sA = ra == 0 ? 1.0 : -1.0;
  ln.sB = rb == 0 ? 1.0 : -1.0;
  ln.rot = ra == 1 && rb == 1;
  const int n2 = ra + 2 * rb;

  constexpr int kTab = JT * KS * kWave;
  double* cosN = lds;
  double* sinN = lds + nphi;
  double* tabC = lds + 2 * nphi;
  double* tabS = tabC + kTab;
  for (int i0 = 0; i0 < nphi; i0 += 2 * kThreads) {
    double c[2], sn[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + threadIdx.x + u * kThreads;
      c[u] = i < nphi ? a.tables.cosN[i] : 0.0;
      sn[u] = i < nphi ? a.tables.sinN[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + threadIdx.x + u * kThreads;
      if (i < nphi) {
        cosN[i] = c[u];
        sinN[i] = sn[u];
      }
    }
  }
  {
    constexpr int kPerThread = kTab / kThreads;
    double tc[kPerThread], ts[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      tc[u] = a.tables.tabC[threadIdx.x + u * kThreads];
      ts[u] = a.tables.tabS[threadIdx.x + u * kThreads];
    }
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      tabC[threadIdx.x + u * kThreads] = tc[u];
      tabS[threadIdx.x + u * kThreads] = ts[u];
    }
  }

  // the lane's points j (clamped to H past the last frequency), P - j, 0
  auto jOf = [&](int v) { return 1 + 16 * (v / 4) + 4 * (v % 4) + ln.quad; };
  auto live = [&](int v) { return ln.valid && (FULL || jOf(v) <= H); };
  auto jc = [&](int v) { return FULL ? jOf(v) : min(jOf(v), H); };
  const size_t base = (static_cast<size_t>(node) * a.ntheta + (ln.valid ? th : th0)) * nphi +
                      static_cast<size_t>(P) * n2;
  const double mu = a.muNode[node];
  double yjr[KS], yji[KS], yqr[KS], yqi[KS], y0r, y0i;
  if constexpr (START) {
    const double2* d0 = reinterpret_cast<const double2*>(a.gridIn) + base;
    const double2* mIn = reinterpret_cast<const double2*>(a.memIn) + base;
    double2* mOut = reinterpret_cast<double2*>(a.memOut) + base;
    const double* itau = a.itau + base;
    const double href = a.dt * a.itauRefNode[node];
    auto phi1 = [](double h) {
      return fabs(h) < 1e-3 ? 1.0 - h / 2.0 + h * h / 6.0 - h * h * h / 24.0 +
                                  h * h * h * h / 120.0
                            : (1.0 - exp(-h)) / h;
    };
    const double refBeta = a.scheme == 1 ? 1.0 - 1.0 / (1.0 + href) : 1.0 - phi1(href);
    // m~ at a point from its loaded values (stored where `store`), and W d0
    auto start = [&](int n1, double2 d, double2 m, double it, bool store, double& wr,
                     double& wi) {
      const double h = a.dt * it;
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
      if (store) mOut[n1] = make_double2(ew * m.x + aw * d.x, ew * m.y + aw * d.y);
      const double W = mu * (bw - refBeta);
      wr = W * d.x;
      wi = W * d.y;
    };
    start(0, d0[0], mIn[0], itau[0], ln.valid && ln.quad == 0, y0r, y0i);
    if (!ln.valid) y0r = y0i = 0.0;
    // a group's loads issue together: each store of m~ (to the buffer the
    // memory is read from) orders the loads after it behind it
    constexpr int kGroup = KS % 6 == 0 ? 6 : 4;
#pragma unroll
    for (int g0 = 0; g0 < KS; g0 += kGroup) {
      double2 dj[kGroup], dq[kGroup], mj[kGroup], mq[kGroup];
      double hj[kGroup], hq[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const int j = jc(g0 + u);
        dj[u] = d0[j];
        dq[u] = d0[P - j];
        mj[u] = mIn[j];
        mq[u] = mIn[P - j];
        hj[u] = itau[j];
        hq[u] = itau[P - j];
      }
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const int v = g0 + u, j = jc(v);
        start(j, dj[u], mj[u], hj[u], live(v), yjr[v], yji[v]);
        start(P - j, dq[u], mq[u], hq[u], live(v), yqr[v], yqi[v]);
        if (!live(v)) yjr[v] = yji[v] = yqr[v] = yqi[v] = 0.0;
      }
    }
  } else {
    const double2* m = reinterpret_cast<const double2*>(a.memOut) + base;
    const double2 v0 = m[0];
    y0r = ln.valid ? mu * v0.x : 0.0;
    y0i = ln.valid ? mu * v0.y : 0.0;
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      const int j = jc(v);
      const double2 xj = m[j], xq = m[P - j];
      yjr[v] = live(v) ? mu * xj.x : 0.0;
      yji[v] = live(v) ? mu * xj.y : 0.0;
      yqr[v] = live(v) ? mu * xq.x : 0.0;
      yqi[v] = live(v) ? mu * xq.y : 0.0;
    }
  }
  __syncthreads();  // the tables are staged
  PhaseFence();
  ComplexForwardStore<JT>(a, ln, th, node, cosN, sinN, tabC, tabS, yjr, yji, yqr, yqi, y0r, y0i,
                          reinterpret_cast<double2*>(START ? a.fourierOut2 : a.fourierOut),
                          FULL);
}

}  // namespace complex_lanes
}  // namespace sl3amd
