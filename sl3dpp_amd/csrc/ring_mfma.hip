// The longitude stage of the grid transforms on the FP64 matrix cores
// (gfx950, v_mfma_f64_16x16x4f64), for grids with nphi = R P, P odd.
//
// A ring of nphi points splits into R interleaved sequences x_r[j] =
// x[R j + r] of P points (Cooley-Tukey with R outer); each P-point DFT is
// computed in the symmetric form
//   x_j = h_0 + sum_k (h_k + h_{P-k}) cos(2 pi jk/P) + i (h_k - h_{P-k}) sin(..)
// and its forward twin, as products of the constant (H x H) cos and sin
// matrices (H = (P-1)/2, padded to 16) with the sequences. One wave holds
// 8 complex sequences as the 16 real columns of a matrix-core tile; the
// matrices are the A operands, the data the B operands, so the inverse
// transform's accumulators are already the forward transform's operands
// and the pointwise product between them never leaves the registers.
// The spin-0 block packs two real rings into one complex sequence.
//
// Per workgroup (4 waves): one memory node, a run of consecutive rings.
//  1. stage the rings' Fourier coefficients in LDS (synthesis ops);
//  2. fold: h_r[k] = sum_{m = k mod P} F_m e^{2 pi i r m / nphi};
//  3. inverse sub-DFTs on the matrix cores, x at j and P - j;
//  4. the pointwise operation of the op (weight, step start, finish ...);
//  5. forward sub-DFTs on the matrix cores, G_r[k] staged in LDS;
//  6. combine: F_m = (2 pi / nphi) w_theta sum_r e^{-2 pi i r m / nphi}
//     G_r[m mod P], written with consecutive rings in consecutive lanes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "transform.hpp"

namespace sl3amd {

namespace {

constexpr int kWave = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kWave * kWaves;
constexpr int kSeq = 8;  // complex sequences per wave

typedef double Acc4 __attribute__((ext_vector_type(4)));

__device__ inline double Shfl(double v, int mask) { return __shfl_xor(v, mask, kWave); }

// The sum over the four lanes of a column (lane, lane ^ 16, ^ 32, ^ 48),
// the same value in each of them.
__device__ inline double ColumnSum(double v) {
  v += Shfl(v, 16);
  v += Shfl(v, 32);
  return v;
}

struct Geometry {
  int R, P, H, nphi, L, mlo, orders;
  int ringsPerWave, ringsPerGroup;
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
  return g;
}

// z times i^n (n = 0..3).
__device__ inline double2 IPow(double2 z, int n) {
  double a = (n & 1) ? -z.y : z.x;
  double b = (n & 1) ? z.x : z.y;
  if (n & 2) {
    a = -a;
    b = -b;
  }
  return make_double2(a, b);
}

// The fold of the lane's sequence at frequency k (complex, both parts):
// with m = k + q P and nphi = R P, e^{2 pi i r m / nphi} =
// e^{2 pi i r k / nphi} (e^{2 pi i / R})^{r q}, the second factor a power
// of i, so one complex product per fold. Fs: the group's staged
// coefficients [ring][m - mlo]; cosN / sinN: cos / sin(2 pi n / nphi).
template <bool REAL>
__device__ inline double2 Fold(const Geometry& g, const double2* Fs, const double* cosN,
                               const double* sinN, int ringA, int r, int k) {
  double2 acc = make_double2(0.0, 0.0);
  int m = k, q = 0;
  while (m - g.P >= -g.L) {
    m -= g.P;
    --q;
  }
  const int step4 = 4 / g.R;
  for (; m <= g.L; m += g.P, ++q) {
    double2 z;
    if (REAL) {
      // Z_m = F1_m + i F2_m, F_{-m} = conj(F_m), the real part at m = 0
      const int am = m < 0 ? -m : m;
      const double2 f1 = Fs[ringA * g.orders + am];
      const double2 f2 = Fs[(ringA + 1) * g.orders + am];
      const double sg = m < 0 ? -1.0 : 1.0;
      const double i1 = m == 0 ? 0.0 : sg * f1.y, i2 = m == 0 ? 0.0 : sg * f2.y;
      z = make_double2(f1.x - i2, i1 + f2.x);
    } else {
      z = Fs[ringA * g.orders + (m - g.mlo)];
    }
    const double2 t = IPow(z, (r * q * step4) & 3);
    acc.x += t.x;
    acc.y += t.y;
  }
  const double c = cosN[r * k], sn = sinN[r * k];
  return make_double2(acc.x * c - acc.y * sn, acc.x * sn + acc.y * c);
}

// Per-lane indexing: the lane's column is sequence s = col / 2, part p.
struct Lane {
  int lane, col, quad, s, p, r;
  int seqGroup;  // the lane's sequence within the group
  int ring;      // REAL: the lane's own real ring; else the complex ring
  int ringA;     // the ring the fold reads first (REAL: the pair's first)
  bool valid;    // the lane's ring exists
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
  l.r = l.s % g.R;
  l.seqGroup = wave * kSeq + l.s;
  if (REAL) {
    l.ringA = wave * g.ringsPerWave + 2 * (l.s / g.R);
    l.ring = l.ringA + l.p;
  } else {
    l.ringA = wave * g.ringsPerWave + l.s / g.R;
    l.ring = l.ringA;
  }
  l.valid = th0 + l.ring < ntheta;
  return l;
}

// The forward sub-DFTs of y (the lane's part at j, at P - j, at 0) on the
// matrix cores, staged as G[seq][k] in LDS. SYNC: the stage overlays the
// operand tables, so every wave first finishes its products.
template <int JT, bool SYNC>
__device__ inline void ForwardToStage(const Geometry& g, const Lane& ln,
                                      const double* tabC, const double* tabS,
                                      const double (&yj)[4 * JT], const double (&yq)[4 * JT],
                                      double y0, double* Gd) {
  constexpr int KS = 4 * JT;
  double av[KS], bv[KS];
  double sum = 0.0;
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    av[v] = yj[v] + yq[v];
    bv[v] = yj[v] - yq[v];
    sum += av[v];
  }
  const double g0 = y0 + ColumnSum(sum);
  Acc4 C[JT], S[JT];
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    C[t] = Acc4{0.0, 0.0, 0.0, 0.0};
    S[t] = Acc4{0.0, 0.0, 0.0, 0.0};
  }
  // 2 JT independent accumulation chains, interleaved
#pragma unroll
  for (int kb = 0; kb < KS; ++kb) {
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int at = (t * KS + kb) * kWave + ln.lane;
      C[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabC[at], av[kb], C[t], 0, 0, 0);
      S[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabS[at], bv[kb], S[t], 0, 0, 0);
    }
  }
  if constexpr (SYNC) __syncthreads();
  // G(k) = y0 + C'_k - i S'_k, G(P - k) = y0 + C'_k + i S'_k
  const double sg = ln.p == 0 ? 1.0 : -1.0;
  double* row = Gd + static_cast<size_t>(ln.seqGroup) * g.P * 2 + ln.p;
  if (ln.quad == 0) row[0] = g0;
#pragma unroll
  for (int t = 0; t < JT; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double sx = Shfl(S[t][q], 1);
      const int k = 1 + 16 * t + 4 * q + ln.quad;
      if (k <= g.H) {
        row[2 * k] = y0 + C[t][q] + sg * sx;
        row[2 * (g.P - k)] = y0 + C[t][q] - sg * sx;
      }
    }
  }
}

// Combine the staged G into F_m for every ring of the group: per
// frequency k < P, u_r = G_r(k) e^{-2 pi i r k / nphi}, then every order
// m = k + q P is the R-point sum of u_r i^{-(4/R) r q}; consecutive
// threads take consecutive rings.
template <bool REAL>
__device__ inline void CombineFromStage(const Geometry& g, const RingLaunch& a,
                                        const double* cosN, const double* sinN,
                                        const double2* Gs, int node, int th0,
                                        double* out) {
  const double scale = 2.0 * 3.14159265358979323846 / g.nphi;
  const int tasks = g.P * g.ringsPerGroup;
  const int step4 = 4 / g.R;
  for (int e = threadIdx.x; e < tasks; e += blockDim.x) {
    const int ring = e % g.ringsPerGroup;
    const int k = e / g.ringsPerGroup;
    const int th = th0 + ring;
    if (th >= a.ntheta) continue;
    const int wave = ring / g.ringsPerWave, gl = ring % g.ringsPerWave;
    const double w = scale * a.wtheta[th];
    auto at = [&](int mo) {
      return out + ((static_cast<size_t>(mo) * a.nodes + node) * a.ntheta + th) * 2;
    };
    if (REAL) {
      const int base = wave * kSeq + (gl / 2) * g.R;
      const int kk = (g.P - k) % g.P;
      double2 u[4], v[4];
      for (int r = 0; r < g.R; ++r) {
        const double2 gp = Gs[(base + r) * g.P + k];
        const double2 gm = Gs[(base + r) * g.P + kk];
        const double c = cosN[r * k], sn = sinN[r * k];
        const double c2 = cosN[r * kk], s2 = sinN[r * kk];
        u[r] = make_double2(gp.x * c + gp.y * sn, gp.y * c - gp.x * sn);
        v[r] = make_double2(gm.x * c2 + gm.y * s2, gm.y * c2 - gm.x * s2);
      }
      // m = k + q P >= 0; -m = kk + q' P with q' = -q (k = 0) or -(q + 1)
      for (int m = k, q = 0; m <= g.L; m += g.P, ++q) {
        const int qm = k == 0 ? -q : -(q + 1);
        double2 zp = make_double2(0.0, 0.0), zm = make_double2(0.0, 0.0);
        for (int r = 0; r < g.R; ++r) {
          const double2 tp = IPow(u[r], (-r * q * step4) & 3);
          const double2 tm = IPow(v[r], (-r * qm * step4) & 3);
          zp.x += tp.x;
          zp.y += tp.y;
          zm.x += tm.x;
          zm.y += tm.y;
        }
        double2 res;
        if (gl % 2 == 0) {
          res = make_double2(0.5 * (zp.x + zm.x), 0.5 * (zp.y - zm.y));
        } else {
          // (Z_m - conj(Z_{-m})) / 2i
          res = make_double2(0.5 * (zp.y + zm.y), -0.5 * (zp.x - zm.x));
        }
        double* dst = at(m);
        dst[0] = res.x * w;
        dst[1] = res.y * w;
      }
    } else {
      const int base = wave * kSeq + gl * g.R;
      double2 u[4];
      for (int r = 0; r < g.R; ++r) {
        const double2 gv = Gs[(base + r) * g.P + k];
        const double c = cosN[r * k], sn = sinN[r * k];
        u[r] = make_double2(gv.x * c + gv.y * sn, gv.y * c - gv.x * sn);
      }
      int m = k, q = 0;
      while (m - g.P >= -g.L) {
        m -= g.P;
        --q;
      }
      for (; m <= g.L; m += g.P, ++q) {
        double2 res = make_double2(0.0, 0.0);
        for (int r = 0; r < g.R; ++r) {
          const double2 t = IPow(u[r], (-r * q * step4) & 3);
          res.x += t.x;
          res.y += t.y;
        }
        double* dst = at(m - g.mlo);
        dst[0] = res.x * w;
        dst[1] = res.y * w;
      }
    }
  }
}

// The op classes, one kernel each so that each carries only its own live
// registers: synthesis only (synthesise, finish), one analysis (analyse,
// weighted, node scaled, pass), the step start (two analyses).
enum { kClassSynth = 0, kClassOne = 1, kClassStart = 2 };

template <int JT, bool REAL, int CLASS>
__global__ void __launch_bounds__(kThreads) RingMfmaKernel(RingLaunch a) {
  constexpr int KS = 4 * JT;
  extern __shared__ double lds[];
  const Geometry g = MakeGeometry<REAL>(a);
  const int node = blockIdx.y;
  const int th0 = blockIdx.x * g.ringsPerGroup;
  const Lane ln = MakeLane<REAL>(g, th0, a.ntheta);
  const double* cosN = a.tables.cosN;  // small, cache-resident
  const double* sinN = a.tables.sinN;
  // LDS: the operand tables (all classes but the step start, which keeps
  // them in the cache), then the staged coefficients; the staged G of the
  // forward transforms overlays both once they are spent.
  constexpr bool kTabLds = CLASS != kClassStart;
  constexpr int kTab = JT * KS * kWave;
  double* stageD = lds + (kTabLds ? 2 * kTab : 0);
  double2* stage = reinterpret_cast<double2*>(stageD);
  double* Gd = lds;
  const double2* Gs = reinterpret_cast<const double2*>(lds);
  const double* tabC = a.tables.tabC;
  const double* tabS = a.tables.tabS;
  if constexpr (kTabLds) {
    for (int i = threadIdx.x; i < kTab; i += blockDim.x) {
      lds[i] = a.tables.tabC[i];
      lds[kTab + i] = a.tables.tabS[i];
    }
    tabC = lds;
    tabS = lds + kTab;
  }
  const bool synth = a.op == kRingSynthesise || a.op == kRingPass ||
                     a.op == kRingFinish || a.op == kRingStartFromSpectral;
  if (synth) {
    const int count = g.orders * g.ringsPerGroup;
    for (int e = threadIdx.x; e < count; e += blockDim.x) {
      const int ring = e % g.ringsPerGroup, mo = e / g.ringsPerGroup;
      const int th = th0 + ring;
      double2 v = make_double2(0.0, 0.0);
      if (th < a.ntheta) {
        const double* src =
            a.fourierIn + ((static_cast<size_t>(mo) * a.nodes + node) * a.ntheta + th) * 2;
        v = make_double2(src[0], src[1]);
      }
      stage[ring * g.orders + mo] = v;
    }
  }
  __syncthreads();

  // x at j = 1 + 16 t + 4 q + quad (index v = 4 t + q), at P - j, and at 0
  double xj[KS], xq[KS], x0 = 0.0;
  const int th = th0 + ln.ring;
  const size_t ringBase = (static_cast<size_t>(node) * a.ntheta + (ln.valid ? th : 0)) * g.nphi;
  // The lane's points: j = 1 + 16 t + 4 q + quad (v = 4 t + q), P - j, and
  // 0. Where a lane has no point (padding, a ring past the last) the index
  // is its ring's point 0, always in bounds: loads are unconditional and
  // issue back to back, only stores are predicated.
  const size_t p0 = ringBase + ln.r;
  size_t pj[KS], pq[KS];
  bool lv[KS];
#pragma unroll
  for (int v = 0; v < KS; ++v) {
    const int j = 1 + 16 * (v / 4) + 4 * (v % 4) + ln.quad;
    lv[v] = ln.valid && j <= g.H;
    pj[v] = lv[v] ? ringBase + g.R * j + ln.r : p0;
    pq[v] = lv[v] ? ringBase + g.R * (g.P - j) + ln.r : p0;
  }
  auto element = [&](size_t pt) -> size_t { return REAL ? pt : 2 * pt + ln.p; };
  if (synth) {
    double av[KS], bv[KS];
    double sum = 0.0;
#pragma unroll
    for (int kb = 0; kb < KS; ++kb) {
      const int k = 1 + 4 * kb + ln.quad;
      av[kb] = 0.0;
      bv[kb] = 0.0;
      if (k <= g.H) {
        const double2 hk = Fold<REAL>(g, stage, cosN, sinN, ln.ringA, ln.r, k);
        const double2 hm = Fold<REAL>(g, stage, cosN, sinN, ln.ringA, ln.r, g.P - k);
        av[kb] = ln.p == 0 ? hk.x + hm.x : hk.y + hm.y;
        bv[kb] = ln.p == 0 ? hk.x - hm.x : hk.y - hm.y;
      }
      sum += av[kb];
    }
    const double2 h0c = Fold<REAL>(g, stage, cosN, sinN, ln.ringA, ln.r, 0);
    const double h0 = ln.p == 0 ? h0c.x : h0c.y;
    x0 = h0 + ColumnSum(sum);
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
        const int at = (t * KS + kb) * kWave + ln.lane;
        C[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabC[at], av[kb], C[t], 0, 0, 0);
        S[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(tabS[at], bv[kb], S[t], 0, 0, 0);
      }
    }
    // x_j = h0 + C_j + i S_j, x_{P-j} = h0 + C_j - i S_j
    const double sg = ln.p == 0 ? -1.0 : 1.0;
#pragma unroll
    for (int t = 0; t < JT; ++t) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double sx = Shfl(S[t][q], 1);
        xj[4 * t + q] = h0 + C[t][q] + sg * sx;
        xq[4 * t + q] = h0 + C[t][q] - sg * sx;
      }
    }
  } else {
    // a stored ring (complex: interleaved parts; real: the lane's own ring)
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      const double vj = a.gridIn[element(pj[v])], vq = a.gridIn[element(pq[v])];
      xj[v] = lv[v] ? vj : 0.0;
      xq[v] = lv[v] ? vq : 0.0;
    }
    const double v0 = a.gridIn[element(p0)];
    x0 = ln.valid ? v0 : 0.0;
  }
  __syncthreads();  // the staged coefficients are read; the stage is reused

  // the pointwise operation at the lane's points
  const bool own0 = ln.valid && ln.quad == 0;  // the lane that stores point 0

  if constexpr (CLASS == kClassSynth) {
    if (a.op == kRingFinish) {
      // d on the grid and m+ = m~ + beta d in place
      double bj[KS], bq[KS], mj[KS], mq[KS];
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        bj[v] = a.beta[pj[v]];
        bq[v] = a.beta[pq[v]];
        mj[v] = a.memOut[element(pj[v])];
        mq[v] = a.memOut[element(pq[v])];
      }
      const double b0 = a.beta[p0], m0 = a.memOut[element(p0)];
      if (own0) {
        a.gridOut[element(p0)] = x0;
        a.memOut[element(p0)] = m0 + b0 * x0;
      }
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        if (lv[v]) {
          a.gridOut[element(pj[v])] = xj[v];
          a.gridOut[element(pq[v])] = xq[v];
          a.memOut[element(pj[v])] = mj[v] + bj[v] * xj[v];
          a.memOut[element(pq[v])] = mq[v] + bq[v] * xq[v];
        }
      }
    } else {
      if (own0) a.gridOut[element(p0)] = x0;
#pragma unroll
      for (int v = 0; v < KS; ++v) {
        if (lv[v]) {
          a.gridOut[element(pj[v])] = xj[v];
          a.gridOut[element(pq[v])] = xq[v];
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
    // every point's itau and m first (unconditional loads, issued together)
    double hj[KS], hq[KS], mInJ[KS], mInQ[KS];
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      hj[v] = a.itau[pj[v]];
      hq[v] = a.itau[pq[v]];
      mInJ[v] = a.memIn[element(pj[v])];
      mInQ[v] = a.memIn[element(pq[v])];
    }
    const double h0 = a.itau[p0], mIn0 = a.memIn[element(p0)];
    // returns W d0 and writes m~, (REAL) beta and W; mt gets mu m~
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
      mt = mu * mNew;
      return mu * (bw - refBeta) * d0;
    };
    double yj[KS], yq[KS], mj[KS], mq[KS], y0 = 0.0, m0 = 0.0;
    y0 = start(p0, h0, mIn0, x0, own0, m0);
    if (!ln.valid) y0 = m0 = 0.0;
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      yj[v] = start(pj[v], hj[v], mInJ[v], xj[v], lv[v], mj[v]);
      yq[v] = start(pq[v], hq[v], mInQ[v], xq[v], lv[v], mq[v]);
      if (!lv[v]) yj[v] = yq[v] = mj[v] = mq[v] = 0.0;
    }
    // the analysis of W d0 (:1361-1376) ...
    ForwardToStage<JT, false>(g, ln, tabC, tabS, yj, yq, y0, Gd);
    __syncthreads();
    CombineFromStage<REAL>(g, a, cosN, sinN, Gs, node, th0, a.fourierOut2);
    __syncthreads();
    // ... and of the fixed memory force mu m~ (:1355-1356)
    ForwardToStage<JT, false>(g, ln, tabC, tabS, mj, mq, m0, Gd);
    __syncthreads();
    CombineFromStage<REAL>(g, a, cosN, sinN, Gs, node, th0, a.fourierOut);
  } else {
  // y = factor x (pass and weighted: the point's weight; node scaled: mu)
  double yj[KS], yq[KS], y0 = x0;
  if (a.op == kRingPass || a.op == kRingWeighted) {
    double wj[KS], wq[KS];
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      wj[v] = a.weight[pj[v]];
      wq[v] = a.weight[pq[v]];
    }
    const double w0 = a.weight[p0];
    y0 = ln.valid ? w0 * x0 : 0.0;
#pragma unroll
    for (int v = 0; v < KS; ++v) {
      yj[v] = lv[v] ? wj[v] * xj[v] : 0.0;
      yq[v] = lv[v] ? wq[v] * xq[v] : 0.0;
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
  ForwardToStage<JT, true>(g, ln, tabC, tabS, yj, yq, y0, Gd);
  __syncthreads();
  CombineFromStage<REAL>(g, a, cosN, sinN, Gs, node, th0, a.fourierOut);
  }
}

template <int JT, int CLASS>
void LaunchClass(const RingLaunch& r, dim3 grid, size_t lds, hipStream_t stream) {
  if (r.block == 2) {
    hipLaunchKernelGGL((RingMfmaKernel<JT, true, CLASS>), grid, dim3(kThreads), lds, stream,
                       r);
  } else {
    hipLaunchKernelGGL((RingMfmaKernel<JT, false, CLASS>), grid, dim3(kThreads), lds, stream,
                       r);
  }
}

template <int JT>
void LaunchJT(const RingLaunch& r, dim3 grid, size_t lds, hipStream_t stream) {
  switch (r.op) {
    case kRingSynthesise:
    case kRingFinish:
      LaunchClass<JT, kClassSynth>(r, grid, lds, stream);
      break;
    case kRingStart:
    case kRingStartFromSpectral:
      LaunchClass<JT, kClassStart>(r, grid, lds, stream);
      break;
    default:
      LaunchClass<JT, kClassOne>(r, grid, lds, stream);
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
  const bool real = r.block == 2;
  const int rpg = kWaves * (real ? 2 * kSeq : kSeq) / r.tables.R;
  const int orders = real ? r.lmax + 1 : 2 * r.lmax + 1;
  const bool synth = r.op == kRingSynthesise || r.op == kRingPass || r.op == kRingFinish ||
                     r.op == kRingStartFromSpectral;
  const bool start = r.op == kRingStart || r.op == kRingStartFromSpectral;
  const bool analysis = r.op != kRingSynthesise && r.op != kRingFinish;
  const size_t tables =
      start ? 0 : 2 * sizeof(double) * (4 * r.tables.jt * r.tables.jt) * kWave;  // 2 x JT x KS x 64
  const size_t staged = synth ? 16 * static_cast<size_t>(rpg) * orders : 0;
  const size_t g = analysis ? 16 * static_cast<size_t>(kWaves) * kSeq * r.tables.P : 0;
  return std::max(tables + staged, g);
}

void LaunchRingMfma(const RingLaunch& r, hipStream_t stream) {
  const int R = r.tables.R;
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
    default: throw ArgumentError("ring transform: no matrix-core tiling for this grid");
  }
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
