// Generalised spherical-harmonic transforms for gfx950, FP64.
//
// Synthesis of spin block N: F_m(theta) = sum_l T[m,l](theta) f_lm, then
// f(theta, phi_j) = sum_m F_m(theta) e^{i m phi_j}; analysis the adjoint
// with the quadrature weights (SurfaceGrid.hpp:141-252 is the contract;
// the reference runs GSHTrans loops on the CPU).
//
// Legendre stage: per order m a real GEMM on the matrix cores
// (legendre_mfma.hip).
//
// Longitude stage: nphi = R P (R = 4 or 2), decimation in time: the R
// P-point DFTs of x_{R n + r} by the symmetric split (pairs n, P - n give a
// cosine sum of x_n + x_{P-n} and a sine sum of x_n - x_{P-n}, half the
// multiplies of a direct DFT), then the twiddled combination for the
// orders |m| <= lmax only. A workgroup holds RB rings (consecutive
// colatitudes of one radial node) in LDS.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "transform.hpp"

namespace sl3amd {

// ---- longitude DFT ---------------------------------------------------------

namespace {

constexpr int kThreads = 256;

struct RingArgs {
  DftTables t;
  int lmax, nodes, ntheta, nphi;
  int fstride;  // the Fourier layout's colatitude stride
  int block;  // 0, 1 complex spins; 2 real spin 0
  int mlo;
  double scale;           // analysis: 2 pi / nphi
  const double* wtheta;   // analysis: colatitude weights
  // fields
  const double* fourierIn;  // synthesis input
  double* fourierOut;       // analysis output
  const double* gridIn;     // analysis input ring (complex or real)
  double* gridOut;          // synthesis output ring
  const double* weight;     // pointwise weight applied between (pass)
  int op;                   // RingOp
  // step start / finish operands
  const double* memIn;   // m (grid, this block)
  double* memOut;        // m~ (start) / m+ (finish, in place)
  const double* itau;    // per point (start)
  double* betaOut;       // per point (start, block 2 writes)
  double* weightOut;     // per point (start, block 2 writes)
  const double* beta;    // per point (finish)
  const double* muNode;  // branch modulus per node (start)
  const double* itauRefNode;
  double dt;
  int scheme;
  double* fourierOut2;   // start: W d0's transform beside the fixed force's
  const int* live;       // per realisation, 0: converged, skipped (null: none)
  int liveNodes;
};

// The LDS of one workgroup: RB rings of nphi complex, their R P-point
// sub-transforms, and the twiddle tables.
constexpr int kRings = 4;

struct RingShared {
  double2* ring;   // [RB][nphi]
  double2* sub;    // [RB][nphi] (R sub-transforms of P)
  double* cosP;    // [P]
  double* sinP;    // [P]
  double* cosN;    // [nphi]
  double* sinN;    // [nphi]
};

__device__ inline double2 cmul(double2 a, double c, double s) {
  // a * (c + i s)
  return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// Forward: sub[ring][r * P + k] = sum_n ring[ring][R n + r] e^{-2 pi i n k / P}.
__device__ void ForwardSubDfts(const RingShared& sh, int rings, int R, int P) {
  const int H = (P - 1) / 2;
  const bool even = (P % 2) == 0;
  const int nphi = R * P;
  const int kCount = H + 1 + (even ? 1 : 0);  // k = 0, pairs 1..H, P/2
  const int tasks = rings * R * kCount;
  for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
    const int kk = task % kCount;
    const int r = (task / kCount) % R;
    const int g = task / (kCount * R);
    const double2* x = sh.ring + g * nphi;
    double2* out = sh.sub + g * nphi + r * P;
    const double2 x0 = x[r];
    if (kk == 0) {
      double2 s = x0;
      for (int n = 1; n < P; ++n) {
        s.x += x[R * n + r].x;
        s.y += x[R * n + r].y;
      }
      out[0] = s;
    } else if (kk <= H) {
      const int k = kk;
      double2 C = make_double2(0.0, 0.0), S = make_double2(0.0, 0.0);
      int idx = 0;
      for (int n = 1; n <= H; ++n) {
        idx += k;
        if (idx >= P) idx -= P;
        const double2 xn = x[R * n + r], xm = x[R * (P - n) + r];
        const double c = sh.cosP[idx], s = sh.sinP[idx];
        C.x += (xn.x + xm.x) * c;
        C.y += (xn.y + xm.y) * c;
        S.x += (xn.x - xm.x) * s;
        S.y += (xn.y - xm.y) * s;
      }
      if (even) {
        const double2 xh = x[R * (P / 2) + r];
        const double sg = (k & 1) ? -1.0 : 1.0;
        C.x += sg * xh.x;
        C.y += sg * xh.y;
      }
      // e^{-i a} = cos - i sin: G(k) = x0 + C - i S, G(P-k) = x0 + C + i S
      out[k] = make_double2(x0.x + C.x + S.y, x0.y + C.y - S.x);
      out[P - k] = make_double2(x0.x + C.x - S.y, x0.y + C.y + S.x);
    } else {  // k = P/2, even P
      double2 s = x0;
      for (int n = 1; n < P; ++n) {
        const double sg = (n & 1) ? -1.0 : 1.0;
        s.x += sg * x[R * n + r].x;
        s.y += sg * x[R * n + r].y;
      }
      out[P / 2] = s;
    }
  }
}

// Inverse: ring[g][R j + r] = sum_k sub[g][r P + k] e^{+2 pi i j k / P}.
__device__ void InverseSubDfts(const RingShared& sh, int rings, int R, int P) {
  const int H = (P - 1) / 2;
  const bool even = (P % 2) == 0;
  const int nphi = R * P;
  const int jCount = H + 1 + (even ? 1 : 0);
  const int tasks = rings * R * jCount;
  for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
    const int jj = task % jCount;
    const int r = (task / jCount) % R;
    const int g = task / (jCount * R);
    const double2* h = sh.sub + g * nphi + r * P;
    double2* x = sh.ring + g * nphi;
    const double2 h0 = h[0];
    if (jj == 0) {
      double2 s = h0;
      for (int k = 1; k < P; ++k) {
        s.x += h[k].x;
        s.y += h[k].y;
      }
      x[r] = s;
    } else if (jj <= H) {
      const int j = jj;
      double2 C = make_double2(0.0, 0.0), S = make_double2(0.0, 0.0);
      int idx = 0;
      for (int k = 1; k <= H; ++k) {
        idx += j;
        if (idx >= P) idx -= P;
        const double2 hk = h[k], hm = h[P - k];
        const double c = sh.cosP[idx], s = sh.sinP[idx];
        C.x += (hk.x + hm.x) * c;
        C.y += (hk.y + hm.y) * c;
        S.x += (hk.x - hm.x) * s;
        S.y += (hk.y - hm.y) * s;
      }
      if (even) {
        const double2 hh = h[P / 2];
        const double sg = (j & 1) ? -1.0 : 1.0;
        C.x += sg * hh.x;
        C.y += sg * hh.y;
      }
      // e^{+i a}: f_j = h0 + C + i S, f_{P-j} = h0 + C - i S
      x[R * j + r] = make_double2(h0.x + C.x - S.y, h0.y + C.y + S.x);
      x[R * (P - j) + r] = make_double2(h0.x + C.x + S.y, h0.y + C.y - S.x);
    } else {
      double2 s = h0;
      for (int k = 1; k < P; ++k) {
        const double sg = (k & 1) ? -1.0 : 1.0;
        s.x += sg * h[k].x;
        s.y += sg * h[k].y;
      }
      x[R * (P / 2) + r] = s;
    }
  }
}

// Synthesis: ring values from F_m (m = mlo..lmax; the real block's negative
// orders are the conjugates) of RB rings.
__device__ void FoldForInverse(const RingShared& sh, const RingArgs& a, int node,
                               int th0, int rings, int R, int P) {
  const int nphi = R * P;
  const int tasks = rings * R * P;
  const bool real = a.block == 2;
  for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
    const int k = task % P;
    const int r = (task / P) % R;
    const int g = task / (P * R);
    const int th = th0 + g;
    double2 s = make_double2(0.0, 0.0);
    // m = k + q P within [-lmax, lmax]
    int m = k;
    while (m - P >= -a.lmax) m -= P;
    for (; m <= a.lmax; m += P) {
      double2 f;
      if (real) {
        const int am = m < 0 ? -m : m;
        const double* src =
            a.fourierIn + ((static_cast<size_t>(am) * a.nodes + node) * a.fstride + th) * 2;
        f = make_double2(src[0], m < 0 ? -src[1] : src[1]);
      } else {
        const double* src =
            a.fourierIn +
            ((static_cast<size_t>(m - a.mlo) * a.nodes + node) * a.fstride + th) * 2;
        f = make_double2(src[0], src[1]);
      }
      // e^{+2 pi i r m / nphi}
      int idx = (r * m) % nphi;
      if (idx < 0) idx += nphi;
      const double c = sh.cosN[idx], s2 = sh.sinN[idx];
      const double2 t = cmul(f, c, s2);
      s.x += t.x;
      s.y += t.y;
    }
    sh.sub[g * nphi + r * P + k] = s;
  }
}

// Analysis: F_m = scale w_theta sum_r e^{-2 pi i r m / nphi} G_r(m mod P).
__device__ void CombineForward(const RingShared& sh, const RingArgs& a,
                               double* out, int node, int th0, int rings, int R,
                               int P) {
  const int nphi = R * P;
  const int orders = a.lmax - a.mlo + 1;
  const int tasks = rings * orders;
  for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
    const int g = task % rings;
    const int m = a.mlo + task / rings;
    const int th = th0 + g;
    int k = m % P;
    if (k < 0) k += P;
    double2 s = make_double2(0.0, 0.0);
    for (int r = 0; r < R; ++r) {
      int idx = (r * m) % nphi;
      if (idx < 0) idx += nphi;
      // e^{-i a} = cos - i sin
      const double2 t = cmul(sh.sub[g * nphi + r * P + k], sh.cosN[idx], -sh.sinN[idx]);
      s.x += t.x;
      s.y += t.y;
    }
    const double w = a.scale * a.wtheta[th];
    double* dst =
        out + ((static_cast<size_t>(m - a.mlo) * a.nodes + node) * a.fstride + th) * 2;
    dst[0] = s.x * w;
    dst[1] = s.y * w;
  }
}

__device__ void LoadTables(const RingShared& sh, const RingArgs& a, int P) {
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    sh.cosP[i] = a.t.cosP[i];
    sh.sinP[i] = a.t.sinP[i];
  }
  for (int i = threadIdx.x; i < a.nphi; i += blockDim.x) {
    sh.cosN[i] = a.t.cosN[i];
    sh.sinN[i] = a.t.sinN[i];
  }
}

__device__ RingShared CarveShared(double* lds, int nphi, int P) {
  RingShared sh;
  sh.ring = reinterpret_cast<double2*>(lds);
  sh.sub = sh.ring + kRings * nphi;
  double* rest = reinterpret_cast<double*>(sh.sub + kRings * nphi);
  sh.cosP = rest;
  sh.sinP = rest + P;
  sh.cosN = rest + 2 * P;
  sh.sinN = rest + 2 * P + nphi;
  return sh;
}

// One launch shape for every ring operation: block (node, theta chunk).
__global__ void __launch_bounds__(kThreads) RingKernel(RingArgs a) {
  extern __shared__ double lds[];
  const int R = a.t.R, P = a.t.P, nphi = a.nphi;
  const RingShared sh = CarveShared(lds, nphi, P);
  const int node = blockIdx.y;
  if (a.live && a.live[node / a.liveNodes] == 0) return;  // a converged realisation
  const int th0 = blockIdx.x * kRings;
  const int rings = min(kRings, a.ntheta - th0);
  const bool real = a.block == 2;
  const size_t ringBase = (static_cast<size_t>(node) * a.ntheta + th0) * nphi;
  const int count = rings * nphi;
  LoadTables(sh, a, P);
  __syncthreads();  // every twiddle in LDS before any lane reads one

  // 1. the ring values
  if (a.op == kRingSynthesise || a.op == kRingPass || a.op == kRingFinish ||
      a.op == kRingStartFromSpectral) {
    FoldForInverse(sh, a, node, th0, rings, R, P);
    __syncthreads();
    InverseSubDfts(sh, rings, R, P);
    __syncthreads();
    if (real) {
      for (int i = threadIdx.x; i < count; i += blockDim.x) sh.ring[i].y = 0.0;
    }
  } else {
    // analysis of a stored ring
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      sh.ring[i] = real ? make_double2(a.gridIn[ringBase + i], 0.0)
                        : reinterpret_cast<const double2*>(a.gridIn)[ringBase + i];
    }
  }
  __syncthreads();

  // 2. the pointwise operation
  if (a.op == kRingSynthesise) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      if (real) {
        a.gridOut[ringBase + i] = sh.ring[i].x;
      } else {
        reinterpret_cast<double2*>(a.gridOut)[ringBase + i] = sh.ring[i];
      }
    }
    return;
  }
  if (a.op == kRingNodeScaled) {
    const double mu = a.muNode[node];
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      sh.ring[i].x *= mu;
      sh.ring[i].y *= mu;
    }
  } else if (a.op == kRingPass || a.op == kRingWeighted) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      const double w = a.weight[ringBase + i];
      sh.ring[i].x *= w;
      sh.ring[i].y *= w;
    }
  } else if (a.op == kRingFinish) {
    // d on the grid (the new state's strain) and m+ = m~ + beta d in place
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      const double bt = a.beta[ringBase + i];
      const double2 d = sh.ring[i];
      if (real) {
        a.gridOut[ringBase + i] = d.x;
        a.memOut[ringBase + i] = a.memOut[ringBase + i] + bt * d.x;
      } else {
        reinterpret_cast<double2*>(a.gridOut)[ringBase + i] = d;
        double2* mo = reinterpret_cast<double2*>(a.memOut) + ringBase + i;
        const double2 mv = *mo;
        *mo = make_double2(mv.x + bt * d.x, mv.y + bt * d.y);
      }
    }
    return;
  } else if (a.op == kRingStart || a.op == kRingStartFromSpectral) {
    // src/Viscoelastic3D.cpp:1283-1332, one branch: the weights of
    // h = dt itau(x), m~ = e m + alpha d0, beta, W = mu (beta - beta_ref),
    // the fixed force mu m~; then W d0 (:1361-1365). ring holds d0.
    const double mu = a.muNode[node];
    const double href = a.dt * a.itauRefNode[node];
    double refBeta;
    if (a.scheme == 1) {
      refBeta = 1.0 - 1.0 / (1.0 + href);
    } else {
      const double e = exp(-href);
      const double p = fabs(href) < 1e-3
                           ? 1.0 - href / 2.0 + href * href / 6.0 -
                                 href * href * href / 24.0 +
                                 href * href * href * href / 120.0
                           : (1.0 - e) / href;
      refBeta = 1.0 - p;
    }
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      const double h = a.dt * a.itau[ringBase + i];
      double ew, aw, bw;
      if (a.scheme == 1) {
        const double w = 1.0 / (1.0 + h);
        ew = w;
        aw = 0.0;
        bw = 1.0 - w;
      } else {
        const double e = exp(-h);
        const double p = fabs(h) < 1e-3 ? 1.0 - h / 2.0 + h * h / 6.0 -
                                              h * h * h / 24.0 +
                                              h * h * h * h / 120.0
                                        : (1.0 - e) / h;
        ew = e;
        aw = p - e;
        bw = 1.0 - p;
      }
      const double2 d0 = sh.ring[i];
      if (real) {
        const double mv = a.memIn[ringBase + i];
        a.memOut[ringBase + i] = ew * mv + aw * d0.x;
      } else {
        const double2 mv = reinterpret_cast<const double2*>(a.memIn)[ringBase + i];
        reinterpret_cast<double2*>(a.memOut)[ringBase + i] =
            make_double2(ew * mv.x + aw * d0.x, ew * mv.y + aw * d0.y);
      }
      const double W = mu * (bw - refBeta);
      if (real) {
        a.betaOut[ringBase + i] = bw;
        a.weightOut[ringBase + i] = W;
      }
      sh.ring[i] = make_double2(W * d0.x, W * d0.y);
    }
    __syncthreads();
    // the analysis of W d0 (:1361-1376) ...
    ForwardSubDfts(sh, rings, R, P);
    __syncthreads();
    CombineForward(sh, a, a.fourierOut2, node, th0, rings, R, P);
    __syncthreads();
    // ... and of the fixed memory force mu m~ (:1355-1356); each thread
    // reads back the m~ values it wrote.
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      const double2 mt =
          real ? make_double2(a.memOut[ringBase + i], 0.0)
               : reinterpret_cast<const double2*>(a.memOut)[ringBase + i];
      sh.ring[i] = make_double2(mu * mt.x, mu * mt.y);
    }
    __syncthreads();
    ForwardSubDfts(sh, rings, R, P);
    __syncthreads();
    CombineForward(sh, a, a.fourierOut, node, th0, rings, R, P);
    return;
  }
  __syncthreads();
  // 3. analysis of the ring
  ForwardSubDfts(sh, rings, R, P);
  __syncthreads();
  CombineForward(sh, a, a.fourierOut, node, th0, rings, R, P);
}

}  // namespace

// ---- host side --------------------------------------------------------------

size_t RingLds(int nphi, int P) {
  return sizeof(double) * (4 * kRings * nphi + 2 * P + 2 * nphi);
}

void LaunchLegendreSynthesis(const LegendrePlan& p, const double* const spectral[3],
                             double* const fourier[3], const int* reach,
                             hipStream_t stream) {
  ProfileScope prof("legendre_synthesis", stream);
  LaunchLegendreSynthesisMfma(p, spectral, fourier, reach, stream);
}

void LaunchLegendreAnalysis(const LegendrePlan& p, const double* const fourier[3],
                            double* const spectral[3], const double* const addend[3],
                            const int* reach, const int* firstNode, hipStream_t stream) {
  ProfileScope prof("legendre_analysis", stream);
  LaunchLegendreAnalysisMfma(p, fourier, spectral, addend, reach, firstNode, stream);
}

void LaunchRing(const RingLaunch& r, hipStream_t stream) {
  if (r.tables.jt > 0 && RingMfmaLds(r) <= 160 * 1024) {
    LaunchRingMfma(r, stream);
    return;
  }
  RingArgs a{};
  a.t = r.tables;
  a.lmax = r.lmax;
  a.nodes = r.nodes;
  a.ntheta = r.ntheta;
  a.fstride = r.fstride;
  a.nphi = r.nphi;
  a.block = r.block;
  a.mlo = r.block == 2 ? 0 : -r.lmax;
  a.scale = 2.0 * 3.14159265358979323846 / r.nphi;
  a.wtheta = r.wtheta;
  a.fourierIn = r.fourierIn;
  a.fourierOut = r.fourierOut;
  a.fourierOut2 = r.fourierOut2;
  a.gridIn = r.gridIn;
  a.gridOut = r.gridOut;
  a.weight = r.weight;
  a.op = r.op;
  a.memIn = r.memIn;
  a.memOut = r.memOut;
  a.itau = r.itau;
  a.betaOut = r.betaOut;
  a.weightOut = r.weightOut;
  a.beta = r.beta;
  a.muNode = r.muNode;
  a.itauRefNode = r.itauRefNode;
  a.dt = r.dt;
  a.scheme = r.scheme;
  a.live = r.live;
  a.liveNodes = r.liveNodes;
  const dim3 grid(CeilDiv(r.ntheta, kRings), r.nodes);
  ProfileScope prof692("ring", stream);
  hipLaunchKernelGGL(RingKernel, grid, dim3(kThreads), RingLds(r.nphi, r.tables.P),
                     stream, a);
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
