// The grid coupling's per-degree maps and the fixed point's vector sweeps
// for gfx950, FP64 (all HBM-bound):
//  * the memory force (GridCoupling::SpheroidalForce / ToroidalForce with
//    Gather and ForceRowsFromComponents, src/GridCoupling.cpp:613-809,
//    include/sl3dpp/SpinComponents.hpp:124-145), written as rows of the
//    memory dofs, the right-hand sides of the memory response;
//  * strain recovery (GridCoupling::AddStrainOfSolution and Scatter,
//    src/GridCoupling.cpp:595-689; DeviatoricComponents,
//    SpinComponents.hpp:53-72) from the response on the memory dofs plus
//    the load channel (AddUnitChannels, src/Viscoelastic3D.cpp:1029-1073);
//  * RelativeChange (src/Viscoelastic3D.cpp:1636-1669) and the Anderson
//    sweeps (include/sl3dpp/detail/AndersonMixer.hpp:123-218), reduced in a
//    fixed block order so that results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <vector>

#include "common.hpp"
#include "coupling.hpp"

namespace sl3amd {

namespace {

constexpr int kThreads = 256;
constexpr double kSqrt2 = 1.41421356237309504880;

struct Cplx {
  double re, im;
};
__device__ inline Cplx C(double re, double im) { return {re, im}; }
__device__ inline Cplx operator+(Cplx a, Cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline Cplx operator-(Cplx a, Cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline Cplx operator*(double s, Cplx a) { return {s * a.re, s * a.im}; }
__device__ inline Cplx TimesI(Cplx a) { return {-a.im, a.re}; }
__device__ inline Cplx Conj(Cplx a) { return {a.re, -a.im}; }
__device__ inline Cplx Div(Cplx a, double d) { return {a.re / d, a.im / d}; }

__device__ inline int LMin(int m, int spin) {
  const int am = m < 0 ? -m : m, an = spin < 0 ? -spin : spin;
  return am > an ? am : an;
}

__device__ inline size_t RowOf(const SpectralView& v, int b, int l, int m) {
  const int mlo = b == 2 ? 0 : -v.lmax;
  const int spin = b == 0 ? -2 : (b == 1 ? -1 : 0);
  return static_cast<size_t>(v.rowOffset[b][m - mlo] + l - LMin(m, spin));
}

__device__ inline Cplx Load(const SpectralView& v, int b, int l, int m, int node,
                            int nodes) {
  const double* p = v.block[b] + 2 * (RowOf(v, b, l, m) * nodes + node);
  return C(p[0], p[1]);
}

__device__ inline void Store(const SpectralView& v, int b, int l, int m, int node,
                             int nodes, Cplx z) {
  double* p = v.block[b] + 2 * (RowOf(v, b, l, m) * nodes + node);
  p[0] = z.re;
  p[1] = z.im;
}

// GridCoupling::Gather (src/GridCoupling.cpp:613-633): the five components
// of column m >= 0, the derived ones by A^{(N)}_{lm} = (-1)^m
// conj(A^{(-N)}_{l,-m}).
__device__ inline void Gather(const SpectralView& v, int l, int m, int node,
                              int nodes, Cplx A[5]) {
  const double sign = (m & 1) ? -1.0 : 1.0;
  A[0] = A[1] = A[3] = A[4] = C(0.0, 0.0);
  A[2] = Load(v, 2, l, m, node, nodes);
  if (l >= 1) {
    A[1] = Load(v, 1, l, m, node, nodes);
    A[3] = sign * Conj(Load(v, 1, l, -m, node, nodes));
  }
  if (l >= 2) {
    A[0] = Load(v, 0, l, m, node, nodes);
    A[4] = sign * Conj(Load(v, 0, l, -m, node, nodes));
  }
}

// ForceRowsFromComponents: u, v, w, du, dv, dw.
__device__ inline void ForceRows(int l, const Cplx A[5], double r, Cplx rows[6]) {
  const double zeta2 = l * (l + 1.0);
  const double zeta = sqrt(zeta2);
  const double root = sqrt(zeta2 - 2.0 > 0.0 ? zeta2 - 2.0 : 0.0);
  const Cplx p13 = A[1] + A[3], m13 = A[1] - A[3];
  const Cplx p04 = A[0] + A[4], m04 = A[0] - A[4];
  rows[0] = (kSqrt2 * zeta) * p13 - 4.0 * A[2];
  rows[1] = (zeta * root) * p04 + (2.0 * zeta2) * A[2] - (kSqrt2 * zeta) * p13;
  rows[2] = TimesI((zeta * root) * m04 - (kSqrt2 * zeta) * m13);
  rows[3] = (4.0 * r) * A[2];
  rows[4] = (kSqrt2 * zeta * r) * p13;
  rows[5] = TimesI((kSqrt2 * zeta * r) * m13);
}

constexpr int kChunkElements = 4;
constexpr int kMaxPoints = 6;  // memory points per element the LDS holds
constexpr int kChunkOrders = 8;
constexpr int kMaxBasis = 96;  // np x ngll GLL table entries the LDS holds

// One workgroup: a chunk of memory elements, one degree l and kChunkOrders
// orders, for the spheroidal and the toroidal system together (one read of
// the analysed tensor). Stage 1 forms the force rows of every point of the
// chunk (and of the element below, whose top node the chunk's bottom node
// shares) in LDS; stage 2 sums each owned GLL node's rows.
__global__ void __launch_bounds__(kThreads)
    ForceKernel(const DegreeTask* tasks, RadialGeometry g, SpectralView w, const int* dofs,
                const int2* work, double* force, const int* live) {
  constexpr int maxPoints = (kChunkElements + 1) * kMaxPoints;
  __shared__ Cplx rowsL[maxPoints * kChunkOrders][6];
  // per staged point its radius and jac weight radius; the GLL value and
  // derivative tables (when they fit) and the reference weights
  __shared__ double sRadius[maxPoints], sPoint[maxPoints];
  __shared__ double sValue[kMaxBasis], sDeriv[kMaxBasis], sWeight[kMaxPoints];
  static_assert(maxPoints <= kThreads && kMaxBasis <= kThreads, "one thread per staged entry");
  const int2 wk = work[blockIdx.y];
  const int l = wk.x, mc = wk.y;
  const int real = blockIdx.z;  // the realisation
  if (live && live[real] == 0) return;  // converged
  const int width = g.nodes * g.realisations;
  const int L = w.lmax;
  const DegreeTask ts = tasks[l];
  const bool hasTor = l >= 1;
  const DegreeTask tt = hasTor ? tasks[L + l] : ts;  // toroidal l at L + 1 + (l - 1)
  const int elemEnd = g.elem0 + g.elements;
  const int eLo = max(g.elem0, ts.firstElement);
  const int e0 = g.elem0 + blockIdx.x * kChunkElements;
  const int e1 = min(e0 + kChunkElements, elemEnd);
  if (e1 <= eLo) return;
  const int eStart = max(eLo, e0 - 1);
  const int points = (e1 - eStart) * g.np;
  const int nm = g.ngll - 1;
  // owned GLL nodes: [gidx(max(e0, eLo), 0), gidx(e1, 0)) plus the top node
  // of the last memory element
  const int gnFirst = max(e0, eLo) * nm;
  const int gnEnd = e1 * nm + (e1 == elemEnd ? 1 : 0);
  const int owned = gnEnd - gnFirst;
  const int mCount = min(kChunkOrders, l + 1 - mc);
  // the geometry stage 2 reads, loaded here and stored after the gathers
  // (one thread per entry: points, the basis and np all fit the workgroup)
  const int basis = g.np * g.ngll;
  const bool stagedBasis = basis <= kMaxBasis;
  const int t = threadIdx.x;
  double tRadius = 0.0, tPoint = 0.0, tValue = 0.0, tDeriv = 0.0, tWeight = 0.0;
  if (t < points) {
    const int elem = eStart + t / g.np, q = t % g.np;
    tRadius = g.radius[(elem - g.elem0) * g.np + q];
    tPoint = g.jac[elem - g.elem0] * g.weight[q] * tRadius;
  }
  if (stagedBasis && t < basis) {
    tValue = g.value[t];
    tDeriv = g.deriv[t];
  }
  if (t < g.np) tWeight = g.weight[t];
  // stage 1: rows of (point, m)
  for (int e = threadIdx.x; e < points * mCount; e += blockDim.x) {
    const int p = e % points, mi = e / points;
    const int elem = eStart + p / g.np, q = p % g.np;
    const int node = (elem - g.elem0) * g.np + q;
    Cplx A[5], rows[6];
    Gather(w, l, mc + mi, real * g.nodes + node, width, A);
    ForceRows(l, A, g.radius[node], rows);
    Cplx* dst = rowsL[mi * maxPoints + p];
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[k] = rows[k];
  }
  if (t < points) {
    sRadius[t] = tRadius;
    sPoint[t] = tPoint;
  }
  if (stagedBasis && t < basis) {
    sValue[t] = tValue;
    sDeriv[t] = tDeriv;
  }
  if (t < g.np) sWeight[t] = tWeight;
  __syncthreads();
  const double* value = stagedBasis ? sValue : g.value;
  const double* derivTable = stagedBasis ? sDeriv : g.deriv;
  // stage 2: each owned node's sums, order fastest (consecutive threads
  // write consecutive columns of one force row)
  double* bs = force + ts.fOff + real * ts.cols;
  double* bt = force + tt.fOff + real * tt.cols;
  const size_t lds = ts.colsPad, ldt = tt.colsPad;
  for (int e = threadIdx.x; e < owned * mCount; e += blockDim.x) {
    const int mi = e % mCount, gl = e / mCount;
    const int gn = gnFirst + gl;
    const int m = mc + mi;
    int elems[2], locals[2], count = 0;
    const int ea = gn / nm, ia = gn % nm;
    if (ia == 0 && ea - 1 >= eStart && ea - 1 < e1) {
      elems[count] = ea - 1;
      locals[count++] = nm;
    }
    if (ea >= eStart && ea < e1) {
      elems[count] = ea;
      locals[count++] = ia;
    }
    // the row's dofs first: their loads in flight under the sums
    const int elemOwn = count > 0 ? elems[count - 1] : ea;
    const int iOwn = count > 0 ? locals[count - 1] : ia;
    const int at = ((elemOwn - g.elem0) * g.ngll + iOwn) * 2;
    const int dU = dofs[ts.dofOff + at], dV = dofs[ts.dofOff + at + 1];
    const int dW = hasTor ? dofs[tt.dofOff + at] : -1;
    Cplx U = C(0.0, 0.0), V = C(0.0, 0.0), W = C(0.0, 0.0);
    for (int k = 0; k < count; ++k) {
      const int elem = elems[k], i = locals[k];
      for (int q = 0; q < g.np; ++q) {
        const int p = (elem - eStart) * g.np + q;
        const double r = sRadius[p];
        const double point = sPoint[p];  // jac weight[q] r
        const double v = point * value[q * g.ngll + i];
        const double deriv = sWeight[q] * derivTable[q * g.ngll + i] * r;
        const Cplx* rw = rowsL[mi * maxPoints + p];
        U = U + v * rw[0] + deriv * rw[3];
        V = V + v * rw[1] + deriv * rw[4];
        W = W + v * rw[2] + deriv * rw[5];
      }
    }
    if (dU >= 0) {
      bs[static_cast<size_t>(dU) * lds + m] = U.re;
      bs[static_cast<size_t>(dU) * lds + l + 1 + m] = U.im;
    }
    if (dV >= 0) {
      bs[static_cast<size_t>(dV) * lds + m] = V.re;
      bs[static_cast<size_t>(dV) * lds + l + 1 + m] = V.im;
    }
    if (hasTor) {
      if (dW >= 0) {
        bt[static_cast<size_t>(dW) * ldt + m] = W.re;
        bt[static_cast<size_t>(dW) * ldt + l + 1 + m] = W.im;
      }
    }
  }
}

// One workgroup per (degree, order), a thread per memory node (beyond the
// workgroup's 1024 threads, the nodes 1024 apart in turn).
__global__ void __launch_bounds__(1024)
    StrainKernel(const DegreeTask* sph, const DegreeTask* tor, int lmax,
                 RadialGeometry g, const int* dofs, const double* response,
                 const double* units, const double* amp, const double* tideUnits,
                 const double* tideAmp, SpectralView out, const int* live) {
  const int pair = blockIdx.x;
  int l = static_cast<int>((sqrt(8.0 * pair + 1.0) - 1.0) * 0.5);
  while (l * (l + 1) / 2 > pair) --l;
  while ((l + 1) * (l + 2) / 2 <= pair) ++l;
  const int m = pair - l * (l + 1) / 2;
  const DegreeTask ts = sph[l];
  const int real = blockIdx.y;  // the realisation
  if (live && live[real] == 0) return;  // converged
  const int width = g.nodes * g.realisations;
  for (int node = threadIdx.x; node < g.nodes; node += blockDim.x) {
    Cplx A[5] = {C(0, 0), C(0, 0), C(0, 0), C(0, 0), C(0, 0)};
    if (node >= ts.firstNode) {
      const int elem = g.elem0 + node / g.np, q = node % g.np;
      const double r = g.radius[node];
      const double jac = g.jac[elem - g.elem0];
      Cplx U = C(0, 0), V = C(0, 0), W = C(0, 0), rdU = C(0, 0), rdV = C(0, 0),
           rdW = C(0, 0);
      const Cplx a = amp ? C(amp[2 * pair], amp[2 * pair + 1]) : C(0.0, 0.0);
      const Cplx at = tideAmp ? C(tideAmp[2 * pair], tideAmp[2 * pair + 1]) : C(0.0, 0.0);
      const bool tide = tideAmp != nullptr;
      const bool tor1 = l >= 1;
      DegreeTask tt;
      if (tor1) tt = tor[l - 1];
      for (int i = 0; i < g.ngll; ++i) {
        const int* d = dofs + ts.dofOff + ((elem - g.elem0) * g.ngll + i) * 2;
        Cplx Ui = C(0, 0), Vi = C(0, 0), Wi = C(0, 0);
        // the response columns m (Re) and l + 1 + m (Im), column-major
        const double* re = response + ts.fOff + static_cast<size_t>(real * ts.cols + m) * ts.nmPad;
        const double* im =
            response + ts.fOff + static_cast<size_t>(real * ts.cols + l + 1 + m) * ts.nmPad;
        if (d[0] >= 0) {
          Ui = C(re[d[0]], im[d[0]]);
          if (units != nullptr && ts.unitOff >= 0) Ui = Ui + units[ts.dmOff + d[0]] * a;
          if (tide) Ui = Ui + tideUnits[ts.dmOff + d[0]] * at;
        }
        if (d[1] >= 0) {
          Vi = C(re[d[1]], im[d[1]]);
          if (units != nullptr && ts.unitOff >= 0) Vi = Vi + units[ts.dmOff + d[1]] * a;
          if (tide) Vi = Vi + tideUnits[ts.dmOff + d[1]] * at;
        }
        if (tor1) {
          const int dw = dofs[tt.dofOff + ((elem - g.elem0) * g.ngll + i) * 2];
          if (dw >= 0) {
            const double* bt = response + tt.fOff;
            Wi = C(bt[static_cast<size_t>(real * tt.cols + m) * tt.nmPad + dw],
                   bt[static_cast<size_t>(real * tt.cols + l + 1 + m) * tt.nmPad + dw]);
          }
        }
        const double v = g.value[q * g.ngll + i];
        U = U + v * Ui;
        V = V + v * Vi;
        W = W + v * Wi;
        const double bb = r * g.deriv[q * g.ngll + i] / jac;
        rdU = rdU + bb * Ui;
        rdV = rdV + bb * Vi;
        rdW = rdW + bb * Wi;
      }
      // DeviatoricComponents
      const double zeta2 = l * (l + 1.0);
      const double zeta = sqrt(zeta2);
      A[2] = Div(rdU - U + (0.5 * zeta2) * V, 3.0 * r);
      if (l >= 1) {
        const Cplx tangential = rdV - V + U;
        const Cplx twist = W - rdW;
        const double den = 2.0 * kSqrt2 * r;
        A[1] = Div(zeta * (tangential + TimesI(twist)), den);
        A[3] = Div(zeta * (tangential - TimesI(twist)), den);
      }
      if (l >= 2) {
        const double f = 0.5 * zeta * sqrt(zeta2 - 2.0) / r;
        A[0] = f * (V - TimesI(W));
        A[4] = f * (V + TimesI(W));
      }
    }
    // Scatter (src/GridCoupling.cpp:595-611), every slot written once.
    const double sign = (m & 1) ? -1.0 : 1.0;
    const int col = real * g.nodes + node;
    if (l >= 2) {
      Store(out, 0, l, m, col, width, A[0]);
      if (m > 0) Store(out, 0, l, -m, col, width, sign * Conj(A[4]));
    }
    if (l >= 1) {
      Store(out, 1, l, m, col, width, A[1]);
      if (m > 0) Store(out, 1, l, -m, col, width, sign * Conj(A[3]));
    }
    Store(out, 2, l, m, col, width, A[2]);
  }
}

// StrainKernel with the workgroup's operands staged in LDS first: the
// response columns m and l + 1 + m of the degree's spheroidal and toroidal
// tasks, the unit load and tide responses, both dof tables and the GLL
// value and derivative tables, each read once with consecutive threads on
// consecutive entries (the node's radius, element size and the channel
// amplitudes are loaded beside them). A thread then reads its node's GLL
// values from LDS (no dependent global loads) and computes
// exactly what StrainKernel computes, in the same order. Sizes: capS >=
// every spheroidal task's memory dofs (at most elements x ngll x 2, each
// memory dof being some (element, GLL node, component) of the table), capT
// likewise for the toroidal tasks (elements x ngll).
__global__ void __launch_bounds__(1024)
    StrainStagedKernel(const DegreeTask* sph, const DegreeTask* tor, int lmax,
                       RadialGeometry g, const int* dofs, const double* response,
                       const double* units, const double* amp, const double* tideUnits,
                       const double* tideAmp, SpectralView out, int capS, int capT,
                       const int* live) {
  extern __shared__ double staged[];
  double* sRe = staged;
  double* sIm = sRe + capS;
  double* sUnit = sIm + capS;
  double* sTide = sUnit + capS;
  double* tRe = sTide + capS;
  double* tIm = tRe + capT;
  const int basis = g.np * g.ngll;
  double* sValue = tIm + capT;  // g.value, then g.deriv
  double* sDeriv = sValue + basis;
  int* sDof = reinterpret_cast<int*>(sDeriv + basis);
  const int table = g.elements * g.ngll;
  int* tDof = sDof + 2 * table;
  const int pair = blockIdx.x;
  int l = static_cast<int>((sqrt(8.0 * pair + 1.0) - 1.0) * 0.5);
  while (l * (l + 1) / 2 > pair) --l;
  while ((l + 1) * (l + 2) / 2 <= pair) ++l;
  const int m = pair - l * (l + 1) / 2;
  const DegreeTask ts = sph[l];
  const bool tor1 = l >= 1;
  DegreeTask tt;
  if (tor1) tt = tor[l - 1];
  const bool tide = tideAmp != nullptr;
  const bool unitsOn = units != nullptr && ts.unitOff >= 0;  // no load channel: none
  const int real = blockIdx.y;  // the realisation
  if (live && live[real] == 0) return;  // converged
  // the thread's first node's radius and element size and the channels'
  // amplitudes, loaded before the barrier
  const int first = threadIdx.x;
  double firstRadius = 0.0, firstJac = 0.0;
  if (first < g.nodes) {
    firstRadius = g.radius[first];
    firstJac = g.jac[first / g.np];
  }
  const Cplx a = amp ? C(amp[2 * pair], amp[2 * pair + 1]) : C(0.0, 0.0);
  const Cplx at = tide ? C(tideAmp[2 * pair], tideAmp[2 * pair + 1]) : C(0.0, 0.0);
  const int width = g.nodes * g.realisations;
  {
    // every operand's loads of a round issued before any of its LDS
    // stores: one memory latency per round, not one per operand and entry
    const double* re = response + ts.fOff + static_cast<size_t>(real * ts.cols + m) * ts.nmPad;
    const double* im =
        response + ts.fOff + static_cast<size_t>(real * ts.cols + l + 1 + m) * ts.nmPad;
    const double* tre =
        tor1 ? response + tt.fOff + static_cast<size_t>(real * tt.cols + m) * tt.nmPad : nullptr;
    const double* tim =
        tor1 ? response + tt.fOff + static_cast<size_t>(real * tt.cols + l + 1 + m) * tt.nmPad
             : nullptr;
    const int nS = ts.nm, nD = 2 * table, nT = tor1 ? tt.nm : 0, nTD = tor1 ? table : 0;
    const int most = max(max(max(nS, nD), max(nT, nTD)), basis);
    constexpr int kRound = 3;  // entries per thread and operand in one round
    for (int base = 0; base < most; base += kRound * static_cast<int>(blockDim.x)) {
      double vRe[kRound], vIm[kRound], vUnit[kRound], vTide[kRound], vTRe[kRound], vTIm[kRound];
      double vValue[kRound], vDeriv[kRound];
      int vDof[kRound], vTDof[kRound];
#pragma unroll
      for (int u = 0; u < kRound; ++u) {
        const int j = base + u * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
        if (j < nS) {
          vRe[u] = re[j];
          vIm[u] = im[j];
          if (unitsOn) vUnit[u] = units[ts.dmOff + j];
          if (tide) vTide[u] = tideUnits[ts.dmOff + j];
        }
        if (j < nD) vDof[u] = dofs[ts.dofOff + j];
        if (j < nT) {
          vTRe[u] = tre[j];
          vTIm[u] = tim[j];
        }
        if (j < nTD) vTDof[u] = dofs[tt.dofOff + 2 * j];
        if (j < basis) {
          vValue[u] = g.value[j];
          vDeriv[u] = g.deriv[j];
        }
      }
#pragma unroll
      for (int u = 0; u < kRound; ++u) {
        const int j = base + u * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
        if (j < nS) {
          sRe[j] = vRe[u];
          sIm[j] = vIm[u];
          if (unitsOn) sUnit[j] = vUnit[u];
          if (tide) sTide[j] = vTide[u];
        }
        if (j < nD) sDof[j] = vDof[u];
        if (j < nT) {
          tRe[j] = vTRe[u];
          tIm[j] = vTIm[u];
        }
        if (j < nTD) tDof[j] = vTDof[u];
        if (j < basis) {
          sValue[j] = vValue[u];
          sDeriv[j] = vDeriv[u];
        }
      }
    }
  }
  __syncthreads();
  for (int node = threadIdx.x; node < g.nodes; node += blockDim.x) {
    Cplx A[5] = {C(0, 0), C(0, 0), C(0, 0), C(0, 0), C(0, 0)};
    if (node >= ts.firstNode) {
      const int elem = g.elem0 + node / g.np, q = node % g.np;
      const double r = node == first ? firstRadius : g.radius[node];
      const double jac = node == first ? firstJac : g.jac[elem - g.elem0];
      Cplx U = C(0, 0), V = C(0, 0), W = C(0, 0), rdU = C(0, 0), rdV = C(0, 0),
           rdW = C(0, 0);
      for (int i = 0; i < g.ngll; ++i) {
        const int slot = (elem - g.elem0) * g.ngll + i;
        const int d0 = sDof[2 * slot], d1 = sDof[2 * slot + 1];
        Cplx Ui = C(0, 0), Vi = C(0, 0), Wi = C(0, 0);
        if (d0 >= 0) {
          Ui = C(sRe[d0], sIm[d0]);
          if (unitsOn) Ui = Ui + sUnit[d0] * a;
          if (tide) Ui = Ui + sTide[d0] * at;
        }
        if (d1 >= 0) {
          Vi = C(sRe[d1], sIm[d1]);
          if (unitsOn) Vi = Vi + sUnit[d1] * a;
          if (tide) Vi = Vi + sTide[d1] * at;
        }
        if (tor1) {
          const int dw = tDof[slot];
          if (dw >= 0) Wi = C(tRe[dw], tIm[dw]);
        }
        const double v = sValue[q * g.ngll + i];
        U = U + v * Ui;
        V = V + v * Vi;
        W = W + v * Wi;
        const double bb = r * sDeriv[q * g.ngll + i] / jac;
        rdU = rdU + bb * Ui;
        rdV = rdV + bb * Vi;
        rdW = rdW + bb * Wi;
      }
      const double zeta2 = l * (l + 1.0);
      const double zeta = sqrt(zeta2);
      A[2] = Div(rdU - U + (0.5 * zeta2) * V, 3.0 * r);
      if (l >= 1) {
        const Cplx tangential = rdV - V + U;
        const Cplx twist = W - rdW;
        const double den = 2.0 * kSqrt2 * r;
        A[1] = Div(zeta * (tangential + TimesI(twist)), den);
        A[3] = Div(zeta * (tangential - TimesI(twist)), den);
      }
      if (l >= 2) {
        const double f = 0.5 * zeta * sqrt(zeta2 - 2.0) / r;
        A[0] = f * (V - TimesI(W));
        A[4] = f * (V + TimesI(W));
      }
    }
    const double sign = (m & 1) ? -1.0 : 1.0;
    const int col = real * g.nodes + node;
    if (l >= 2) {
      Store(out, 0, l, m, col, width, A[0]);
      if (m > 0) Store(out, 0, l, -m, col, width, sign * Conj(A[4]));
    }
    if (l >= 1) {
      Store(out, 1, l, m, col, width, A[1]);
      if (m > 0) Store(out, 1, l, -m, col, width, sign * Conj(A[3]));
    }
    Store(out, 2, l, m, col, width, A[2]);
  }
}

__global__ void FillKernel(double* dst, size_t count, double value) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < count; i += stride) {
    dst[i] = value;
  }
}

// StageState (src/Viscoelastic3D.cpp:463-492) for two terms: the sum
// a x + b y accumulated from zero in the terms' order.
__global__ void CombineKernel(double* dst, double a, const double* src, double b,
                              size_t count) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += stride) {
    const double first = a * dst[i];
    const double second = b * src[i];
    dst[i] = first + second;
  }
}

// GridCoupling::ScaleByNode: complex value i of a [row][node] tensor
// belongs to node i mod nodes.
__global__ void ScaleByNodeKernel(const double* __restrict__ mu, double* spectral, size_t count,
                                  int nodes) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += stride) {
    const double f = mu[i % nodes];
    spectral[2 * i] *= f;
    spectral[2 * i + 1] *= f;
  }
}

// GridCoupling::Multiply: the tensor's two complex and one real component
// times the weight at each point.
__global__ void MultiplyKernel(double* c0, double* c1, double* r, const double* __restrict__ w,
                               size_t points) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < points;
       i += stride) {
    const double f = w[i];
    c0[2 * i] *= f;
    c0[2 * i + 1] *= f;
    c1[2 * i] *= f;
    c1[2 * i + 1] *= f;
    r[i] *= f;
  }
}

__global__ void CopyKernel(double* dst, const double* src, size_t count) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < count; i += stride) {
    dst[i] = src[i];
  }
}

// A reduction's few results to host-mapped memory, then the ticket after a
// system-scope fence (the host polls the ticket).
__global__ void PublishKernel(const double* src, int n, HostPublish pub) {
  if (static_cast<int>(threadIdx.x) < n) pub.values[threadIdx.x] = src[threadIdx.x];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_store(pub.seq, pub.ticket, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Block b covers [n b / B, n (b + 1) / B): a fixed partition of the index
// range, so the reductions below group the same terms on every run.
__device__ inline void BlockRange(size_t n, size_t& lo, size_t& hi) {
  lo = n * blockIdx.x / gridDim.x;
  hi = n * (blockIdx.x + 1) / gridDim.x;
}

__global__ void __launch_bounds__(kThreads)
    RelativeChangeKernel(const double* next, const double* now, size_t pairs,
                         double* partials) {
  __shared__ double s1[kThreads], s2[kThreads], s3[kThreads];
  size_t lo, hi;
  BlockRange(pairs, lo, hi);
  double scale = 0.0, diff = 0.0, bad = 0.0;
  for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const double nr = next[2 * i], ni = next[2 * i + 1];
    const double dr = nr - now[2 * i], di = ni - now[2 * i + 1];
    if (!(isfinite(nr) && isfinite(ni))) bad = 1.0;
    scale = fmax(scale, nr * nr + ni * ni);
    diff = fmax(diff, dr * dr + di * di);
  }
  s1[threadIdx.x] = scale;
  s2[threadIdx.x] = diff;
  s3[threadIdx.x] = bad;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
      s3[threadIdx.x] = fmax(s3[threadIdx.x], s3[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = s1[0];
    partials[3 * blockIdx.x + 1] = s2[0];
    partials[3 * blockIdx.x + 2] = s3[0];
  }
}

// Sums (or maxima) of `width` columns of partials: one workgroup per
// column, thread t taking blocks t, t + 256, ..., then a fixed tree, so the
// grouping of the terms never changes.
__global__ void __launch_bounds__(kThreads)
    ReducePartialsKernel(const double* partials, int blocks, int width, int useMax,
                         double* out) {
  __shared__ double red[kThreads];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    const double v = partials[b * width + c];
    s = useMax ? fmax(s, v) : s + v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double a = red[threadIdx.x], b = red[threadIdx.x + w];
      red[threadIdx.x] = useMax ? fmax(a, b) : a + b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// The run of compact index c (compact[r] <= c < compact[r + 1]).
__device__ inline int FindRun(const ActiveRuns& a, long long c) {
  int lo = 0, hi = a.runs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (a.compact[mid] <= c) {
      lo = mid;
    } else {
      hi = mid;
    }
  }
  return lo;
}

// Walks the compact range with stride blockDim.x in complex pairs (every
// run has even length and offset), calling body(c, j) with c and j the
// pair's first double in the compact vector and in the device layout; the
// run advances monotonically from one binary search. The workgroup's pairs
// are a fixed slice of all of them, so reductions group the same terms on
// every run.
template <class Body>
__device__ inline void ForActivePairs(const ActiveRuns& a, Body&& body) {
  size_t lo, hi;
  BlockRange(static_cast<size_t>(a.total / 2), lo, hi);
  long long p = static_cast<long long>(lo) + threadIdx.x;
  if (p >= static_cast<long long>(hi)) return;
  int run = FindRun(a, 2 * p);
  for (; p < static_cast<long long>(hi); p += blockDim.x) {
    const long long c = 2 * p;
    while (c >= a.compact[run + 1]) ++run;
    body(c, a.full[run] + (c - a.compact[run]));
  }
}

__device__ inline double2 Load2(const double* v, long long i) {
  return *reinterpret_cast<const double2*>(v + i);
}
__device__ inline void Store2(double* v, long long i, double2 x) {
  *reinterpret_cast<double2*>(v + i) = x;
}

__device__ inline void FirstBody(const double* x, const double* gx, double* xPrev,
                                 double* fPrev, const ActiveRuns& a) {
  ForActivePairs(a, [&](long long c, long long j) {
    const double2 xv = Load2(x, j), gv = Load2(gx, j);
    Store2(xPrev, c, xv);
    Store2(fPrev, c, make_double2(gv.x - xv.x, gv.y - xv.y));
  });
}

__global__ void AndersonFirstKernel(const double* x, const double* gx, double* xPrev,
                                    double* fPrev, ActiveRuns a) {
  FirstBody(x, gx, xPrev, fPrev, a);
}

__global__ void CopyRunsKernel(double* dst, const double* src, ActiveRuns a) {
  ForActivePairs(a, [&](long long, long long j) { Store2(dst, j, Load2(src, j)); });
}

// Entry e of a batch: realisation r's runs and Anderson history.
__device__ inline ActiveRuns EntryRuns(const ActiveRuns& a, int r) {
  ActiveRuns v = a;
  v.full += static_cast<size_t>(r) * a.runs;
  return v;
}
// An entry's history within the base sweep's buffers: dF and dG moved by
// `history`, xPrev and fPrev by `prev` doubles, its own m and slot (the
// base stays in the kernel arguments: indexing a modified copy of its
// pointer arrays would put the copy in scratch memory).
struct EntryHistory {
  long long history = 0, prev = 0;
  int m = 0, slot = 0;
};
__device__ inline EntryHistory EntryOf(const RealisationBatch& b, int e) {
  EntryHistory h;
  const long long r = b.r[e];
  h.history = r * b.historyStride;
  h.prev = r * b.prevStride;
  h.m = b.m[e];
  h.slot = b.slot[e];
  return h;
}

// RESIDUAL: the pass's RelativeChange (max |gx|^2, max |gx - x|^2, a
// non-finite flag) from the same loads, into residualPartials (3 per block).
// W: the sums' width, 8 or 16 (AndersonWidth).
template <bool RESIDUAL, int W>
__device__ inline void SweepBody(const AndersonSweep& s, const EntryHistory& h,
                                 const ActiveRuns& a, double* partials,
                                 double* residualPartials) {
  __shared__ double red[kThreads][2 * W + 1];
  double col[W] = {}, gam[W] = {};
  double scale = 0.0, diff = 0.0, bad = 0.0;
  double* dFs = s.dF[h.slot] + h.history;
  double* dGs = s.dG[h.slot] + h.history;
  double* xPrev = s.xPrev + h.prev;
  double* fPrev = s.fPrev + h.prev;
  ForActivePairs(a, [&](long long c, long long j) {
    const double2 xs = Load2(s.x, j), gx = Load2(s.gx, j);
    const double2 xp = Load2(xPrev, c), fp = Load2(fPrev, c);
    const double2 f = make_double2(gx.x - xs.x, gx.y - xs.y);
    const double2 df = make_double2(f.x - fp.x, f.y - fp.y);
    if constexpr (RESIDUAL) {
      if (!(isfinite(gx.x) && isfinite(gx.y))) bad = 1.0;
      scale = fmax(scale, gx.x * gx.x + gx.y * gx.y);
      diff = fmax(diff, f.x * f.x + f.y * f.y);
    }
    Store2(dFs, c, df);
    Store2(dGs, c, make_double2(xs.x - xp.x + df.x, xs.y - xp.y + df.y));
    Store2(xPrev, c, xs);
    Store2(fPrev, c, f);
#pragma unroll
    for (int i = 0; i < W; ++i) {
      if (i < h.m) {
        const double2 di = i == h.slot ? df : Load2(s.dF[i] + h.history, c);
        col[i] += di.x * df.x + di.y * df.y;
        gam[i] += di.x * f.x + di.y * f.y;
      }
    }
  });
  for (int i = 0; i < W; ++i) {
    red[threadIdx.x][i] = col[i];
    red[threadIdx.x][W + i] = gam[i];
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      for (int i = 0; i < 2 * W; ++i) red[threadIdx.x][i] += red[threadIdx.x + w][i];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2 * W) partials[blockIdx.x * 2 * W + threadIdx.x] = red[0][threadIdx.x];
  if constexpr (RESIDUAL) {
    __syncthreads();
    red[threadIdx.x][0] = scale;
    red[threadIdx.x][1] = diff;
    red[threadIdx.x][2] = bad;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        for (int i = 0; i < 3; ++i) {
          red[threadIdx.x][i] = fmax(red[threadIdx.x][i], red[threadIdx.x + w][i]);
        }
      }
      __syncthreads();
    }
    if (threadIdx.x < 3) residualPartials[blockIdx.x * 3 + threadIdx.x] = red[0][threadIdx.x];
  }
}

template <bool RESIDUAL, int W>
__global__ void __launch_bounds__(kThreads)
    AndersonSweepKernel(AndersonSweep s, ActiveRuns a, double* partials,
                        double* residualPartials) {
  EntryHistory h;
  h.m = s.m;
  h.slot = s.slot;
  SweepBody<RESIDUAL, W>(s, h, a, partials, residualPartials);
}

// Every entry's sweep (blockIdx.y the entry), its partials at e entryStride:
// the sums' partialBlocks x 2 W, then (RESIDUAL) the maxima's x 3.
template <bool RESIDUAL, int W>
__global__ void __launch_bounds__(kThreads)
    AndersonSweepBatchKernel(AndersonSweep base, ActiveRuns a, RealisationBatch b,
                             double* partials, size_t entryStride) {
  const int e = blockIdx.y;
  double* p = partials + e * entryStride;
  SweepBody<RESIDUAL, W>(base, EntryOf(b, e), EntryRuns(a, b.r[e]), p,
                         p + static_cast<size_t>(gridDim.x) * 2 * W);
}

// RelativeChange over the active entries (the others are zero in both).
__device__ inline void RelativeChangeBody(const double* next, const double* now,
                                          const ActiveRuns& a, double* partials) {
  __shared__ double s1[kThreads], s2[kThreads], s3[kThreads];
  double scale = 0.0, diff = 0.0, bad = 0.0;
  ForActivePairs(a, [&](long long, long long j) {
    const double2 n = Load2(next, j), o = Load2(now, j);
    const double dr = n.x - o.x, di = n.y - o.y;
    if (!(isfinite(n.x) && isfinite(n.y))) bad = 1.0;
    scale = fmax(scale, n.x * n.x + n.y * n.y);
    diff = fmax(diff, dr * dr + di * di);
  });
  s1[threadIdx.x] = scale;
  s2[threadIdx.x] = diff;
  s3[threadIdx.x] = bad;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
      s3[threadIdx.x] = fmax(s3[threadIdx.x], s3[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = s1[0];
    partials[3 * blockIdx.x + 1] = s2[0];
    partials[3 * blockIdx.x + 2] = s3[0];
  }
}

__global__ void __launch_bounds__(kThreads)
    RelativeChangeActiveKernel(const double* next, const double* now, ActiveRuns a,
                               double* partials) {
  RelativeChangeBody(next, now, a, partials);
}

// Every entry's RelativeChange (over its runs, or every[e]: over every entry
// of its columns), partials at e entryStride.
__global__ void __launch_bounds__(kThreads)
    RelativeChangeBatchKernel(const double* next, const double* now, ActiveRuns runs,
                              ActiveRuns every, RealisationBatch b, double* partials,
                              size_t entryStride) {
  const int e = blockIdx.y;
  RelativeChangeBody(next, now, EntryRuns(b.every[e] ? every : runs, b.r[e]),
                     partials + e * entryStride);
}

struct Gamma {
  double v[kAndersonMax];
};

// gamma: the entry's m values (held in registers before the sweep).
__device__ inline void MixBody(const AndersonSweep& s, const EntryHistory& h,
                               const double* gamma, const ActiveRuns& a) {
  double g[kAndersonMax];
#pragma unroll
  for (int i = 0; i < kAndersonMax; ++i) g[i] = i < h.m ? gamma[i] : 0.0;
  ForActivePairs(a, [&](long long c, long long j) {
    double2 value = Load2(s.gx, j);
#pragma unroll
    for (int i = 0; i < kAndersonMax; ++i) {
      if (i < h.m) {
        const double2 d = Load2(s.dG[i] + h.history, c);
        value.x -= g[i] * d.x;
        value.y -= g[i] * d.y;
      }
    }
    Store2(s.x, j, value);
  });
}

// One realisation's mix: the same sums in the same order as MixBody, the
// history loop over the run-time m (faster here than MixBody's unrolled one).
__global__ void AndersonMixKernel(AndersonSweep s, Gamma gamma, ActiveRuns a) {
  ForActivePairs(a, [&](long long c, long long j) {
    double2 value = Load2(s.gx, j);
    for (int i = 0; i < s.m; ++i) {
      const double2 d = Load2(s.dG[i], c);
      value.x -= gamma.v[i] * d.x;
      value.y -= gamma.v[i] * d.y;
    }
    Store2(s.x, j, value);
  });
}

// Every entry's mix, its gamma at gammas + e kAndersonMax (device).
__global__ void AndersonMixBatchKernel(AndersonSweep base, RealisationBatch b,
                                       const double* gammas, ActiveRuns a) {
  const int e = blockIdx.y;
  MixBody(base, EntryOf(b, e), gammas + e * kAndersonMax, EntryRuns(a, b.r[e]));
}

// Every entry's first mixed pass (the history of base moved to its
// realisation); x = gx follows in a launch of its own (the two partition the
// entries differently: one launch could overwrite x before another
// workgroup reads it).
__global__ void AndersonFirstBatchKernel(AndersonSweep base, RealisationBatch b, ActiveRuns a) {
  const int e = blockIdx.y;
  const EntryHistory h = EntryOf(b, e);
  FirstBody(base.x, base.gx, base.xPrev + h.prev, base.fPrev + h.prev, EntryRuns(a, b.r[e]));
}

__global__ void CopyRunsBatchKernel(double* dst, const double* src, ActiveRuns every,
                                    RealisationBatch b) {
  const int e = blockIdx.y;
  ForActivePairs(EntryRuns(every, b.r[e]),
                 [&](long long, long long j) { Store2(dst, j, Load2(src, j)); });
}

// ReducePartialsKernel for every entry (blockIdx.y): column c of entry e's
// partials (at e entryStride) into out[r stride + offset + c].
__global__ void __launch_bounds__(kThreads)
    ReducePartialsBatchKernel(const double* partials, size_t entryStride, int blocks, int width,
                              int useMax, RealisationBatch b, double* out, int stride,
                              int offset) {
  __shared__ double red[kThreads];
  const int c = blockIdx.x, e = blockIdx.y;
  const double* p = partials + e * entryStride;
  double s = 0.0;
  for (int k = threadIdx.x; k < blocks; k += blockDim.x) {
    const double v = p[k * width + c];
    s = useMax ? fmax(s, v) : s + v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double x = red[threadIdx.x], y = red[threadIdx.x + w];
      red[threadIdx.x] = useMax ? fmax(x, y) : x + y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b.r[e] * stride + offset + c] = red[0];
}

// Every entry's n results (result + r stride) to the host-mapped values +
// r valuesStride, then the ticket to seq[r].
__global__ void PublishBatchKernel(const double* result, int stride, int n, RealisationBatch b,
                                   HostPublish pub, int valuesStride) {
  const int r = b.r[blockIdx.x];
  if (static_cast<int>(threadIdx.x) < n) {
    pub.values[r * valuesStride + threadIdx.x] = result[r * stride + threadIdx.x];
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_store(pub.seq + r, pub.ticket, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__device__ inline void RefIndexToLm(int spin, long long idx, int& l, int& m) {
  if (spin == 0) {
    l = static_cast<int>((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while (static_cast<long long>(l) * (l + 1) / 2 > idx) --l;
    while (static_cast<long long>(l + 1) * (l + 2) / 2 <= idx) ++l;
    m = static_cast<int>(idx - static_cast<long long>(l) * (l + 1) / 2);
  } else {
    const long long t = idx + spin * spin;
    l = static_cast<int>(sqrt(static_cast<double>(t)));
    while (static_cast<long long>(l) * l > t) --l;
    while (static_cast<long long>(l + 1) * (l + 1) <= t) ++l;
    m = static_cast<int>(t - static_cast<long long>(l) * l - l);
  }
}

__global__ void LayoutKernel(const double* src, double* dst, const int* rowOffset,
                             int spin, int lmax, int nodes, long long size,
                             int toDevice) {
  const long long total = size * nodes;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  const int mlo = spin == 0 ? 0 : -lmax;
  for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
       e < total; e += stride) {
    const long long idx = e % size;
    const int node = static_cast<int>(e / size);
    int l, m;
    RefIndexToLm(spin, idx, l, m);
    const long long row = rowOffset[m - mlo] + l - LMin(m, spin);
    const long long devAt = 2 * (row * nodes + node);
    const long long refAt = 2 * (static_cast<long long>(node) * size + idx);
    if (toDevice) {
      dst[devAt] = src[refAt];
      dst[devAt + 1] = src[refAt + 1];
    } else {
      dst[refAt] = src[devAt];
      dst[refAt + 1] = src[devAt + 1];
    }
  }
}

int GridFor(size_t n) {
  const size_t blocks = (n + kThreads - 1) / kThreads;
  return static_cast<int>(blocks < 65536 ? (blocks == 0 ? 1 : blocks) : 65536);
}

}  // namespace

void LaunchPublish(const double* src, int n, const HostPublish& pub, hipStream_t stream) {
  if (n < 0 || n > 64) throw ArgumentError("publish at most 64 doubles");
  hipLaunchKernelGGL(PublishKernel, dim3(1), dim3(64), 0, stream, src, n, pub);
  SL3_HIP(hipGetLastError());
}

void LaunchFill(double* dst, size_t count, double value, hipStream_t stream) {
  if (count == 0) return;
  ProfileScope prof475("fill", stream);
  hipLaunchKernelGGL(FillKernel, dim3(GridFor(count)), dim3(kThreads), 0, stream,
                     dst, count, value);
  SL3_HIP(hipGetLastError());
}

void LaunchCombine(double* dst, double a, const double* src, double b, size_t count,
                   hipStream_t stream) {
  if (count == 0) return;
  ProfileScope prof("combine", stream);
  hipLaunchKernelGGL(CombineKernel, dim3(GridFor(count)), dim3(kThreads), 0, stream, dst, a,
                     src, b, count);
  SL3_HIP(hipGetLastError());
}

void LaunchScaleByNode(const double* mu, double* spectral, size_t count, int nodes,
                       hipStream_t stream) {
  if (count == 0) return;
  if (nodes < 1 || count % static_cast<size_t>(nodes) != 0) {
    throw ArgumentError("a [row][node] tensor holds whole rows of nodes");
  }
  ProfileScope prof("scale_by_node", stream);
  hipLaunchKernelGGL(ScaleByNodeKernel, dim3(GridFor(count)), dim3(kThreads), 0, stream, mu,
                     spectral, count, nodes);
  SL3_HIP(hipGetLastError());
}

void LaunchMultiply(double* c0, double* c1, double* r, const double* weight, size_t points,
                    hipStream_t stream) {
  if (points == 0) return;
  ProfileScope prof("multiply", stream);
  hipLaunchKernelGGL(MultiplyKernel, dim3(GridFor(points)), dim3(kThreads), 0, stream, c0, c1,
                     r, weight, points);
  SL3_HIP(hipGetLastError());
}

void LaunchCopy(double* dst, const double* src, size_t count, hipStream_t stream) {
  if (count == 0) return;
  ProfileScope prof483("copy", stream);
  hipLaunchKernelGGL(CopyKernel, dim3(GridFor(count)), dim3(kThreads), 0, stream,
                     dst, src, count);
  SL3_HIP(hipGetLastError());
}

std::vector<int2> ForceWork(int lmax) {
  std::vector<int2> work;
  for (int l = 0; l <= lmax; ++l) {
    for (int m = 0; m <= l; m += kChunkOrders) work.push_back(make_int2(l, m));
  }
  return work;
}

void LaunchForce(const DegreeTask* tasks_dev, const RadialGeometry& g,
                 const SpectralView& weighted, const int* dofs, const int2* work_dev,
                 int workCount, double* force, hipStream_t stream, const int* live) {
  if (workCount == 0 || g.elements == 0) return;
  if (g.np > kMaxPoints) throw ArgumentError("at most 6 memory points per element");
  const dim3 grid(CeilDiv(g.elements, kChunkElements), workCount, g.realisations);
  ProfileScope prof("force", stream);
  hipLaunchKernelGGL(ForceKernel, grid, dim3(kThreads), 0, stream, tasks_dev, g, weighted,
                     dofs, work_dev, force, live);
  SL3_HIP(hipGetLastError());
}

void LaunchStrain(const DegreeTask* sphTasks_dev, const DegreeTask* torTasks_dev,
                  int lmax, const RadialGeometry& g, const int* dofs,
                  const double* response, const double* units, const double* amp,
                  const double* tideUnits, const double* tideAmp, SpectralView out,
                  hipStream_t stream, const int* live) {
  const int pairs = (lmax + 1) * (lmax + 2) / 2;
  const int threads = std::min(1024, 64 * CeilDiv(g.nodes, 64));
  ProfileScope prof509("strain", stream);
  // every task's memory dofs fit these bounds (see StrainStagedKernel)
  const int table = g.elements * g.ngll;
  const int capS = 2 * table, capT = table;
  const size_t lds =
      sizeof(double) * (4 * static_cast<size_t>(capS) + 2 * capT + 2 * g.np * g.ngll) +
      sizeof(int) * (2 * static_cast<size_t>(table) + table);
  if (lds <= 64 * 1024) {
    hipLaunchKernelGGL(StrainStagedKernel, dim3(pairs, g.realisations), dim3(threads), lds, stream,
                       sphTasks_dev, torTasks_dev, lmax, g, dofs, response, units, amp,
                       tideUnits, tideAmp, out, capS, capT, live);
  } else {
    hipLaunchKernelGGL(StrainKernel, dim3(pairs, g.realisations), dim3(threads), 0, stream, sphTasks_dev,
                       torTasks_dev, lmax, g, dofs, response, units, amp, tideUnits, tideAmp,
                       out, live);
  }
  SL3_HIP(hipGetLastError());
}

void LaunchRelativeChange(const double* next, const double* now, size_t n,
                          double* partials, int partialBlocks, double* result,
                          hipStream_t stream, const HostPublish* pub) {
  ProfileScope prof518("residual", stream);
  hipLaunchKernelGGL(RelativeChangeKernel, dim3(partialBlocks), dim3(kThreads), 0,
                     stream, next, now, n / 2, partials);
  SL3_HIP(hipGetLastError());
  hipLaunchKernelGGL(ReducePartialsKernel, dim3(3), dim3(kThreads), 0, stream, partials,
                     partialBlocks, 3, 1, result);
  SL3_HIP(hipGetLastError());
  if (pub) LaunchPublish(result, 3, *pub, stream);
}

void LaunchRelativeChangeActive(const double* next, const double* now, const ActiveRuns& runs,
                                double* partials, int partialBlocks, double* result,
                                hipStream_t stream, const HostPublish* pub) {
  ProfileScope prof("residual", stream);
  hipLaunchKernelGGL(RelativeChangeActiveKernel, dim3(partialBlocks), dim3(kThreads), 0,
                     stream, next, now, runs, partials);
  SL3_HIP(hipGetLastError());
  hipLaunchKernelGGL(ReducePartialsKernel, dim3(3), dim3(kThreads), 0, stream, partials,
                     partialBlocks, 3, 1, result);
  SL3_HIP(hipGetLastError());
  if (pub) LaunchPublish(result, 3, *pub, stream);
}

void LaunchAndersonFirst(double* x, const double* gx, double* xPrev, double* fPrev,
                         const ActiveRuns& runs, size_t n, hipStream_t stream) {
  ProfileScope prof("anderson", stream);
  hipLaunchKernelGGL(AndersonFirstKernel, dim3(1024), dim3(kThreads), 0, stream, x, gx,
                     xPrev, fPrev, runs);
  SL3_HIP(hipGetLastError());
  hipLaunchKernelGGL(CopyKernel, dim3(GridFor(n)), dim3(kThreads), 0, stream, x, gx, n);
  SL3_HIP(hipGetLastError());
}

void LaunchAndersonFirst(double* x, const double* gx, double* xPrev, double* fPrev,
                         const ActiveRuns& runs, const ActiveRuns& every, hipStream_t stream) {
  ProfileScope prof("anderson", stream);
  hipLaunchKernelGGL(AndersonFirstKernel, dim3(1024), dim3(kThreads), 0, stream, x, gx,
                     xPrev, fPrev, runs);
  SL3_HIP(hipGetLastError());
  LaunchCopyRuns(x, gx, every, stream);
}

void LaunchCopyRuns(double* dst, const double* src, const ActiveRuns& runs, hipStream_t stream) {
  if (runs.total == 0) return;
  hipLaunchKernelGGL(CopyRunsKernel, dim3(1024), dim3(kThreads), 0, stream, dst, src, runs);
  SL3_HIP(hipGetLastError());
}

void LaunchAndersonSweep(const AndersonSweep& s, const ActiveRuns& runs, double* partials,
                         int partialBlocks, double* sums, double* residual,
                         hipStream_t stream, const HostPublish* pub) {
  if (s.m < 1 || s.m > kAndersonMax) throw ArgumentError("Anderson history of 1 to 16");
  ProfileScope prof("anderson", stream);
  const int W = AndersonWidth(s.m);
  double* residualPartials = partials + static_cast<size_t>(partialBlocks) * 2 * W;
  if (W == 8) {
    if (residual) {
      hipLaunchKernelGGL((AndersonSweepKernel<true, 8>), dim3(partialBlocks), dim3(kThreads), 0,
                         stream, s, runs, partials, residualPartials);
    } else {
      hipLaunchKernelGGL((AndersonSweepKernel<false, 8>), dim3(partialBlocks), dim3(kThreads),
                         0, stream, s, runs, partials, nullptr);
    }
  } else {
    if (residual) {
      hipLaunchKernelGGL((AndersonSweepKernel<true, kAndersonMax>), dim3(partialBlocks),
                         dim3(kThreads), 0, stream, s, runs, partials, residualPartials);
    } else {
      hipLaunchKernelGGL((AndersonSweepKernel<false, kAndersonMax>), dim3(partialBlocks),
                         dim3(kThreads), 0, stream, s, runs, partials, nullptr);
    }
  }
  SL3_HIP(hipGetLastError());
  hipLaunchKernelGGL(ReducePartialsKernel, dim3(2 * W), dim3(kThreads), 0, stream, partials,
                     partialBlocks, 2 * W, 0, sums);
  SL3_HIP(hipGetLastError());
  if (residual) {
    hipLaunchKernelGGL(ReducePartialsKernel, dim3(3), dim3(kThreads), 0, stream,
                       residualPartials, partialBlocks, 3, 1, residual);
    SL3_HIP(hipGetLastError());
  }
  // the sums then the residual's values, laid out as in the device result
  // when residual follows the sums directly
  if (pub) {
    if (residual && residual != sums + 2 * W) throw ArgumentError("published residual follows the sums");
    LaunchPublish(sums, 2 * W + (residual ? 3 : 0), *pub, stream);
  }
}

void LaunchRelativeChangeBatch(const double* next, const double* now, const ActiveRuns& runs,
                               const ActiveRuns& every, const RealisationBatch& b,
                               double* partials, int partialBlocks, double* result, int stride,
                               hipStream_t stream) {
  if (b.count == 0) return;
  if (b.count > kBatchMax) throw ArgumentError("at most 64 realisations per batch");
  ProfileScope prof("residual", stream);
  const size_t entryStride = static_cast<size_t>(partialBlocks) * (2 * kAndersonMax + 3);
  hipLaunchKernelGGL(RelativeChangeBatchKernel, dim3(partialBlocks, b.count), dim3(kThreads), 0,
                     stream, next, now, runs, every, b, partials, entryStride);
  SL3_HIP(hipGetLastError());
  hipLaunchKernelGGL(ReducePartialsBatchKernel, dim3(3, b.count), dim3(kThreads), 0, stream,
                     partials, entryStride, partialBlocks, 3, 1, b, result, stride, 0);
  SL3_HIP(hipGetLastError());
}

int LaunchAndersonSweepBatch(const AndersonSweep& base, const ActiveRuns& runs,
                             const RealisationBatch& b, bool residual, double* partials,
                             int partialBlocks, double* result, int stride, hipStream_t stream) {
  if (b.count == 0) return 0;
  if (b.count > kBatchMax) throw ArgumentError("at most 64 realisations per batch");
  int mMax = 1;
  for (int e = 0; e < b.count; ++e) {
    if (b.m[e] < 1 || b.m[e] > kAndersonMax) throw ArgumentError("Anderson history of 1 to 16");
    mMax = std::max(mMax, b.m[e]);
  }
  ProfileScope prof("anderson", stream);
  const int W = AndersonWidth(mMax);
  const size_t entryStride = static_cast<size_t>(partialBlocks) * (2 * kAndersonMax + 3);
  const dim3 grid(partialBlocks, b.count);
  if (W == 8) {
    if (residual) {
      hipLaunchKernelGGL((AndersonSweepBatchKernel<true, 8>), grid, dim3(kThreads), 0, stream,
                         base, runs, b, partials, entryStride);
    } else {
      hipLaunchKernelGGL((AndersonSweepBatchKernel<false, 8>), grid, dim3(kThreads), 0, stream,
                         base, runs, b, partials, entryStride);
    }
  } else {
    if (residual) {
      hipLaunchKernelGGL((AndersonSweepBatchKernel<true, kAndersonMax>), grid, dim3(kThreads), 0,
                         stream, base, runs, b, partials, entryStride);
    } else {
      hipLaunchKernelGGL((AndersonSweepBatchKernel<false, kAndersonMax>), grid, dim3(kThreads),
                         0, stream, base, runs, b, partials, entryStride);
    }
  }
  SL3_HIP(hipGetLastError());
  hipLaunchKernelGGL(ReducePartialsBatchKernel, dim3(2 * W, b.count), dim3(kThreads), 0, stream,
                     partials, entryStride, partialBlocks, 2 * W, 0, b, result, stride, 0);
  SL3_HIP(hipGetLastError());
  if (residual) {
    hipLaunchKernelGGL(ReducePartialsBatchKernel, dim3(3, b.count), dim3(kThreads), 0, stream,
                       partials + static_cast<size_t>(partialBlocks) * 2 * W, entryStride,
                       partialBlocks, 3, 1, b, result, stride, 2 * W);
    SL3_HIP(hipGetLastError());
  }
  return W;
}

void LaunchAndersonMixBatch(const AndersonSweep& base, const ActiveRuns& runs,
                            const RealisationBatch& b, const double* gammas,
                            hipStream_t stream) {
  if (b.count == 0) return;
  ProfileScope prof("anderson", stream);
  hipLaunchKernelGGL(AndersonMixBatchKernel, dim3(2048, b.count), dim3(kThreads), 0, stream,
                     base, b, gammas, runs);
  SL3_HIP(hipGetLastError());
}

void LaunchAndersonFirstBatch(const AndersonSweep& base, const ActiveRuns& runs,
                              const ActiveRuns& every, const RealisationBatch& b,
                              hipStream_t stream) {
  if (b.count == 0) return;
  ProfileScope prof("anderson", stream);
  hipLaunchKernelGGL(AndersonFirstBatchKernel, dim3(1024, b.count), dim3(kThreads), 0, stream,
                     base, b, runs);
  SL3_HIP(hipGetLastError());
  LaunchCopyRunsBatch(base.x, base.gx, every, b, stream);
}

void LaunchCopyRunsBatch(double* dst, const double* src, const ActiveRuns& every,
                         const RealisationBatch& b, hipStream_t stream) {
  if (b.count == 0) return;
  hipLaunchKernelGGL(CopyRunsBatchKernel, dim3(1024, b.count), dim3(kThreads), 0, stream, dst,
                     src, every, b);
  SL3_HIP(hipGetLastError());
}

void LaunchPublishBatch(const double* result, int stride, int n, const RealisationBatch& b,
                        const HostPublish& pub, int valuesStride, hipStream_t stream) {
  if (b.count == 0) return;
  if (n < 0 || n > 64) throw ArgumentError("publish at most 64 doubles");
  hipLaunchKernelGGL(PublishBatchKernel, dim3(b.count), dim3(64), 0, stream, result, stride, n, b,
                     pub, valuesStride);
  SL3_HIP(hipGetLastError());
}

void LaunchAndersonMix(const AndersonSweep& s, const double* gamma, const ActiveRuns& runs,
                       hipStream_t stream) {
  Gamma g{};
  for (int i = 0; i < s.m && i < kAndersonMax; ++i) g.v[i] = gamma[i];
  ProfileScope prof("anderson", stream);
  hipLaunchKernelGGL(AndersonMixKernel, dim3(2048), dim3(kThreads), 0, stream, s, g, runs);
  SL3_HIP(hipGetLastError());
}

void LaunchToDeviceLayout(const double* ref, double* dev, const int* rowOffset,
                          int spin, int lmax, int nodes, hipStream_t stream) {
  const long long size = spin == 0 ? static_cast<long long>(lmax + 1) * (lmax + 2) / 2
                                   : static_cast<long long>(lmax + 1) * (lmax + 1) -
                                         spin * spin;
  ProfileScope prof561("layout", stream);
  hipLaunchKernelGGL(LayoutKernel, dim3(GridFor(size * nodes)), dim3(kThreads), 0,
                     stream, ref, dev, rowOffset, spin, lmax, nodes, size, 1);
  SL3_HIP(hipGetLastError());
}

void LaunchToReferenceLayout(const double* dev, double* ref, const int* rowOffset,
                             int spin, int lmax, int nodes, hipStream_t stream) {
  const long long size = spin == 0 ? static_cast<long long>(lmax + 1) * (lmax + 2) / 2
                                   : static_cast<long long>(lmax + 1) * (lmax + 1) -
                                         spin * spin;
  ProfileScope prof572("layout", stream);
  hipLaunchKernelGGL(LayoutKernel, dim3(GridFor(size * nodes)), dim3(kThreads), 0,
                     stream, dev, ref, rowOffset, spin, lmax, nodes, size, 0);
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
