// The 3D step on the GPU (see step.hpp).
#include "step.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <limits>
#include <cstring>
#include <future>
#include <map>
#include <mutex>
#include <string>
#include <thread>

#include "kernels.hpp"

namespace sl3amd {

namespace {

constexpr double kPi = 3.14159265358979323846;

// itau(x) = itau_ref(node) exp(-delta(x)) (src/Viscoelastic3D.cpp:250-287)
__global__ void ItauKernel(const double* delta, const double* itauRef, double* itau,
                           size_t perNode, size_t total) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const double factor = delta ? exp(-delta[i]) : 1.0;
    itau[i] = itauRef[i / perNode] * factor;
  }
}

// The rows of a spectral block above a node block's degree, at the node
// block's columns: the addend's values there, or zero.
__global__ void CutRowsKernel(const int* rows, int count, double* block, const double* addend,
                              size_t rowDoubles, size_t first, int width) {
  const size_t total = static_cast<size_t>(count) * width;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const size_t at = static_cast<size_t>(rows[i / width]) * rowDoubles + first + i % width;
    block[at] = addend ? addend[at] : 0.0;
  }
}

// Converge's damping (src/Viscoelastic3D.cpp:1949-1957): n = o + d (n - o).
__global__ void DampKernel(double* next, const double* now, double d, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    next[i] = now[i] + d * (next[i] - now[i]);
  }
}

int Blocks(size_t n) {
  const size_t b = (n + 255) / 256;
  return static_cast<int>(std::min<size_t>(b == 0 ? 1 : b, 65536));
}

// dgelss with rcond = -1 on the symmetric Gram system, as the minimum-norm
// least-squares solution through a cyclic Jacobi eigen-decomposition
// (singular values of a symmetric matrix are its |eigenvalues|).
void GramLeastSquares(int m, const double* gram, const double* rhs, double* gamma) {
  double a[kAndersonMax * kAndersonMax], v[kAndersonMax * kAndersonMax];
  for (int i = 0; i < m * m; ++i) a[i] = gram[i];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) v[i * m + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += a[p * m + q] * a[p * m + q];
    if (off == 0.0) break;
    for (int p = 0; p < m; ++p) {
      for (int q = p + 1; q < m; ++q) {
        const double apq = a[p * m + q];
        if (apq == 0.0) continue;
        const double tau = (a[q * m + q] - a[p * m + p]) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = a[k * m + p], akq = a[k * m + q];
          a[k * m + p] = c * akp - s * akq;
          a[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = a[p * m + k], aqk = a[q * m + k];
          a[p * m + k] = c * apk - s * aqk;
          a[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = v[k * m + p], vkq = v[k * m + q];
          v[k * m + p] = c * vkp - s * vkq;
          v[k * m + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  double smax = 0.0;
  for (int i = 0; i < m; ++i) smax = std::max(smax, std::abs(a[i * m + i]));
  const double cut = 2.220446049250313e-16 * smax;
  for (int i = 0; i < m; ++i) gamma[i] = 0.0;
  for (int k = 0; k < m; ++k) {
    const double lam = a[k * m + k];
    if (!(std::abs(lam) > cut)) continue;
    double proj = 0.0;
    for (int i = 0; i < m; ++i) proj += v[i * m + k] * rhs[i];
    proj /= lam;
    for (int i = 0; i < m; ++i) gamma[i] += v[i * m + k] * proj;
  }
}

}  // namespace

// ---- Coupling ---------------------------------------------------------------

Coupling::Coupling(const Earth& earth, int lmax, double oversampling, double eps,
                   const std::vector<int>& transformDegrees, int realisations)
    : lmax_(lmax), grid_(lmax, oversampling) {
  hipStream_t s = Stream();
  if (realisations < 1 || realisations > kRealisationsMax) {
    throw ArgumentError("realisations must lie in [1, " + std::to_string(kRealisationsMax) + "]");
  }
  if (realisations > 1 && !transformDegrees.empty()) {
    throw ArgumentError("realisations side by side take the full degree's transforms: no "
                        "transform degrees");
  }
  const auto& mesh = *earth.mesh;
  const auto& points = earth.reference.points;
  const int np = points.count, n = mesh.NGll();
  // memory elements: solid, with an active branch (GridCoupling.cpp:35-52)
  std::vector<int> elements;
  for (int e = 0; e < mesh.ElementCount(); ++e) {
    if (!earth.material.IsFluid(e) && earth.rheology.ElementActive(mesh, e)) {
      elements.push_back(e);
    }
  }
  if (elements.empty()) throw ArgumentError("no active branch: nothing to couple");
  for (size_t i = 1; i < elements.size(); ++i) {
    if (elements[i] != elements[i - 1] + 1) {
      throw ArgumentError("the relaxing elements must be contiguous");
    }
  }
  geometry_.np = np;
  geometry_.ngll = n;
  geometry_.elem0 = elements.front();
  geometry_.elements = static_cast<int>(elements.size());
  geometry_.nodes = geometry_.elements * np;
  std::vector<double> jac;
  for (int e : elements) {
    jac.push_back(mesh.Jacobian(e));
    for (int p = 0; p < np; ++p) {
      nodeElement_.push_back(e);
      radiusHost_.push_back(points.RadiusAt(mesh, e, p));
    }
  }
  firstElement_.resize(lmax + 1);
  firstNodeHost_.resize(lmax + 1);
  for (int l = 0; l <= lmax; ++l) {
    firstElement_[l] = l > 0 ? mesh.StartElement(l, eps) : 0;
    int first = geometry_.nodes;
    for (int node = 0; node < geometry_.nodes; ++node) {
      if (nodeElement_[node] >= firstElement_[l]) {
        first = node;
        break;
      }
    }
    firstNodeHost_[l] = first;
  }
  reachHost_.assign(geometry_.nodes, 0);
  for (int node = 0; node < geometry_.nodes; ++node) {
    int reach = -1;
    for (int l = 0; l <= lmax; ++l) {
      if (firstNodeHost_[l] <= node) reach = l;
    }
    reachHost_[node] = std::max(reach, 0);
  }
  geometry_.realisations = realisations;
  const int N = geometry_.nodes, total = NodeCount();
  radius_.Upload(radiusHost_.data(), radiusHost_.size(), s);
  jac_.Upload(jac.data(), jac.size(), s);
  weight_.Upload(points.weight.data(), points.weight.size(), s);
  value_.Upload(points.value.data(), points.value.size(), s);
  deriv_.Upload(points.derivative.data(), points.derivative.size(), s);
  geometry_.radius = radius_.data();
  geometry_.jac = jac_.data();
  geometry_.weight = weight_.data();
  geometry_.value = value_.data();
  geometry_.deriv = deriv_.data();
  firstNode_.Upload(firstNodeHost_.data(), firstNodeHost_.size(), s);
  {
    std::vector<int> reach(static_cast<size_t>(total));  // every realisation's nodes
    for (int node = 0; node < total; ++node) reach[node] = reachHost_[node % N];
    reach_.Upload(reach.data(), reach.size(), s);
    SL3_HIP(hipStreamSynchronize(s));  // the host copy is freed next
  }

  // the spectral layout: the full degree's rows, every node
  spectralOffset_[0] = 0;
  for (int b = 0; b < 3; ++b) {
    const auto& lay = grid_.Layout(b);
    rowOffset_[b].Upload(lay.rowOffset.data(), lay.rowOffset.size(), s);
    spectralOffset_[b + 1] = spectralOffset_[b] + static_cast<size_t>(lay.Rows()) * total * 2;
  }
  // the node blocks (GridCoupling::TransformDegrees, src/GridCoupling.cpp:
  // 163-190): each memory element at the least listed degree, or lmax, at or
  // above the largest degree whose subset reaches it; the reach only grows
  // outwards, so each degree's nodes are one contiguous block
  std::vector<int> degrees = transformDegrees;
  for (int d : degrees) {
    if (d < 2 || d > lmax) {
      throw ArgumentError("a transform degree of " + std::to_string(d) + " is outside [2, lmax = " +
                          std::to_string(lmax) + "]");
    }
  }
  degrees.push_back(lmax);
  std::sort(degrees.begin(), degrees.end());
  degrees.erase(std::unique(degrees.begin(), degrees.end()), degrees.end());
  for (int node = 0; node < total; ++node) {
    const int d = *std::lower_bound(degrees.begin(), degrees.end(), reachHost_[node % N]);
    if (blocks_.empty() || blocks_.back()->lmax != d) {
      if (!blocks_.empty() && blocks_.back()->lmax > d) {
        throw ArgumentError("the node blocks' degrees must grow outwards");
      }
      blocks_.push_back(std::make_unique<NodeBlock>(d, oversampling));
      blocks_.back()->firstNode = node;
    }
    ++blocks_.back()->nodes;
  }
  for (auto& k : blocks_) {
    k->fieldOffset = gridValues_;
    gridValues_ += static_cast<size_t>(k->nodes) * k->grid.FieldSize();
    for (int b = 0; b < 3; ++b) {
      k->fourierOffset[b] = fourierDoubles_[b];
      fourierDoubles_[b] += static_cast<size_t>(k->grid.Layout(b).Orders()) * k->nodes *
                            FourierStride(k->grid.NTheta()) * 2;
    }
    BuildBlock(*k, s);
  }
  // the active runs of a spectral tensor: per realisation, per row in
  // memory order, the realisation's nodes at or above the row degree's first
  // memory node (every realisation's compact vector laid out alike); and
  // the runs of every entry of a realisation's columns
  auto runsOf = [&](bool active, DeviceArray<long long>& fullDev,
                    DeviceArray<long long>& compactDev, ActiveRuns& out) {
    std::vector<long long> full, compact;
    long long count = 0;
    for (int r = 0; r < realisations; ++r) {
      for (int b = 0; b < 3; ++b) {
        const auto& lay = grid_.Layout(b);
        for (int m = lay.mlo; m <= lmax; ++m) {
          for (int l = lay.LMin(m); l <= lmax; ++l) {
            const int row = lay.Row(l, m);
            const int fn = active ? std::min(firstNodeHost_[l], N) : 0;
            full.push_back(static_cast<long long>(spectralOffset_[b]) +
                           2LL * (static_cast<long long>(row) * total + r * N + fn));
            if (r == 0) {
              compact.push_back(count);
              count += 2LL * (N - fn);
            }
          }
        }
      }
    }
    compact.push_back(count);
    fullDev.Upload(full.data(), full.size(), s);
    compactDev.Upload(compact.data(), compact.size(), s);
    out.full = fullDev.data();
    out.compact = compactDev.data();
    out.runs = static_cast<int>(compact.size() - 1);
    out.total = count;
    SL3_HIP(hipStreamSynchronize(s));  // the host tables are freed next
  };
  runsOf(true, runFull_, runCompact_, runs_);
  runsOf(false, everyFull_, everyCompact_, everyRuns_);
  SL3_HIP(hipStreamSynchronize(s));
}

// A node block's tables: the Wigner tables on its colatitudes (rows in
// the full degree's layout, up to the block's degree), its orders, its
// Legendre plan, the longitude DFT's tables of its grid.
void Coupling::BuildBlock(NodeBlock& k, hipStream_t s) {
  const SurfaceGrid& g = k.grid;
  const int lmax = k.lmax;
  std::vector<int> orders;
  for (int b = 0; b < 3; ++b) {
    const auto& lay = g.Layout(b);
    const auto table = g.WignerTableIn(b, grid_.Layout(b));
    k.table[b].Upload(table.data(), table.size(), s);
    for (int m = lay.mlo; m <= lmax; ++m) {
      orders.push_back(b);
      orders.push_back(m);
    }
    k.plan.table[b] = k.table[b].data();
    k.plan.rowOffset[b] = rowOffset_[b].data();
    k.plan.mlo[b] = grid_.Layout(b).mlo;
    k.plan.spin[b] = lay.spin;
    k.plan.fourierMlo[b] = lay.mlo;
    SL3_HIP(hipStreamSynchronize(s));  // the host table is freed next
  }
  k.orders.Upload(orders.data(), orders.size(), s);
  k.plan.orders = k.orders.data();
  k.plan.orderCount = static_cast<int>(orders.size() / 2);
  k.plan.lmax = lmax;
  k.plan.nodes = k.nodes;
  k.plan.ntheta = g.NTheta();
  k.plan.fstride = FourierStride(g.NTheta());
  k.plan.nodeOffset = k.firstNode;
  k.plan.nodesTotal = NodeCount();
  for (int b = 0; b < 3; ++b) {
    std::vector<int> cut;
    const auto& lay = grid_.Layout(b);
    for (int m = lay.mlo; m <= lmax_; ++m) {
      for (int l = std::max(lay.LMin(m), lmax + 1); l <= lmax_; ++l) cut.push_back(lay.Row(l, m));
    }
    k.cutRowCount[b] = static_cast<int>(cut.size());
    if (cut.empty()) continue;
    k.cutRows[b].Upload(cut.data(), cut.size(), s);
    SL3_HIP(hipStreamSynchronize(s));  // the host list is freed next
  }

  DftTables& dft = k.dft;
  const int nphi = g.NPhi();
  dft.R = nphi % 4 == 0 ? 4 : 2;
  dft.P = nphi / dft.R;
  std::vector<double> cp(dft.P), sp(dft.P), cn(nphi), sn(nphi);
  for (int j = 0; j < dft.P; ++j) {
    cp[j] = std::cos(2.0 * kPi * j / dft.P);
    sp[j] = std::sin(2.0 * kPi * j / dft.P);
  }
  for (int j = 0; j < nphi; ++j) {
    cn[j] = std::cos(2.0 * kPi * j / nphi);
    sn[j] = std::sin(2.0 * kPi * j / nphi);
  }
  k.cosP.Upload(cp.data(), cp.size(), s);
  k.sinP.Upload(sp.data(), sp.size(), s);
  k.cosN.Upload(cn.data(), cn.size(), s);
  k.sinN.Upload(sn.data(), sn.size(), s);
  k.wtheta.Upload(g.CoLatitudeWeights().data(), g.NTheta(), s);
  dft.cosP = k.cosP.data();
  dft.sinP = k.sinP.data();
  dft.cosN = k.cosN.data();
  dft.sinN = k.sinN.data();
  // the matrix-core operand tables of the P-point transforms
  dft.jt = RingMfmaTiles(dft.P);
  if (dft.jt > 0) {
    const int P = dft.P, H = (P - 1) / 2, ks = 4 * dft.jt;
    std::vector<double> tc(static_cast<size_t>(dft.jt) * ks * 64, 0.0), ts(tc.size(), 0.0);
    for (int t = 0; t < dft.jt; ++t) {
      for (int kb = 0; kb < ks; ++kb) {
        for (int lane = 0; lane < 64; ++lane) {
          const int j = 1 + 16 * t + lane % 16, k = 1 + 4 * kb + lane / 16;
          if (j > H || k > H) continue;
          const size_t at = (static_cast<size_t>(t) * ks + kb) * 64 + lane;
          const int jk = (j * k) % P;
          tc[at] = std::cos(2.0 * kPi * jk / P);
          ts[at] = std::sin(2.0 * kPi * jk / P);
        }
      }
    }
    k.tabC.Upload(tc.data(), tc.size(), s);
    k.tabS.Upload(ts.data(), ts.size(), s);
    dft.tabC = k.tabC.data();
    dft.tabS = k.tabS.data();
    SL3_HIP(hipStreamSynchronize(s));  // the host tables are freed next
  }
  if (RingLds(nphi, dft.P) > 160 * 1024) {
    throw ArgumentError("grid too large for the ring kernel's LDS staging");
  }
  SL3_HIP(hipStreamSynchronize(s));
}

size_t Coupling::FourierDoubles(int block) const { return fourierDoubles_[block]; }

SpectralView Coupling::View(double* spectral) const {
  SpectralView v;
  for (int b = 0; b < 3; ++b) {
    v.block[b] = spectral + spectralOffset_[b];
    v.rowOffset[b] = rowOffset_[b].data();
  }
  v.lmax = lmax_;
  return v;
}

void Coupling::SynthesiseToFourier(const double* spectral, double* const fourier[3],
                                   bool truncated, hipStream_t s, const int* live) const {
  const double* blocks[3] = {spectral + spectralOffset_[0], spectral + spectralOffset_[1],
                             spectral + spectralOffset_[2]};
  for (const auto& k : blocks_) {
    double* out[3] = {fourier[0] + k->fourierOffset[0], fourier[1] + k->fourierOffset[1],
                      fourier[2] + k->fourierOffset[2]};
    LegendrePlan p = k->plan;
    p.live = live;
    p.liveNodes = geometry_.nodes;
    LaunchLegendreSynthesis(p, blocks, out, truncated ? reach_.data() : nullptr, s);
  }
}

void Coupling::AnalyseFromFourier(const double* const fourier[3], double* spectral,
                                  const double* addend, bool truncated, hipStream_t s,
                                  const int* live, bool keepInactive) const {
  double* blocks[3] = {spectral + spectralOffset_[0], spectral + spectralOffset_[1],
                       spectral + spectralOffset_[2]};
  const double* adds[3] = {};
  if (addend) {
    for (int b = 0; b < 3; ++b) adds[b] = addend + spectralOffset_[b];
  }
  for (const auto& k : blocks_) {
    const double* in[3] = {fourier[0] + k->fourierOffset[0], fourier[1] + k->fourierOffset[1],
                           fourier[2] + k->fourierOffset[2]};
    LegendrePlan p = k->plan;
    p.live = live;
    p.liveNodes = geometry_.nodes;
    p.keepInactive = truncated && keepInactive;
    LaunchLegendreAnalysis(p, in, blocks, addend ? adds : nullptr,
                           truncated ? reach_.data() : nullptr,
                           truncated ? firstNode_.data() : nullptr, s);
  }
  // a truncated analysis feeds the force, which reads only the degrees
  // that reach each node (all within its block's): the rows above the
  // block's degree are left as they are
  if (!truncated) FillCutRows(blocks, addend ? adds : nullptr, s);
}

void Coupling::FillCutRows(double* const spectral[3], const double* const addend[3],
                           hipStream_t s) const {
  for (const auto& k : blocks_) {
    for (int b = 0; b < 3; ++b) {
      if (spectral[b] == nullptr || k->cutRowCount[b] == 0) continue;
      const size_t total = static_cast<size_t>(k->cutRowCount[b]) * 2 * k->nodes;
      hipLaunchKernelGGL(CutRowsKernel, dim3(Blocks(total)), dim3(256), 0, s,
                         k->cutRows[b].data(), k->cutRowCount[b], spectral[b],
                         addend ? addend[b] : nullptr, 2 * static_cast<size_t>(NodeCount()),
                         2 * static_cast<size_t>(k->firstNode), 2 * k->nodes);
      SL3_HIP(hipGetLastError());
    }
  }
}

namespace {
int OrderStart(const SurfaceGrid& g, int block) {
  int z = 0;
  for (int b = 0; b < block; ++b) z += g.Layout(b).Orders();
  return z;
}
}  // namespace

void Coupling::SynthesiseBlockToFourier(int block, const double* spectral,
                                        double* fourier, hipStream_t s) const {
  for (const auto& k : blocks_) {
    LegendrePlan p = k->plan;
    p.orders = k->plan.orders + 2 * OrderStart(k->grid, block);
    p.orderCount = k->grid.Layout(block).Orders();
    const double* blocks[3] = {};
    double* fouriers[3] = {};
    blocks[block] = spectral;
    fouriers[block] = fourier + k->fourierOffset[block];
    LaunchLegendreSynthesis(p, blocks, fouriers, nullptr, s);
  }
}

void Coupling::AnalyseBlockFromFourier(int block, const double* fourier, double* spectral,
                                       hipStream_t s) const {
  for (const auto& k : blocks_) {
    LegendrePlan p = k->plan;
    p.orders = k->plan.orders + 2 * OrderStart(k->grid, block);
    p.orderCount = k->grid.Layout(block).Orders();
    const double* fouriers[3] = {};
    double* blocks[3] = {};
    blocks[block] = spectral;
    fouriers[block] = fourier + k->fourierOffset[block];
    LaunchLegendreAnalysis(p, fouriers, blocks, nullptr, nullptr, nullptr, s);
  }
  double* blocks[3] = {};
  blocks[block] = spectral;
  FillCutRows(blocks, nullptr, s);
}

RingLaunch Coupling::BlockLaunch(RingLaunch r, const NodeBlock& k) const {
  const SurfaceGrid& g = k.grid;
  r.tables = k.dft;
  r.lmax = k.lmax;
  r.nodes = k.nodes;
  r.ntheta = g.NTheta();
  r.fstride = FourierStride(g.NTheta());
  r.nphi = g.NPhi();
  r.wtheta = k.wtheta.data();
  const size_t points = k.fieldOffset;
  const size_t values = (r.block == 2 ? 1 : 2) * points;  // a spin block's grid component
  const size_t fourier = k.fourierOffset[r.block];
  auto at = [](auto* p, size_t offset) { return p != nullptr ? p + offset : p; };
  r.fourierIn = at(r.fourierIn, fourier);
  r.fourierOut = at(r.fourierOut, fourier);
  r.fourierOut2 = at(r.fourierOut2, fourier);
  r.gridIn = at(r.gridIn, values);
  r.gridOut = at(r.gridOut, values);
  r.memIn = at(r.memIn, values);
  r.memOut = at(r.memOut, values);
  r.weight = at(r.weight, points);
  r.itau = at(r.itau, points);
  r.betaOut = at(r.betaOut, points);
  r.weightOut = at(r.weightOut, points);
  r.beta = at(r.beta, points);
  r.muNode = at(r.muNode, static_cast<size_t>(k.firstNode));
  r.itauRefNode = at(r.itauRefNode, static_cast<size_t>(k.firstNode));
  return r;
}

void Coupling::Ring(RingLaunch r, hipStream_t s) const {
  for (const auto& k : blocks_) LaunchRing(BlockLaunch(r, *k), s);
}

void Coupling::RingPass(double* const in[3], double* const out[3], const double* weight,
                        hipStream_t s, const int* live) const {
  for (const auto& k : blocks_) {
    RingLaunch base;
    base.op = kRingPass;
    base.weight = weight;
    base.live = live;
    base.liveNodes = geometry_.nodes;
    const RingLaunch r = BlockLaunch(base, *k);
    if (RingPassFusable(r)) {
      RingBlocks blocks;
      for (int b = 0; b < 3; ++b) {
        blocks.in[b] = in[b] + k->fourierOffset[b];
        blocks.out[b] = out[b] + k->fourierOffset[b];
      }
      LaunchRingPassMfma(r, blocks, s);
      continue;
    }
    for (int b = 0; b < 3; ++b) {
      RingLaunch rb = base;
      rb.block = b;
      rb.fourierIn = in[b];
      rb.fourierOut = out[b];
      LaunchRing(BlockLaunch(rb, *k), s);
    }
  }
}

void Coupling::Synthesise(const double* spectral, double* const grid[3],
                          double* const fourier[3], hipStream_t s) const {
  SynthesiseToFourier(spectral, fourier, false, s);
  for (int b = 0; b < 3; ++b) {
    RingLaunch r;
    r.block = b;
    r.op = kRingSynthesise;
    r.fourierIn = fourier[b];
    r.gridOut = grid[b];
    Ring(r, s);
  }
}

void Coupling::Analyse(const double* const grid[3], double* spectral,
                       double* const fourier[3], hipStream_t s) const {
  for (int b = 0; b < 3; ++b) {
    RingLaunch r;
    r.block = b;
    r.op = kRingAnalyse;
    r.gridIn = grid[b];
    r.fourierOut = fourier[b];
    Ring(r, s);
  }
  AnalyseFromFourier(fourier, spectral, nullptr, false, s);
}

// ---- The tile-shape tuner --------------------------------------------------

namespace {
struct TuneCache {
  std::mutex mutex;
  std::map<std::array<long long, 4>, std::vector<double>> chosen;
  int analysisRows = 0, synthesisTheta = 0, bandPanelRows = 0;  // overrides, 0: tune
};
TuneCache& Tunes() {
  static TuneCache c;
  return c;
}
}  // namespace

void SetTuningOverride(int analysisRows, int synthesisTheta, int bandPanelRows) {
  auto valid = [](int v, const int* choices, int count) {
    return v == 0 || std::find(choices, choices + count, v) != choices + count;
  };
  if (!valid(analysisRows, kAnalysisRowsChoices, kTileChoices) ||
      !valid(synthesisTheta, kSynthesisThetaChoices, kTileChoices) ||
      !valid(bandPanelRows, kPanelRowsChoices, kPanelChoices)) {
    throw ArgumentError("tile shapes: analysis rows 16, 32, 64, 128; synthesis colatitudes "
                        "32, 64, 112, 208; band panel rows 4, 8, 16; 0 to tune");
  }
  auto& c = Tunes();
  std::lock_guard<std::mutex> lock(c.mutex);
  c.analysisRows = analysisRows;
  c.synthesisTheta = synthesisTheta;
  c.bandPanelRows = bandPanelRows;
}

int BandPanelOverride() {
  auto& c = Tunes();
  std::lock_guard<std::mutex> lock(c.mutex);
  return c.bandPanelRows;
}

// Each choice: one untimed launch, then five between two events; the
// launches write the same outputs from the same inputs, and every choice's
// outputs are the same bits (the tiles change the work per workgroup, not
// the order of any sum).
void Coupling::TuneLegendre(const double* spectral, double* const fourier[3],
                            double* spectralOut, hipStream_t s) {
  auto& cache = Tunes();
  int forcedRows, forcedTheta;
  {
    std::lock_guard<std::mutex> lock(cache.mutex);
    forcedRows = cache.analysisRows;
    forcedTheta = cache.synthesisTheta;
  }
  long long truncation = 0;
  for (size_t i = 0; i < reachHost_.size(); ++i) truncation = truncation * 31 + reachHost_[i];
  for (size_t i = 0; i < firstNodeHost_.size(); ++i) {
    truncation = truncation * 37 + firstNodeHost_[i];
  }
  const double* blocks[3] = {spectral + spectralOffset_[0], spectral + spectralOffset_[1],
                             spectral + spectralOffset_[2]};
  double* outs[3] = {spectralOut + spectralOffset_[0], spectralOut + spectralOffset_[1],
                     spectralOut + spectralOffset_[2]};
  for (auto& kp : blocks_) {
    NodeBlock& k = *kp;
    const std::array<long long, 4> key{
        lmax_ * 1000003LL + k.lmax, NodeCount() * 1000003LL + k.firstNode,
        k.nodes * 1000003LL + k.grid.NTheta(), truncation};
    std::vector<double> result;
    {
      std::lock_guard<std::mutex> lock(cache.mutex);
      auto it = cache.chosen.find(key);
      if (it != cache.chosen.end()) result = it->second;
    }
    if (result.empty()) {
      result.assign(2 + 2 * kTileChoices, 0.0);
      hipEvent_t ev[2];
      SL3_HIP(hipEventCreate(&ev[0]));
      SL3_HIP(hipEventCreate(&ev[1]));
      auto time = [&](auto&& launch) {
        launch();
        SL3_HIP(hipEventRecord(ev[0], s));
        for (int r = 0; r < 5; ++r) launch();
        SL3_HIP(hipEventRecord(ev[1], s));
        SL3_HIP(hipEventSynchronize(ev[1]));
        float ms = 0.0f;
        SL3_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        return ms / 5.0;
      };
      double* fo[3] = {fourier[0] + k.fourierOffset[0], fourier[1] + k.fourierOffset[1],
                       fourier[2] + k.fourierOffset[2]};
      const double* fin[3] = {fo[0], fo[1], fo[2]};
      double best = 0.0;  // the synthesis choices, then the analysis choices
      for (int i = 0; i < kTileChoices; ++i) {
        LegendrePlan p = k.plan;
        p.synthesisTheta = kSynthesisThetaChoices[i];
        const double ms = time([&] {
          LaunchLegendreSynthesisMfma(p, blocks, fo, reach_.data(), s);
        });
        result[2 + kTileChoices + i] = ms;
        if (i == 0 || ms < best) {
          best = ms;
          result[1] = p.synthesisTheta;
        }
      }
      for (int i = 0; i < kTileChoices; ++i) {
        LegendrePlan p = k.plan;
        p.analysisRows = kAnalysisRowsChoices[i];
        const double ms = time([&] {
          LaunchLegendreAnalysisMfma(p, fin, outs, nullptr, reach_.data(), firstNode_.data(),
                                     s);
        });
        result[2 + i] = ms;
        if (i == 0 || ms < best) {
          best = ms;
          result[0] = p.analysisRows;
        }
      }
      (void)hipEventDestroy(ev[0]);
      (void)hipEventDestroy(ev[1]);
      std::lock_guard<std::mutex> lock(cache.mutex);
      cache.chosen.emplace(key, result);
    }
    k.plan.analysisRows = forcedRows > 0 ? forcedRows : static_cast<int>(result[0]);
    k.plan.synthesisTheta = forcedTheta > 0 ? forcedTheta : static_cast<int>(result[1]);
    k.tuning = result;
    k.tuning[0] = k.plan.analysisRows;
    k.tuning[1] = k.plan.synthesisTheta;
  }
}

std::vector<double> Coupling::Tuning() const {
  std::vector<double> out;
  for (const auto& k : blocks_) out.insert(out.end(), k->tuning.begin(), k->tuning.end());
  return out;
}

// ---- Viscoelastic3D ---------------------------------------------------------

Viscoelastic3D::Viscoelastic3D(std::shared_ptr<const Earth> earthPtr, int lmax,
                               const double* delta_dev, StepOptions options)
    : earth_(std::move(earthPtr)),
      options_(options),
      coupling_(*earth_, lmax, options.oversampling, options.eps, options.transformDegrees,
                options.realisations) {
  const Earth& earth = *earth_;
  hipStream_t s = Stream();
  const auto& geo = coupling_.Geometry();
  const int np = geo.np;
  if (geo.realisations > 1 && options_.solver != kFixedPoint) {
    throw ArgumentError("realisations side by side take the fixed point");
  }
  for (int column = 0; column < coupling_.NodeCount(); ++column) {
    const int node = column % geo.nodes;  // every realisation's nodes alike
    const int e = coupling_.NodeElement()[node];
    const int p = node % np;
    muNodeHost_.push_back(earth.reference.pointMu[static_cast<size_t>(e) * np + p]);
    itauRefHost_.push_back(earth.reference.pointItau[static_cast<size_t>(e) * np + p]);
  }
  muNode_.Upload(muNodeHost_.data(), muNodeHost_.size(), s);
  itauRefNode_.Upload(itauRefHost_.data(), itauRefHost_.size(), s);
  if (geo.realisations > 1) {
    live_.Resize(geo.realisations);
    SetAllLive();
    const size_t g = static_cast<size_t>(geo.realisations) * kAndersonMax;
    SL3_HIP(hipHostMalloc(reinterpret_cast<void**>(&gammaHost_), g * sizeof(double),
                          hipHostMallocDefault));
    gammaDev_.Resize(g);
  }
  const size_t G = coupling_.GridValueCount();
  SetDelta(delta_dev);
  BuildTasks();
  weight_.Resize(G);
  beta_.Resize(G);
  const size_t S = coupling_.SpectralDoubles();
  fixed_.Resize(S);
  analysed_.Resize(S);
  // the rows no truncated analysis writes (above a node block's degree)
  // stay zero
  LaunchFill(fixed_.data(), S, 0.0, s);
  LaunchFill(analysed_.data(), S, 0.0, s);
  s_.Resize(S);
  sNext_.Resize(S);
  for (int b = 0; b < 3; ++b) {
    fourierA_[b].Resize(coupling_.FourierDoubles(b));
    fourierB_[b].Resize(coupling_.FourierDoubles(b));
  }
  force_.Resize(std::max<size_t>(fDoubles_, 1));
  LaunchFill(force_.data(), force_.size(), 0.0, s);  // the padding stays zero
  response_.Resize(std::max<size_t>(fDoubles_, 1));
  ampDoubles_ = static_cast<size_t>(lmax + 1) * (lmax + 2);
  amp_.Resize(2 * ampDoubles_);
  // per block the sweep's sums (at most 2 kAndersonMax) and the residual's 3 maxima
  // (per realisation side by side: their sweeps share a launch)
  partials_.Resize(static_cast<size_t>(geo.realisations) * 1024 * (2 * kAndersonMax + 3));
  result_.Resize(static_cast<size_t>(2 * kAndersonMax + 3) * geo.realisations);
  // the Legendre kernels' tile shapes, timed on the pass's own buffers
  LaunchFill(s_.data(), S, 0.0, s);
  {
    double* A[3] = {fourierA_[0].data(), fourierA_[1].data(), fourierA_[2].data()};
    coupling_.TuneLegendre(s_.data(), A, analysed_.data(), s);
  }
  for (auto& e : events_) SL3_HIP(hipEventCreate(&e));
  SL3_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned_),
                        static_cast<size_t>(kReadBackMax) * geo.realisations * sizeof(double),
                        hipHostMallocDefault));
  SL3_HIP(hipEventCreateWithFlags(&readEvent_, hipEventDisableTiming));
  {
    const size_t R = geo.realisations;
    const size_t bytes = R * kReadBackMax * sizeof(double) + R * sizeof(unsigned long long);
    SL3_HIP(hipHostMalloc(reinterpret_cast<void**>(&mapped_), bytes,
                          hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(mapped_, 0, bytes);
    void* dev = nullptr;
    SL3_HIP(hipHostGetDevicePointer(&dev, mapped_, 0));
    mappedValuesDev_ = static_cast<double*>(dev);
    mappedSeqDev_ = reinterpret_cast<unsigned long long*>(mappedValuesDev_ + R * kReadBackMax);
  }
  SL3_HIP(hipStreamSynchronize(s));
}

Viscoelastic3D::~Viscoelastic3D() {
  DropPassGraphs(0);
  for (auto& e : events_)
    if (e != nullptr) (void)hipEventDestroy(e);
  if (readEvent_ != nullptr) (void)hipEventDestroy(readEvent_);
  if (pinned_ != nullptr) (void)hipHostFree(pinned_);
  if (mapped_ != nullptr) (void)hipHostFree(mapped_);
  if (gammaHost_ != nullptr) (void)hipHostFree(gammaHost_);
}

// Copies `count` doubles of a device result to the host through pinned
// memory and waits by polling an event recorded behind the copy: the poll
// returns as soon as the copy lands, where a stream synchronisation may
// sleep until an interrupt wakes it (a pass reads its residual this way, and
// the device idles until the host has launched the next work).
void Viscoelastic3D::ReadBack(double* dst, const double* src, int count) {
  if (count < 0 || count > kReadBackMax * coupling_.Realisations()) {
    throw ArgumentError("read-back of at most 40 doubles per realisation");
  }
  hipStream_t s = Stream();
  SL3_HIP(hipMemcpyAsync(pinned_, src, count * sizeof(double), hipMemcpyDeviceToHost, s));
  SL3_HIP(hipEventRecord(readEvent_, s));
  hipError_t status;
  while ((status = hipEventQuery(readEvent_)) == hipErrorNotReady) {
  }
  SL3_HIP(status);
  std::copy(pinned_, pinned_ + count, dst);
}

HostPublish Viscoelastic3D::Publish(int r) {
  HostPublish p;
  p.values = mappedValuesDev_ + static_cast<size_t>(r) * kReadBackMax;
  p.seq = mappedSeqDev_ + r;
  p.ticket = ++ticket_;
  return p;
}

// Polls the ticket word; the stream's state is checked now and then, so an
// error ends the wait, and a stream that went idle without the ticket
// arriving is an error rather than a hang.
const double* Viscoelastic3D::Published(int r, unsigned long long ticket) {
  const size_t R = coupling_.Realisations();
  auto* seq = reinterpret_cast<unsigned long long*>(mapped_ + R * kReadBackMax) + r;
  for (unsigned long long spin = 0;; ++spin) {
    if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) == ticket) break;
    if ((spin & 1023) == 1023) {
      const hipError_t q = hipStreamQuery(Stream());
      if (q == hipSuccess && __atomic_load_n(seq, __ATOMIC_ACQUIRE) != ticket) {
        throw std::runtime_error("a reduction's result did not reach the host");
      }
      if (q != hipSuccess && q != hipErrorNotReady) SL3_HIP(q);
    }
  }
  return mapped_ + static_cast<size_t>(r) * kReadBackMax;
}

void Viscoelastic3D::SetDelta(const double* delta_dev) {
  hipStream_t s = Stream();
  const size_t G = coupling_.GridValueCount();
  itau_.Resize(G);
  ProfileScope prof381("itau", s);
  // each node block's nodes on the block's grid
  for (int i = 0; i < coupling_.BlockCount(); ++i) {
    const NodeBlock& k = coupling_.Block(i);
    const size_t per = k.grid.FieldSize(), count = per * k.nodes;
    hipLaunchKernelGGL(ItauKernel, dim3(Blocks(count)), dim3(256), 0, s,
                       delta_dev ? delta_dev + k.fieldOffset : nullptr,
                       itauRefNode_.data() + k.firstNode, itau_.data() + k.fieldOffset, per,
                       count);
    SL3_HIP(hipGetLastError());
  }
  SL3_HIP(hipStreamSynchronize(s));
}

void Viscoelastic3D::BuildTasks() {
  const auto& mesh = *earth_->mesh;
  const auto& geo = coupling_.Geometry();
  const int L = coupling_.LMax();
  const int n = mesh.NGll();
  // the dof maps (the elastic operators': any point shear gives the same)
  const PointShear elastic{&earth_->reference.points, &earth_->reference.elasticShear};
  std::vector<std::unique_ptr<DegreeOperator>> sph(L + 1);
  std::vector<std::unique_ptr<ToroidalOperator>> tor(L + 1);
  ParallelFor(2 * L + 1, [&](int i) {
    if (i <= L) {
      sph[i] = std::make_unique<DegreeOperator>(earth_->reference.nodal, i, options_.eps, &elastic);
    } else {
      const int l = i - L;
      tor[l] = std::make_unique<ToroidalOperator>(earth_->reference.nodal, l, options_.eps, &elastic);
    }
  });
  tasks_.clear();
  dofsHost_.clear();
  dofsMemHost_.clear();
  dmHost_.clear();
  rhsDoubles_ = scaleDoubles_ = unitDoubles_ = 0;
  gDoubles_ = fDoubles_ = 0;
  nmMax_ = nmPadMax_ = colsMax_ = 0;
  auto add = [&](int l, int kind) {
    DegreeTask t;
    t.l = l;
    t.kind = kind;
    t.cols = 2 * (l + 1);
    t.firstElement = coupling_.FirstElementOfDegree(l);
    t.firstNode = coupling_.FirstNodes()[l];
    t.ndof = kind == 0 ? sph[l]->DofCount() : tor[l]->DofCount();
    // the toroidal degree-1 system pins one dof per solid region (its
    // rigid rotation): none of them is a memory dof
    const std::vector<int> pins = kind == 1 ? tor[l]->Pinned() : std::vector<int>{};
    auto pinned = [&](int d) { return std::find(pins.begin(), pins.end(), d) != pins.end(); };
    t.pinned = pins.empty() ? -1 : pins.back();
    t.rhsOff = static_cast<long long>(rhsDoubles_);
    rhsDoubles_ += static_cast<size_t>(t.ndof) * t.cols;
    t.scaleOff = static_cast<long long>(scaleDoubles_);
    scaleDoubles_ += t.ndof;
    if (kind == 0) {
      t.unitOff = static_cast<long long>(unitDoubles_);
      unitDoubles_ += 2 * static_cast<size_t>(t.ndof);
    }
    t.dofOff = static_cast<long long>(dofsHost_.size());
    for (int e = geo.elem0; e < geo.elem0 + geo.elements; ++e) {
      for (int i = 0; i < n; ++i) {
        const int gn = mesh.GlobalIndex(e, i);
        if (e < t.firstElement) {
          dofsHost_.push_back(-1);
          dofsHost_.push_back(-1);
        } else if (kind == 0) {
          dofsHost_.push_back(sph[l]->Dof(gn, Component::U));
          dofsHost_.push_back(sph[l]->Dof(gn, Component::V));
        } else {
          dofsHost_.push_back(tor[l]->Dof(gn));
          dofsHost_.push_back(-1);
        }
      }
    }
    // the memory dofs, in dof order, and the table's memory-dof rows
    std::vector<int> dm;
    for (size_t k = static_cast<size_t>(t.dofOff); k < dofsHost_.size(); ++k) {
      const int d = dofsHost_[k];
      if (d >= 0 && !pinned(d)) dm.push_back(d);
    }
    std::sort(dm.begin(), dm.end());
    dm.erase(std::unique(dm.begin(), dm.end()), dm.end());
    for (size_t k = static_cast<size_t>(t.dofOff); k < dofsHost_.size(); ++k) {
      const int d = dofsHost_[k];
      int row = -1;
      if (d >= 0 && !pinned(d)) {
        row = static_cast<int>(std::lower_bound(dm.begin(), dm.end(), d) - dm.begin());
      }
      dofsMemHost_.push_back(row);
    }
    t.nm = static_cast<int>(dm.size());
    t.nmPad = 16 * CeilDiv(t.nm, 16);
    // every realisation's columns side by side
    t.colsPad = 16 * CeilDiv(geo.realisations * t.cols, 16);
    t.dmOff = static_cast<long long>(dmHost_.size());
    dmHost_.insert(dmHost_.end(), dm.begin(), dm.end());
    t.gOff = static_cast<long long>(gDoubles_);
    gDoubles_ += static_cast<size_t>(t.nmPad) * t.nmPad;
    t.fOff = static_cast<long long>(fDoubles_);
    fDoubles_ += static_cast<size_t>(t.nmPad) * t.colsPad;
    nmMax_ = std::max(nmMax_, t.nm);
    nmPadMax_ = std::max(nmPadMax_, t.nmPad);
    colsMax_ = std::max(colsMax_, t.cols);
    tasks_.push_back(t);
  };
  for (int l = 0; l <= L; ++l) add(l, 0);
  for (int l = 1; l <= L; ++l) add(l, 1);
  std::vector<long long> solOff;
  solutionDoubles_ = 0;
  for (const auto& t : tasks_) {
    solOff.push_back(static_cast<long long>(solutionDoubles_));
    solutionDoubles_ += 2 * static_cast<size_t>(t.ndof) * (t.l + 1);
  }
  hipStream_t s = Stream();
  solutionOffset_.Upload(solOff.data(), solOff.size(), s);
  tasksDev_.Upload(tasks_.data(), tasks_.size(), s);
  dofs_.Upload(dofsHost_.data(), dofsHost_.size(), s);
  dofsMem_.Upload(dofsMemHost_.data(), dofsMemHost_.size(), s);
  dm_.Upload(dmHost_.data(), std::max<size_t>(dmHost_.size(), 1), s);
  const auto work = ResponseWorkList(tasks_);
  work_.Upload(work.data(), work.size(), s);
  workCount_ = static_cast<int>(work.size());
  narrowCount_ = ResponseNarrowCount(work);
  const auto forceWork = ForceWork(L);
  forceWork_.Upload(forceWork.data(), forceWork.size(), s);
  forceWorkCount_ = static_cast<int>(forceWork.size());
  const auto pairWork = PairWorkList(tasks_);
  pairWork_.Upload(pairWork.data(), std::max<size_t>(pairWork.size(), 1), s);
  pairWorkCount_ = static_cast<int>(pairWork.size());
  SL3_HIP(hipStreamSynchronize(s));
}

Flavour& Viscoelastic3D::FlavourOf(int scheme, double dt) {
  if (dt == 0.0) scheme = kBackwardEuler;  // every scheme's zero-length step
  for (auto it = flavours_.begin(); it != flavours_.end(); ++it) {
    if ((*it)->scheme == scheme && (*it)->dt == dt) {
      flavours_.splice(flavours_.begin(), flavours_, it);
      return *flavours_.front();
    }
  }
  while (flavours_.size() >= 4) {
    DropPassGraphs(flavours_.back()->id);
    flavours_.pop_back();
  }
  auto f = std::make_unique<Flavour>();
  f->id = ++flavourCount_;
  f->scheme = scheme;
  f->dt = dt;
  f->pointShear = earth_->reference.PointShearModulus(scheme, dt);
  const int L = coupling_.LMax();
  const PointShear ps{&earth_->reference.points, &f->pointShear};
  std::vector<std::unique_ptr<DegreeOperator>> sph(L + 1);
  std::vector<std::unique_ptr<ToroidalOperator>> tor(L + 1);
  ParallelFor(2 * L + 1, [&](int i) {
    if (i <= L) {
      sph[i] = std::make_unique<DegreeOperator>(earth_->reference.nodal, i, options_.eps, &ps);
    } else {
      const int l = i - L;
      tor[l] = std::make_unique<ToroidalOperator>(earth_->reference.nodal, l, options_.eps, &ps);
    }
  });
  hipStream_t s = Stream();
  // The scaling, and the bands for the Cholesky attempt (spheroidal with a
  // positive diagonal, every toroidal system).
  std::vector<double> scale(scaleDoubles_), bands;
  std::vector<BandSystem> chol;
  std::vector<int> cholTask;
  for (size_t ti = 0; ti < tasks_.size(); ++ti) {
    const auto& t = tasks_[ti];
    const auto& sc = t.kind == 0 ? sph[t.l]->Scale() : tor[t.l]->Scale();
    std::copy(sc.begin(), sc.end(), scale.begin() + t.scaleOff);
    if (t.kind == 0 && !sph[t.l]->PositiveDiagonal()) continue;
    const auto band = t.kind == 0 ? sph[t.l]->ScaledUpperBand() : tor[t.l]->ScaledUpperBand();
    BandSystem b;
    b.n = t.ndof;
    b.kl = t.kind == 0 ? sph[t.l]->Kd() : tor[t.l]->Kd();
    b.nrhs = t.cols;
    b.ab_off = static_cast<long long>(bands.size());
    b.b_off = t.rhsOff;
    b.row_stride = t.cols;
    b.col_stride = 1;
    bands.insert(bands.end(), band.begin(), band.end());
    chol.push_back(b);
    cholTask.push_back(static_cast<int>(ti));
    f->kdMax = std::max(f->kdMax, b.kl);
  }
  f->usedCholesky.assign(tasks_.size(), 0);
  {
    DeviceArray<double> cholBands;
    cholBands.Upload(bands.data(), bands.size(), s);
    DeviceArray<BandSystem> sysDev;
    sysDev.Upload(chol.data(), chol.size(), s);
    DeviceArray<int> info(chol.size());
    LaunchCholeskyFactor(sysDev.data(), static_cast<int>(chol.size()), f->kdMax,
                         cholBands.data(), info.data(), s);
    std::vector<int> infoHost(chol.size());
    SL3_HIP(hipMemcpyAsync(infoHost.data(), info.data(), sizeof(int) * chol.size(),
                           hipMemcpyDeviceToHost, s));
    SL3_HIP(hipMemcpyAsync(bands.data(), cholBands.data(), sizeof(double) * bands.size(),
                           hipMemcpyDeviceToHost, s));
    SL3_HIP(hipStreamSynchronize(s));
    // keep the successes; the spheroidal failures fall back to LU
    // (src/DegreeOperator.cpp:823-851), a toroidal one is an error
    std::vector<double> factors;
    std::vector<BandSystem> keep, lu;
    std::vector<double> luBands;
    std::vector<int> luTask;
    for (size_t k = 0; k < chol.size(); ++k) {
      const auto& t = tasks_[cholTask[k]];
      if (infoHost[k] != 0) {
        if (t.kind == 1) {
          throw NumericalError("toroidal degree " + std::to_string(t.l) +
                               ": Cholesky failed (info = " + std::to_string(infoHost[k]) + ")");
        }
        continue;
      }
      BandSystem b = chol[k];
      b.ab_off = static_cast<long long>(factors.size());
      factors.insert(factors.end(), bands.begin() + chol[k].ab_off,
                     bands.begin() + chol[k].ab_off + static_cast<long long>(b.kl + 1) * b.n);
      keep.push_back(b);
      f->usedCholesky[cholTask[k]] = 1;
      f->choleskyTask.push_back(cholTask[k]);
    }
    for (size_t ti = 0; ti < tasks_.size(); ++ti) {
      const auto& t = tasks_[ti];
      if (t.kind != 0 || f->usedCholesky[ti]) continue;
      const auto band = sph[t.l]->ScaledLuBand();
      BandSystem b;
      b.n = t.ndof;
      b.kl = b.ku = sph[t.l]->Kd();
      b.nrhs = t.cols;
      b.ab_off = static_cast<long long>(luBands.size());
      b.piv_off = static_cast<long long>(luTask.size() ? 0 : 0);
      b.b_off = t.rhsOff;
      b.row_stride = t.cols;
      b.col_stride = 1;
      luBands.insert(luBands.end(), band.begin(), band.end());
      lu.push_back(b);
      luTask.push_back(static_cast<int>(ti));
    }
    long long pivTotal = 0;
    for (auto& b : lu) {
      b.piv_off = pivTotal;
      pivTotal += b.n;
    }
    const size_t luBase = factors.size();
    for (auto& b : lu) b.ab_off += static_cast<long long>(luBase);
    factors.insert(factors.end(), luBands.begin(), luBands.end());
    f->factors.Upload(factors.data(), factors.size(), s);
    f->pivots.Resize(std::max<long long>(pivTotal, 1));
    if (!lu.empty()) {
      DeviceArray<BandSystem> luDev;
      luDev.Upload(lu.data(), lu.size(), s);
      DeviceArray<int> luInfo(lu.size());
      int klMax = 0;
      for (const auto& b : lu) klMax = std::max(klMax, b.kl);
      LaunchLuFactor(luDev.data(), static_cast<int>(lu.size()), klMax, klMax,
                     f->factors.data(), f->pivots.data(), luInfo.data(), s);
      std::vector<int> li(lu.size());
      SL3_HIP(hipMemcpyAsync(li.data(), luInfo.data(), sizeof(int) * lu.size(),
                             hipMemcpyDeviceToHost, s));
      SL3_HIP(hipStreamSynchronize(s));
      for (size_t k = 0; k < lu.size(); ++k) {
        if (li[k] != 0) {
          throw NumericalError("degree " + std::to_string(tasks_[luTask[k]].l) +
                               ": the banded LU factorisation failed (info = " +
                               std::to_string(li[k]) + ")");
        }
      }
    }
    f->cholesky = keep;
    f->lu = lu;
    f->luTask = luTask;
    SL3_HIP(hipStreamSynchronize(s));  // the host staging vectors end here
  }
  f->scale.Upload(scale.data(), scale.size(), s);
  f->choleskyDev.Upload(f->cholesky.data(), f->cholesky.size(), s);
  f->luDev.Upload(f->lu.data(), f->lu.size(), s);
  const auto ct = SolveTasks(f->cholesky);
  const auto lt = SolveTasks(f->lu);
  f->choleskyTasks.Upload(ct.data(), ct.size(), s);
  f->luTasks.Upload(lt.data(), lt.size(), s);
  f->choleskyTaskCount = static_cast<int>(ct.size());
  f->luTaskCount = static_cast<int>(lt.size());

  // The unit load and tide responses (Viscoelastic3D::AddUnitChannels'
  // cache, src/Viscoelastic3D.cpp:1042-1055): Load() and Tide(2 at l = 0,
  // else l) solved together with this flavour.
  std::vector<double> unitRhs(unitDoubles_, 0.0);
  std::vector<std::vector<double>> unitForce(L + 1);  // load, then tide, per degree
  for (const auto& t : tasks_) {
    if (t.kind != 0) continue;
    const auto load = sph[t.l]->Load();
    const auto tide = sph[t.l]->Tide(t.l == 0 ? 2 : -1);
    unitForce[t.l] = load;
    unitForce[t.l].insert(unitForce[t.l].end(), tide.begin(), tide.end());
    for (int d = 0; d < t.ndof; ++d) {  // BeginSolve
      unitRhs[t.unitOff + d] = load[d] * scale[t.scaleOff + d];
      unitRhs[t.unitOff + t.ndof + d] = tide[d] * scale[t.scaleOff + d];
    }
  }
  DeviceArray<double> unitDev;
  unitDev.Upload(unitRhs.data(), unitRhs.size(), s);
  auto single = [&](std::vector<BandSystem> sys) {
    std::vector<BandSystem> out;
    for (auto b : sys) {
      int ti = -1;
      for (size_t k = 0; k < tasks_.size(); ++k) {
        if (tasks_[k].rhsOff == b.b_off) ti = static_cast<int>(k);
      }
      if (ti < 0 || tasks_[ti].kind != 0) continue;
      b.nrhs = 2;
      b.b_off = tasks_[ti].unitOff;
      b.row_stride = 1;
      b.col_stride = b.n;
      out.push_back(b);
    }
    return out;
  };
  const auto uc = single(f->cholesky), ul = single(f->lu);
  DeviceArray<BandSystem> ucDev, ulDev;
  ucDev.Upload(uc.data(), uc.size(), s);
  ulDev.Upload(ul.data(), ul.size(), s);
  const auto uct = SolveTasks(uc), ult = SolveTasks(ul);
  DeviceArray<int2> uctDev, ultDev;
  uctDev.Upload(uct.data(), uct.size(), s);
  ultDev.Upload(ult.data(), ult.size(), s);
  LaunchCholeskySolve(ucDev.data(), uctDev.data(), static_cast<int>(uct.size()),
                      f->kdMax, f->factors.data(), unitDev.data(), s);
  int kvMax = 0;
  for (const auto& b : f->lu) kvMax = std::max(kvMax, b.kl + b.ku);
  LaunchLuSolve(ulDev.data(), ultDev.data(), static_cast<int>(ult.size()), kvMax,
                f->factors.data(), f->pivots.data(), unitDev.data(), s);
  std::vector<double> units(unitDoubles_);
  SL3_HIP(hipMemcpyAsync(units.data(), unitDev.data(), sizeof(double) * units.size(),
                         hipMemcpyDeviceToHost, s));
  SL3_HIP(hipStreamSynchronize(s));
  for (const auto& t : tasks_) {
    if (t.kind != 0) continue;
    for (int d = 0; d < t.ndof; ++d) {  // FinishSolve
      units[t.unitOff + d] *= scale[t.scaleOff + d];
      units[t.unitOff + t.ndof + d] *= scale[t.scaleOff + d];
    }
  }
  f->units.Upload(units.data(), units.size(), s);
  // the channels' part of the forcing's energy <f, K^-1 f> (the conjugate
  // gradients' scale): each unit force against each unit response
  f->channelEnergy.assign(4 * static_cast<size_t>(L + 1), 0.0);
  for (const auto& t : tasks_) {
    if (t.kind != 0) continue;
    const double* u = unitForce[t.l].data();
    const double* z = units.data() + t.unitOff;
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    for (int d = 0; d < t.ndof; ++d) {
      e[0] += u[d] * z[d];
      e[1] += u[d] * z[t.ndof + d];
      e[2] += u[t.ndof + d] * z[d];
      e[3] += u[t.ndof + d] * z[t.ndof + d];
    }
    std::copy(e, e + 4, f->channelEnergy.begin() + 4 * t.l);
  }
  const size_t memDofs = dmHost_.size();
  std::vector<double> unitsMem(std::max<size_t>(2 * memDofs, 1), 0.0);
  for (const auto& t : tasks_) {
    if (t.kind != 0) continue;
    for (int i = 0; i < t.nm; ++i) {
      const int d = dmHost_[t.dmOff + i];
      unitsMem[t.dmOff + i] = units[t.unitOff + d];
      unitsMem[memDofs + t.dmOff + i] = units[t.unitOff + t.ndof + d];
    }
  }
  f->unitsMem.Upload(unitsMem.data(), unitsMem.size(), s);
  BuildResponse(*f);
  SL3_HIP(hipStreamSynchronize(s));
  flavours_.push_front(std::move(f));
  return *flavours_.front();
}

int Viscoelastic3D::CholeskyFallbacks(int scheme, double dt) {
  auto& f = FlavourOf(scheme, dt);
  int count = 0;
  for (size_t k = 0; k < tasks_.size(); ++k) count += f.usedCholesky[k] ? 0 : 1;
  return count;
}

std::unique_ptr<DeviceState> Viscoelastic3D::ZeroState() const {
  auto state = std::make_unique<DeviceState>();
  hipStream_t s = Stream();
  for (int b = 0; b < 3; ++b) {
    state->m[b].Resize(coupling_.GridDoubles(b));
    LaunchFill(state->m[b].data(), state->m[b].size(), 0.0, s);
  }
  return state;
}

// The memory response of every task: unit columns on the memory dofs,
// scaled (BeginSolve), solved with the flavour's factors, unscaled
// (FinishSolve) on the memory rows and packed for the response GEMM.
void Viscoelastic3D::BuildResponse(Flavour& f) {
  hipStream_t s = Stream();
  const int count = static_cast<int>(tasks_.size());
  std::vector<long long> zOff(tasks_.size());
  long long zTotal = 0;
  for (size_t ti = 0; ti < tasks_.size(); ++ti) {
    zOff[ti] = zTotal;
    zTotal += static_cast<long long>(tasks_[ti].ndof) * tasks_[ti].nm;
  }
  DeviceArray<double> z(static_cast<size_t>(std::max<long long>(zTotal, 1)));
  LaunchFill(z.data(), z.size(), 0.0, s);
  DeviceArray<long long> zOffDev;
  zOffDev.Upload(zOff.data(), zOff.size(), s);
  LaunchUnitColumns(tasksDev_.data(), count, nmMax_, dm_.data(), f.scale.data(),
                    zOffDev.data(), z.data(), s);
  auto columns = [&](const std::vector<BandSystem>& systems, const std::vector<int>& task) {
    std::vector<BandSystem> out;
    for (size_t k = 0; k < systems.size(); ++k) {
      const auto& t = tasks_[task[k]];
      if (t.nm == 0) continue;
      BandSystem b = systems[k];
      b.nrhs = t.nm;
      b.b_off = zOff[task[k]];
      b.row_stride = t.nm;
      b.col_stride = 1;
      out.push_back(b);
    }
    return out;
  };
  const auto zc = columns(f.cholesky, f.choleskyTask), zl = columns(f.lu, f.luTask);
  DeviceArray<BandSystem> zcDev, zlDev;
  zcDev.Upload(zc.data(), zc.size(), s);
  zlDev.Upload(zl.data(), zl.size(), s);
  const auto zct = SolveTasks(zc), zlt = SolveTasks(zl);
  DeviceArray<int2> zctDev, zltDev;
  zctDev.Upload(zct.data(), zct.size(), s);
  zltDev.Upload(zlt.data(), zlt.size(), s);
  LaunchCholeskySolve(zcDev.data(), zctDev.data(), static_cast<int>(zct.size()), f.kdMax,
                      f.factors.data(), z.data(), s);
  int kvMax = 0;
  for (const auto& b : f.lu) kvMax = std::max(kvMax, b.kl + b.ku);
  LaunchLuSolve(zlDev.data(), zltDev.data(), static_cast<int>(zlt.size()), kvMax,
                f.factors.data(), f.pivots.data(), z.data(), s);
  f.response.Resize(std::max<size_t>(gDoubles_, 1));
  LaunchPackResponse(tasksDev_.data(), count, nmPadMax_, dm_.data(), f.scale.data(),
                     zOffDev.data(), z.data(), f.response.data(), s);
  SL3_HIP(hipStreamSynchronize(s));  // z and the descriptors end here
}

void Viscoelastic3D::SolveAll(const double* weighted, Flavour& f) {
  hipStream_t s = Stream();
  rhsCurrent_ = false;
  const SpectralView w = coupling_.View(const_cast<double*>(weighted));
  LaunchForce(tasksDev_.data(), coupling_.Geometry(), w, dofsMem_.data(), forceWork_.data(),
              forceWorkCount_, force_.data(), s, Live());
  LaunchResponse(tasksDev_.data(), work_.data(), workCount_, narrowCount_, f.response.data(),
                 force_.data(), response_.data(), s, Live(), coupling_.Realisations());
}

// TideColumns' check (src/Viscoelastic3D.cpp:142-154) on degree 1's columns.
static void RejectDegreeOneTide(const double* cols) {
  for (int k = 0; k < 4; ++k) {
    if (cols[k] != 0.0) {
      throw ArgumentError(
          "a tide channel has no degree-1 part: tides are taken relative to the "
          "centre of mass, its degree-1 columns must vanish");
    }
  }
}

void Viscoelastic3D::CheckTide(const Channels& channels, double t) const {
  if (!channels.tide || coupling_.LMax() < 1) return;
  double cols[4] = {0.0, 0.0, 0.0, 0.0};
  if (channels.tide(1, t, cols)) RejectDegreeOneTide(cols);
}

// The columns of both channels at t (AddUnitChannels' amplitudes,
// src/Viscoelastic3D.cpp:1058-1071); a tide with a nonzero degree-1 column
// is rejected as TideColumns does (:142-154).
void Viscoelastic3D::Amplitudes(const Channels& channels, double t) {
  const int L = coupling_.LMax();
  std::vector<double> amp(2 * ampDoubles_, 0.0);
  std::vector<double> cols(2 * (L + 1));
  auto gather = [&](const LoadFn& fn, double* into, bool tide) {
    for (int l = 0; l <= L; ++l) {
      std::fill(cols.begin(), cols.end(), 0.0);
      if (!fn(l, t, cols.data())) continue;
      if (tide && l == 1) RejectDegreeOneTide(cols.data());
      for (int m = 0; m <= l; ++m) {
        const size_t at = static_cast<size_t>(l) * (l + 1) / 2 + m;
        into[2 * at] = cols[2 * m];
        into[2 * at + 1] = cols[2 * m + 1];
      }
    }
  };
  if (channels.load) gather(channels.load, amp.data(), false);
  hasTide_ = static_cast<bool>(channels.tide);
  if (hasTide_) gather(channels.tide, amp.data() + ampDoubles_, true);
  ampHost_ = amp;
  hipStream_t s = Stream();
  SL3_HIP(hipMemcpyAsync(amp_.data(), amp.data(), sizeof(double) * amp.size(),
                         hipMemcpyHostToDevice, s));
  SL3_HIP(hipStreamSynchronize(s));
}

double Viscoelastic3D::RelativeChange(const double* next, const double* now, bool everyEntry) {
  hipStream_t s = Stream();
  const HostPublish pub = Publish(0);
  if (everyEntry) {
    LaunchRelativeChange(next, now, coupling_.SpectralDoubles(), partials_.data(), 1024,
                         result_.data(), s, &pub);
  } else {
    LaunchRelativeChangeActive(next, now, coupling_.Runs(), partials_.data(), 1024,
                               result_.data(), s, &pub);
  }
  const double* r = Published(0, pub.ticket);
  if (r[2] != 0.0) return std::numeric_limits<double>::quiet_NaN();
  return r[0] > 0.0 ? std::sqrt(r[1]) / std::sqrt(r[0]) : 0.0;
}

void Viscoelastic3D::DropPassGraphs(unsigned long long flavour) {
  for (auto it = passGraphs_.begin(); it != passGraphs_.end();) {
    if (flavour == 0 || it->flavour == flavour) {
      if (it->exec != nullptr) (void)hipGraphExecDestroy(it->exec);
      it = passGraphs_.erase(it);
    } else {
      ++it;
    }
  }
}

// A pass replays the graph captured for its tensors, flavour and tide
// channel (captured on first use: the kernels and their arguments are
// those LaunchPass records). The first pass after a stored strain's start
// reads the start's own force (analysed_) and is launched directly.
void Viscoelastic3D::Pass(const double* s, double* sNext, bool useStartForce, Flavour& f) {
  rhsCurrent_ = false;
  if (!options_.graphPasses || useStartForce) {
    LaunchPass(s, sNext, useStartForce, f);
    return;
  }
  hipStream_t st = Stream();
  const bool tide = TideAmplitudes() != nullptr;
  PassGraph* g = nullptr;
  for (auto& p : passGraphs_) {
    if (p.s == s && p.next == sNext && p.flavour == f.id && p.tide == tide) g = &p;
  }
  if (g == nullptr) {
    hipGraph_t graph = nullptr;
    SL3_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      LaunchPass(s, sNext, false, f);
    } catch (...) {
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(st, &partial);
      if (partial != nullptr) (void)hipGraphDestroy(partial);
      throw;
    }
    SL3_HIP(hipStreamEndCapture(st, &graph));
    PassGraph p{s, sNext, f.id, tide, nullptr};
    const hipError_t made = hipGraphInstantiate(&p.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SL3_HIP(made);
    passGraphs_.push_back(p);
    g = &passGraphs_.back();
  }
  SL3_HIP(hipGraphLaunch(g->exec, st));
}

void Viscoelastic3D::LaunchPass(const double* s, double* sNext, bool useStartForce,
                                Flavour& f) {
  hipStream_t st = Stream();
  if (!useStartForce) {
    double* A[3] = {fourierA_[0].data(), fourierA_[1].data(), fourierA_[2].data()};
    double* B[3] = {fourierB_[0].data(), fourierB_[1].data(), fourierB_[2].data()};
    coupling_.SynthesiseToFourier(s, A, true, st, Live());
    coupling_.RingPass(A, B, weight_.data(), st, Live());
    coupling_.AnalyseFromFourier(B, analysed_.data(), fixed_.data(), true, st, Live(), true);
  }
  SolveAll(analysed_.data(), f);
  const auto* sph = tasksDev_.data();
  const auto* tor = tasksDev_.data() + (coupling_.LMax() + 1);
  LaunchStrain(sph, tor, coupling_.LMax(), coupling_.Geometry(), dofsMem_.data(),
               response_.data(), f.unitsMem.data(), amp_.data(),
               f.unitsMem.data() + dmHost_.size(), TideAmplitudes(), coupling_.View(sNext), st,
               Live());
}

void Viscoelastic3D::ElasticStrainCoefficients(DeviceState& state, const Channels& channels,
                                               double* out) {
  hipStream_t s = Stream();
  Flavour& f = FlavourOf(kBackwardEuler, 0.0);
  double* B[3] = {fourierB_[0].data(), fourierB_[1].data(), fourierB_[2].data()};
  for (int b = 0; b < 3; ++b) {
    RingLaunch r;
    r.block = b;
    r.op = kRingNodeScaled;
    r.gridIn = state.m[b].data();
    r.fourierOut = B[b];
    r.muNode = muNode_.data();
    coupling_.Ring(r, s);
  }
  coupling_.AnalyseFromFourier(B, analysed_.data(), nullptr, true, s);
  SolveAll(analysed_.data(), f);
  Amplitudes(channels, state.t);
  const auto* sph = tasksDev_.data();
  const auto* tor = tasksDev_.data() + (coupling_.LMax() + 1);
  LaunchStrain(sph, tor, coupling_.LMax(), coupling_.Geometry(), dofsMem_.data(),
               response_.data(), f.unitsMem.data(), amp_.data(),
               f.unitsMem.data() + dmHost_.size(), TideAmplitudes(), coupling_.View(out), s);
}

void Viscoelastic3D::StartStrain(DeviceState& state, const Channels& channels, double* out) {
  if (open_.state != nullptr) throw ArgumentError("a step is open on this operator");
  hipStream_t s = Stream();
  if (state.hasCoefficients) {
    LaunchCopy(out, state.coefficients.data(), coupling_.SpectralDoubles(), s);
  } else if (state.hasStrain) {
    double* B[3] = {fourierB_[0].data(), fourierB_[1].data(), fourierB_[2].data()};
    const double* grid[3] = {state.strain[0].data(), state.strain[1].data(),
                             state.strain[2].data()};
    coupling_.Analyse(grid, out, B, s);
  } else {
    ElasticStrainCoefficients(state, channels, out);
  }
  SL3_HIP(hipStreamSynchronize(s));
}

StepInfo Viscoelastic3D::Step(DeviceState& state, double dt, int scheme,
                              const Channels& channels, const double* start,
                              double warmGuard) {
  if (!(dt >= 0.0)) throw ArgumentError("dt must be non-negative");
  if (!(warmGuard >= 0.0)) throw ArgumentError("the warm-start guard must be non-negative");
  if (scheme != kExponentialTrapezoid && scheme != kBackwardEuler && scheme != kSdirk2) {
    throw ArgumentError(
        "the 3D step is implicit: ExponentialTrapezoid, BackwardEuler or Sdirk2");
  }
  if (dt == 0.0) scheme = kBackwardEuler;  // every scheme's zero-length step
  if (scheme == kSdirk2) return StepSdirk2(state, dt, channels);
  return StepOnce(state, dt, scheme, channels, start, warmGuard);
}

std::unique_ptr<DeviceState> Viscoelastic3D::CloneState(const DeviceState& from) const {
  auto to = std::make_unique<DeviceState>();
  CopyState(from, *to);
  SL3_HIP(hipStreamSynchronize(Stream()));
  return to;
}

void Viscoelastic3D::CopyState(const DeviceState& from, DeviceState& to) const {
  hipStream_t s = Stream();
  to.t = from.t;
  for (int b = 0; b < 3; ++b) {
    to.m[b].Resize(from.m[b].size());
    LaunchCopy(to.m[b].data(), from.m[b].data(), from.m[b].size(), s);
    if (from.hasStrain) {
      to.strain[b].Resize(from.strain[b].size());
      LaunchCopy(to.strain[b].data(), from.strain[b].data(), from.strain[b].size(), s);
    }
  }
  to.hasStrain = from.hasStrain;
  if (from.hasCoefficients) {
    to.coefficients.Resize(from.coefficients.size());
    LaunchCopy(to.coefficients.data(), from.coefficients.data(), from.coefficients.size(), s);
  }
  to.hasCoefficients = from.hasCoefficients;
}

// Sdirk2 by its recipe (src/StepRecipe.cpp:25-32 with StepByRecipe,
// include/sl3dpp/StepRecipe.hpp:164-192, and StageState,
// src/Viscoelastic3D.cpp:463-492): Y1 = BE(gamma h) from the state; the
// second stage starts from m = (1 - r) m_n + r m(Y1), r = (1 - gamma) /
// gamma, with Y1's strain and coefficients, at t_n + (1 - gamma) h, and is
// BE(gamma h) again. The passes add up; the last stage's residual stands.
StepInfo Viscoelastic3D::StepSdirk2(DeviceState& state, double dt, const Channels& channels) {
  hipStream_t s = Stream();
  const double tn = state.t;
  const double length = kSdirk2Gamma * dt;
  // both stages' tides checked before the state changes
  CheckTide(channels, tn + length);
  CheckTide(channels, tn + dt);
  const double r = (1.0 - kSdirk2Gamma) / kSdirk2Gamma;
  if (!stage_) stage_ = std::make_unique<DeviceState>();
  DeviceState& y1 = *stage_;
  CopyState(state, y1);
  const StepInfo first = StepOnce(y1, length, kBackwardEuler, channels, nullptr, 0.0);
  const auto firstHistory = residualHistory;
  const auto firstInfos = realisationInfos_;
  const Timing firstTiming = lastTiming;
  for (int b = 0; b < 3; ++b) {
    LaunchCombine(state.m[b].data(), 1.0 - r, y1.m[b].data(), r, state.m[b].size(), s);
  }
  DeviceState staged;
  staged.t = tn + (1.0 - kSdirk2Gamma) * dt;
  for (int b = 0; b < 3; ++b) {
    staged.m[b] = std::move(state.m[b]);
    staged.strain[b] = std::move(y1.strain[b]);
  }
  staged.hasStrain = y1.hasStrain;
  staged.coefficients = std::move(y1.coefficients);
  staged.hasCoefficients = y1.hasCoefficients;
  const StepInfo second = StepOnce(staged, length, kBackwardEuler, channels, nullptr, 0.0);
  // the stepped stage becomes the state; Y1's buffers return to the cache
  for (int b = 0; b < 3; ++b) {
    state.m[b] = std::move(staged.m[b]);
    y1.strain[b] = std::move(state.strain[b]);
    state.strain[b] = std::move(staged.strain[b]);
  }
  y1.coefficients = std::move(state.coefficients);
  state.coefficients = std::move(staged.coefficients);
  state.hasStrain = staged.hasStrain;
  state.hasCoefficients = staged.hasCoefficients;
  state.t = tn + dt;
  auto stages = [](const StepInfo& a, const StepInfo& b) {
    StepInfo info;
    info.iterations = a.iterations + b.iterations;
    info.mixedPasses = a.mixedPasses + b.mixedPasses;
    info.solveIterations = a.solveIterations + b.solveIterations;
    info.residual = b.residual;
    info.converged = a.converged && b.converged;
    return info;
  };
  StepInfo info = stages(first, second);
  for (size_t r = 0; r < realisationInfos_.size() && r < firstInfos.size(); ++r) {
    realisationInfos_[r] = stages(firstInfos[r], realisationInfos_[r]);
  }
  residualHistory.insert(residualHistory.begin(), firstHistory.begin(), firstHistory.end());
  lastTiming.startStep += firstTiming.startStep;
  lastTiming.passes += firstTiming.passes;
  lastTiming.finish += firstTiming.finish;
  return info;
}

StepInfo Viscoelastic3D::StepOnce(DeviceState& state, double dt, int scheme,
                                  const Channels& channels, const double* start,
                                  double warmGuard) {
  OpenStart(state, dt, scheme, channels);
  StepInfo info;
  try {
    info = OpenClose(start, warmGuard);
  } catch (...) {
    open_.state = nullptr;
    throw;
  }
  if (coupling_.Realisations() == 1) realisationInfos_.assign(1, info);
  OpenFinish(nullptr);
  hipStream_t s = Stream();
  SL3_HIP(hipStreamSynchronize(s));
  SL3_HIP(hipEventElapsedTime(&lastTiming.startStep, events_[0], events_[1]));
  SL3_HIP(hipEventElapsedTime(&lastTiming.passes, events_[1], events_[2]));
  SL3_HIP(hipEventElapsedTime(&lastTiming.finish, events_[2], events_[3]));
  return info;
}

// StartStep (src/Viscoelastic3D.cpp:1223-1378).
void Viscoelastic3D::OpenStart(DeviceState& state, double dt, int scheme,
                               const Channels& channels) {
  if (open_.state != nullptr) throw ArgumentError("a step is already open on this operator");
  if (!(dt >= 0.0)) throw ArgumentError("the step length must be non-negative");
  if (scheme != kExponentialTrapezoid && scheme != kBackwardEuler) {
    throw ArgumentError("a step is opened stage by stage: exponential trapezoid or backward Euler");
  }
  hipStream_t s = Stream();
  const size_t S = coupling_.SpectralDoubles();
  double* A[3] = {fourierA_[0].data(), fourierA_[1].data(), fourierA_[2].data()};
  double* B[3] = {fourierB_[0].data(), fourierB_[1].data(), fourierB_[2].data()};
  Flavour& f = FlavourOf(scheme, dt);
  lastStep_ = {scheme, dt};
  rhsCurrent_ = false;
  const double tEnd = state.t + dt;
  SetAllLive();  // a converge cut short by an error may have left some marked

  SL3_HIP(hipEventRecord(events_[0], s));
  // s0 (src/Viscoelastic3D.cpp:1252-1263)
  if (state.hasCoefficients) {
    LaunchCopy(s_.data(), state.coefficients.data(), S, s);
  } else if (state.hasStrain) {
    const double* grid[3] = {state.strain[0].data(), state.strain[1].data(),
                             state.strain[2].data()};
    coupling_.Analyse(grid, s_.data(), B, s);
  } else {
    ElasticStrainCoefficients(state, channels, s_.data());
  }
  // the channels at the end time, gathered (and a degree-1 tide rejected)
  // before the sweep rewrites the state's memory
  Amplitudes(channels, tEnd);
  // StartStep's sweep (:1283-1376)
  if (!state.hasStrain) coupling_.SynthesiseToFourier(s_.data(), A, false, s);
  for (int b = 0; b < 3; ++b) {
    RingLaunch r;
    r.block = b;
    r.op = state.hasStrain ? kRingStart : kRingStartFromSpectral;
    r.gridIn = state.hasStrain ? state.strain[b].data() : nullptr;
    r.fourierIn = A[b];
    r.fourierOut = B[b];   // the fixed force mu m~
    r.fourierOut2 = A[b];  // W d0 (each workgroup rewrites only its own rings)
    r.memIn = state.m[b].data();
    r.memOut = state.m[b].data();
    r.itau = itau_.data();
    r.betaOut = beta_.data();
    r.weightOut = weight_.data();
    r.muNode = muNode_.data();
    r.itauRefNode = itauRefNode_.data();
    r.dt = dt;
    r.scheme = scheme;
    coupling_.Ring(r, s);
  }
  coupling_.AnalyseFromFourier(B, fixed_.data(), nullptr, true, s);
  const bool cg = options_.solver == kConjugateGradients;
  // analysed_ stays zero below each degree's truncation (only truncated
  // analyses write it, and the force reads it only above): its analyses
  // leave those entries
  // the first pass's start force (the fixed point's only)
  if (!cg) coupling_.AnalyseFromFourier(A, analysed_.data(), fixed_.data(), true, s, nullptr, true);
  SL3_HIP(hipEventRecord(events_[1], s));
  open_.state = &state;
  open_.flavour = &f;
  open_.scheme = scheme;
  open_.dt = dt;
  open_.tEnd = tEnd;
  open_.startForce = !cg;
}

void Viscoelastic3D::RequireOpen() const {
  if (open_.state == nullptr) throw ArgumentError("no step is open on this operator");
}

// One pass of the open step: Respond + AddChannels + StrainOf
// (src/Viscoelastic3D.cpp:1758-1786, 1029-1073).
void Viscoelastic3D::OpenPass(const double* s, double* next) {
  RequireOpen();
  const bool useStartForce = open_.startForce && s == s_.data();
  open_.startForce = false;
  Pass(s, next, useStartForce, *open_.flavour);
}

void Viscoelastic3D::OpenTransposedPass(const double* z, const double* const src[3],
                                        double* out) {
  RequireOpen();
  hipStream_t st = Stream();
  Flavour& f = *open_.flavour;
  const size_t S = coupling_.SpectralDoubles();
  if (zeroAmp_.size() != 2 * ampDoubles_) {
    zeroAmp_.Resize(2 * ampDoubles_);
    LaunchFill(zeroAmp_.data(), zeroAmp_.size(), 0.0, st);
  }
  strainT_.Resize(S);
  double* A[3] = {fourierA_[0].data(), fourierA_[1].data(), fourierA_[2].data()};
  double* B[3] = {fourierB_[0].data(), fourierB_[1].data(), fourierB_[2].data()};
  // the source's analysis first (it shares the Fourier buffers)
  const double* addend = nullptr;
  if (src != nullptr) {
    sourceT_.Resize(S);
    for (int b = 0; b < 3; ++b) {
      RingLaunch r;
      r.block = b;
      r.op = kRingAnalyse;
      r.gridIn = src[b];
      r.fourierOut = B[b];
      coupling_.Ring(r, st);
    }
    coupling_.AnalyseFromFourier(B, sourceT_.data(), nullptr, true, st);
    addend = sourceT_.data();
  }
  // Entry and SolveForces: the forces of z and the memory response
  SolveAll(z, f);
  // StrainOf of the dual solutions, no channels
  const auto* sph = tasksDev_.data();
  const auto* tor = tasksDev_.data() + (coupling_.LMax() + 1);
  LaunchStrain(sph, tor, coupling_.LMax(), coupling_.Geometry(), dofsMem_.data(),
               response_.data(), f.unitsMem.data(), zeroAmp_.data(),
               f.unitsMem.data() + dmHost_.size(), nullptr, coupling_.View(strainT_.data()), st);
  // Exit's synthesis, the lateral weight, the analysis (+ the source)
  coupling_.SynthesiseToFourier(strainT_.data(), A, true, st);
  coupling_.RingPass(A, B, weight_.data(), st);
  coupling_.AnalyseFromFourier(B, out, addend, true, st);
}

StepInfo Viscoelastic3D::OpenClose(const double* start, double warmGuard) {
  RequireOpen();
  StepInfo info;
  residualHistory.clear();
  if (options_.solver == kConjugateGradients) {
    info = CloseByConjugateGradients(*open_.flavour);  // takes no warm start, as the reference's
    SL3_HIP(hipEventRecord(events_[2], Stream()));
    return info;
  }
  return OpenConverge(PassFn(), start, warmGuard);
}

// Converge (src/Viscoelastic3D.cpp:1916-1974): the fixed point s = pass(s)
// from the start strain (or `start`), adaptively Anderson-mixed; the
// result in s_.
StepInfo Viscoelastic3D::OpenConverge(const PassFn& pass, const double* start,
                                      double warmGuard) {
  RequireOpen();
  if (coupling_.Realisations() > 1) return ConvergeBatch(pass, start, warmGuard);
  hipStream_t s = Stream();
  const size_t S = coupling_.SpectralDoubles();
  StepInfo info;
  residualHistory.clear();
  bool warm = false;
  const bool guarded = start != nullptr && warmGuard > 0.0;
  if (guarded) {
    // the step's own start, kept in case the warm one is dropped
    coldStart_.Resize(S);
    LaunchCopy(coldStart_.data(), s_.data(), S, s);
  }
  const bool startForce = open_.startForce;
  if (start != nullptr) {
    LaunchCopy(s_.data(), start, S, s);
    warm = true;
    open_.startForce = false;
  }
  int first = 0;  // the first pass of the sequence being converged
  const bool adaptive = options_.andersonRate > 0.0;
  const int memory = options_.andersonMemory;
  bool mixing = memory > 0 && !adaptive;
  bool mixerStarted = false;
  int count = 0, slot = 0;
  std::vector<double> gram(static_cast<size_t>(std::max(memory, 1)) * std::max(memory, 1), 0.0);
  const ActiveRuns& runs = coupling_.Runs();
  const size_t C = static_cast<size_t>(runs.total);  // the compact history's length
  if (memory > 0 && history_.size() != 2 * static_cast<size_t>(memory) * C) {
    history_.Resize(2 * static_cast<size_t>(memory) * C);
    xPrev_.Resize(C);
    fPrev_.Resize(C);
  }
  double previous = 0.0;
  auto sweepOf = [&](int m) {
    AndersonSweep sw{};
    sw.x = s_.data();
    sw.gx = sNext_.data();
    sw.xPrev = xPrev_.data();
    sw.fPrev = fPrev_.data();
    for (int i = 0; i < memory && i < kAndersonMax; ++i) {
      sw.dF[i] = history_.data() + static_cast<size_t>(i) * C;
      sw.dG[i] = history_.data() + static_cast<size_t>(memory + i) * C;
    }
    sw.m = m;
    sw.slot = slot;
    return sw;
  };
  for (int it = 0; it < options_.maxIterations; ++it) {
    if (pass) {
      pass(s_.data(), sNext_.data());
    } else {
      open_.startForce = false;
      Pass(s_.data(), sNext_.data(), it == 0 && !warm && startForce, *open_.flavour);
    }
    // Once the mixer runs (and without damping) every pass that does not
    // converge takes a sweep over the same two iterates: the sweep then also
    // forms the pass's RelativeChange, one read of the data and one
    // synchronisation. A sweep on the converging pass only writes history
    // that the next step starts afresh.
    const bool fused = mixing && mixerStarted && options_.damping == 1.0;
    const int m = std::min(count + 1, memory);
    AndersonSweep sw{};
    double sums[2 * kAndersonMax];
    const int width = AndersonWidth(m);  // the Gram column at [0, w), the right-hand side after
    double residual;
    if (fused) {
      sw = sweepOf(m);
      const HostPublish pub = Publish(0);
      LaunchAndersonSweep(sw, runs, partials_.data(), 1024, result_.data(),
                          result_.data() + 2 * width, s, &pub);
      const double* r = Published(0, pub.ticket);
      std::copy(r, r + 2 * width, sums);
      const double* rc = r + 2 * width;
      residual = rc[2] != 0.0 ? std::numeric_limits<double>::quiet_NaN()
                 : rc[0] > 0.0 ? std::sqrt(rc[1]) / std::sqrt(rc[0])
                               : 0.0;
    } else {
      // Every pass's image is zero below each degree's truncation; the start
      // (a warm start, or one analysed from a stored strain grid) may hold
      // anything there, so the first pass compares every entry.
      residual = RelativeChange(sNext_.data(), s_.data(), it == first);
    }
    info.iterations = it + 1;
    info.residual = residual;
    residualHistory.push_back(residual);
    if (guarded && it == 0 && !(residual <= warmGuard)) {
      // the warm start's first change exceeds the guard: converge from the
      // step's own start instead (the pass counts)
      LaunchCopy(s_.data(), coldStart_.data(), S, s);
      info.warmFallback = true;
      first = 1;
      continue;
    }
    if (!std::isfinite(residual)) break;
    if (residual <= options_.rtol) {
      info.converged = true;
      std::swap(s_, sNext_);
      break;
    }
    if (options_.damping != 1.0) {
      ProfileScope prof854("damp", s);
      hipLaunchKernelGGL(DampKernel, dim3(Blocks(S)), dim3(256), 0, s, sNext_.data(),
                         s_.data(), options_.damping, S);
      SL3_HIP(hipGetLastError());
    }
    if (adaptive && !mixing && memory > 0 && it > first &&
        residual > options_.andersonRate * previous) {
      mixing = true;
    }
    previous = residual;
    if (mixing) {
      if (!mixerStarted) {
        LaunchAndersonFirst(s_.data(), sNext_.data(), xPrev_.data(), fPrev_.data(), runs, S,
                            s);
        mixerStarted = true;
      } else {
        if (!fused) {
          sw = sweepOf(m);
          const HostPublish pub = Publish(0);
          LaunchAndersonSweep(sw, runs, partials_.data(), 1024, result_.data(), nullptr, s,
                              &pub);
          const double* r = Published(0, pub.ticket);
          std::copy(r, r + 2 * width, sums);
        }
        count = m;
        for (int i = 0; i < m; ++i) {
          gram[slot * memory + i] = sums[i];
          gram[i * memory + slot] = sums[i];
        }
        slot = (slot + 1) % memory;
        double g[kAndersonMax * kAndersonMax], gamma[kAndersonMax];
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) g[i * m + j] = gram[i * memory + j];
        GramLeastSquares(m, g, sums + width, gamma);
        LaunchAndersonMix(sw, gamma, runs, s);
      }
      ++info.mixedPasses;
    } else {
      std::swap(s_, sNext_);
    }
  }
  SL3_HIP(hipEventRecord(events_[2], s));
  return info;
}

// Converge (src/Viscoelastic3D.cpp:1916-1974) for realisations side by
// side. Each pass runs every realisation through the same kernels; each
// realisation then takes what OpenConverge takes alone, on its own runs:
// its residual (from a fused sweep once its mixer runs), its adaptive
// switch, its Anderson history and mix, the Picard image copied into its
// columns of the iterate in place of the buffers' swap. A converged
// realisation's image becomes its columns of the iterate and nothing
// writes them again (its columns still go through the passes).
StepInfo Viscoelastic3D::ConvergeBatch(const PassFn& pass, const double* start,
                                       double warmGuard) {
  if (warmGuard > 0.0) {
    throw ArgumentError("realisations side by side take no guarded warm start");
  }
  hipStream_t s = Stream();
  const int R = coupling_.Realisations();
  const size_t S = coupling_.SpectralDoubles();
  SetAllLive();
  residualHistory.clear();
  const bool startForce = open_.startForce;
  bool warm = false;
  if (start != nullptr) {
    LaunchCopy(s_.data(), start, S, s);
    warm = true;
    open_.startForce = false;
  }
  const bool adaptive = options_.andersonRate > 0.0;
  const int memory = options_.andersonMemory;
  const size_t C = static_cast<size_t>(coupling_.Runs().total);  // per realisation
  const size_t historyEach = 2 * static_cast<size_t>(memory) * C;
  if (memory > 0 && history_.size() != historyEach * R) {
    history_.Resize(historyEach * R);
    xPrev_.Resize(C * R);
    fPrev_.Resize(C * R);
  }
  struct Control {
    bool live = true, mixing = false, started = false, fused = false;
    int count = 0, slot = 0, m = 0, width = 0;
    double previous = 0.0, residual = 0.0;
    std::vector<double> gram;
    StepInfo info;
  };
  std::vector<Control> ctl(R);
  for (auto& c : ctl) {
    c.mixing = memory > 0 && !adaptive;
    c.gram.assign(static_cast<size_t>(std::max(memory, 1)) * std::max(memory, 1), 0.0);
  }
  // realisation 0's sweep pointers; an entry of a batch moves them to its
  // realisation's history
  AndersonSweep base{};
  base.x = s_.data();
  base.gx = sNext_.data();
  base.xPrev = xPrev_.data();
  base.fPrev = fPrev_.data();
  for (int i = 0; i < memory && i < kAndersonMax; ++i) {
    base.dF[i] = history_.data() + static_cast<size_t>(i) * C;
    base.dG[i] = history_.data() + static_cast<size_t>(memory + i) * C;
  }
  auto batch = [&]() {
    RealisationBatch b;
    b.prevStride = static_cast<long long>(C);
    b.historyStride = static_cast<long long>(historyEach);
    return b;
  };
  auto add = [&](RealisationBatch& b, int r, int every) {
    const int e = b.count++;
    b.r[e] = r;
    b.m[e] = ctl[r].m;
    b.slot[e] = ctl[r].slot;
    b.every[e] = every;
  };
  const ActiveRuns runs = coupling_.Runs(0), every = coupling_.EveryRuns(0);
  const int stride = 2 * kAndersonMax + 3;  // one realisation's results
  std::vector<double> back(static_cast<size_t>(stride) * R);
  // the batch's results to the host (one launch) and back
  auto collect = [&](const RealisationBatch& b) {
    if (b.count == 0) return;
    HostPublish pub;
    pub.values = mappedValuesDev_;
    pub.seq = mappedSeqDev_;
    pub.ticket = ++ticket_;
    LaunchPublishBatch(result_.data(), stride, stride, b, pub, kReadBackMax, s);
    for (int e = 0; e < b.count; ++e) {
      const double* v = Published(b.r[e], pub.ticket);
      std::copy(v, v + stride, back.data() + static_cast<size_t>(b.r[e]) * stride);
    }
  };
  for (int it = 0; it < options_.maxIterations; ++it) {
    if (std::none_of(ctl.begin(), ctl.end(), [](const Control& c) { return c.live; })) break;
    if (pass) {
      pass(s_.data(), sNext_.data());
    } else {
      open_.startForce = false;
      Pass(s_.data(), sNext_.data(), it == 0 && !warm && startForce, *open_.flavour);
    }
    // every live realisation's residual: a fused sweep once its mixer runs
    // (the sums at [0, 2 w), the maxima after), else RelativeChange (the
    // maxima at 0; the first pass over every entry); one launch of each
    RealisationBatch fused = batch(), plain = batch(), live = batch();
    for (int r = 0; r < R; ++r) {
      Control& c = ctl[r];
      if (!c.live) continue;
      c.fused = c.mixing && c.started && options_.damping == 1.0;
      c.m = std::min(c.count + 1, memory);
      if (c.fused) {
        add(fused, r, 0);
      } else {
        add(plain, r, it == 0 ? 1 : 0);
      }
      add(live, r, 0);
    }
    const int fusedWidth = LaunchAndersonSweepBatch(base, runs, fused, true, partials_.data(),
                                                    1024, result_.data(), stride, s);
    LaunchRelativeChangeBatch(sNext_.data(), s_.data(), runs, every, plain, partials_.data(), 1024,
                              result_.data(), stride, s);
    collect(live);
    double largest = 0.0;
    RealisationBatch done = batch();
    for (int r = 0; r < R; ++r) {
      Control& c = ctl[r];
      if (!c.live) continue;
      if (c.fused) c.width = fusedWidth;
      const double* rc = back.data() + static_cast<size_t>(r) * stride + (c.fused ? 2 * c.width : 0);
      c.residual = rc[2] != 0.0 ? std::numeric_limits<double>::quiet_NaN()
                   : rc[0] > 0.0 ? std::sqrt(rc[1]) / std::sqrt(rc[0])
                                 : 0.0;
      c.info.iterations = it + 1;
      c.info.residual = c.residual;
      largest = std::isnan(c.residual) || std::isnan(largest) ? std::numeric_limits<double>::quiet_NaN()
                                                            : std::max(largest, c.residual);
      if (!std::isfinite(c.residual)) {
        c.live = false;
      } else if (c.residual <= options_.rtol) {
        c.info.converged = true;
        c.live = false;
        add(done, r, 0);
      }
      // the next passes skip a realisation that stopped
      if (!c.live) SL3_HIP(hipMemsetAsync(live_.data() + r, 0, sizeof(int), s));
    }
    LaunchCopyRunsBatch(s_.data(), sNext_.data(), every, done, s);
    residualHistory.push_back(largest);
    if (options_.damping != 1.0) {
      ProfileScope prof("damp", s);
      hipLaunchKernelGGL(DampKernel, dim3(Blocks(S)), dim3(256), 0, s, sNext_.data(),
                         s_.data(), options_.damping, S);
      SL3_HIP(hipGetLastError());
    }
    // the sweeps a damped pass takes after its residual
    RealisationBatch damped = batch();
    for (int r = 0; r < R; ++r) {
      Control& c = ctl[r];
      if (!c.live) continue;
      if (adaptive && !c.mixing && memory > 0 && it > 0 &&
          c.residual > options_.andersonRate * c.previous) {
        c.mixing = true;
      }
      c.previous = c.residual;
      if (c.mixing && c.started && !c.fused) add(damped, r, 0);
    }
    if (damped.count > 0) {
      const int w = LaunchAndersonSweepBatch(base, runs, damped, false, partials_.data(), 1024,
                                             result_.data(), stride, s);
      for (int e = 0; e < damped.count; ++e) ctl[damped.r[e]].width = w;
      collect(damped);
    }
    // each realisation's move: its Picard image, its first mixed pass, or
    // its mix with the gamma solved here (uploaded for all at once)
    RealisationBatch picard = batch(), first = batch(), mix = batch();
    for (int r = 0; r < R; ++r) {
      Control& c = ctl[r];
      if (!c.live) continue;
      if (!c.mixing) {
        add(picard, r, 0);
        continue;
      }
      if (!c.started) {
        add(first, r, 0);
        c.started = true;
      } else {
        const double* sums = back.data() + static_cast<size_t>(r) * stride;
        const int m = c.m;
        c.count = m;
        for (int i = 0; i < m; ++i) {
          c.gram[c.slot * memory + i] = sums[i];
          c.gram[i * memory + c.slot] = sums[i];
        }
        double g[kAndersonMax * kAndersonMax];
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) g[i * m + j] = c.gram[i * memory + j];
        double* gamma = gammaHost_ + static_cast<size_t>(mix.count) * kAndersonMax;
        std::fill(gamma, gamma + kAndersonMax, 0.0);
        GramLeastSquares(m, g, sums + c.width, gamma);
        add(mix, r, 0);  // m as swept; the slot no longer matters to the mix
        c.slot = (c.slot + 1) % memory;
      }
      ++c.info.mixedPasses;
    }
    LaunchCopyRunsBatch(s_.data(), sNext_.data(), every, picard, s);
    LaunchAndersonFirstBatch(base, runs, every, first, s);
    if (mix.count > 0) {
      SL3_HIP(hipMemcpyAsync(gammaDev_.data(), gammaHost_,
                             sizeof(double) * kAndersonMax * mix.count, hipMemcpyHostToDevice, s));
      LaunchAndersonMixBatch(base, runs, mix, gammaDev_.data(), s);
    }
  }
  SetAllLive();
  realisationInfos_.clear();
  StepInfo info;
  info.converged = true;
  for (const auto& c : ctl) {
    realisationInfos_.push_back(c.info);
    info.iterations = std::max(info.iterations, c.info.iterations);
    info.mixedPasses = std::max(info.mixedPasses, c.info.mixedPasses);
    info.converged = info.converged && c.info.converged;
    info.residual = std::isnan(c.info.residual) || std::isnan(info.residual)
                        ? std::numeric_limits<double>::quiet_NaN()
                        : std::max(info.residual, c.info.residual);
  }
  SL3_HIP(hipEventRecord(events_[2], s));
  return info;
}

// Finish (src/Viscoelastic3D.cpp:1885-1914): d = Synthesise(strain),
// m+ = m~ + beta d, the strain and its coefficients kept, the state at the
// end time. strain: device coefficients, or null for the iterate.
void Viscoelastic3D::OpenFinish(const double* strain) {
  RequireOpen();
  DeviceState& state = *open_.state;
  hipStream_t s = Stream();
  const size_t S = coupling_.SpectralDoubles();
  if (strain != nullptr && strain != s_.data()) LaunchCopy(s_.data(), strain, S, s);
  double* A[3] = {fourierA_[0].data(), fourierA_[1].data(), fourierA_[2].data()};
  coupling_.SynthesiseToFourier(s_.data(), A, true, s);
  for (int b = 0; b < 3; ++b) {
    if (!state.hasStrain) state.strain[b].Resize(coupling_.GridDoubles(b));
    RingLaunch r;
    r.block = b;
    r.op = kRingFinish;
    r.fourierIn = A[b];
    r.gridOut = state.strain[b].data();
    r.memOut = state.m[b].data();
    r.beta = beta_.data();
    coupling_.Ring(r, s);
  }
  state.hasStrain = true;
  if (state.coefficients.size() != S) state.coefficients.Resize(S);
  LaunchCopy(state.coefficients.data(), s_.data(), S, s);
  state.hasCoefficients = true;
  state.t = open_.tEnd;
  SL3_HIP(hipEventRecord(events_[3], s));
  open_.state = nullptr;
}

void Viscoelastic3D::SetAllLive() {
  if (live_.size() == 0) return;
  SL3_HIP(hipMemsetAsync(live_.data(), 0xff, live_.size() * sizeof(int), Stream()));
}

void Viscoelastic3D::RequireOneRealisation(const char* what) const {
  if (coupling_.Realisations() != 1) {
    throw ArgumentError(std::string(what) + " takes an operator of one realisation");
  }
}

void Viscoelastic3D::ForceOf(const double* weighted, double* forces) {
  RequireOneRealisation("the force in the solution layout");
  hipStream_t s = Stream();
  boundaryRows_.Resize(std::max<size_t>(fDoubles_, 1));
  LaunchFill(boundaryRows_.data(), boundaryRows_.size(), 0.0, s);
  LaunchForce(tasksDev_.data(), coupling_.Geometry(),
              coupling_.View(const_cast<double*>(weighted)), dofsMem_.data(),
              forceWork_.data(), forceWorkCount_, boundaryRows_.data(), s);
  LaunchFill(forces, solutionDoubles_, 0.0, s);
  LaunchRowsToSolutions(tasksDev_.data(), static_cast<int>(tasks_.size()), nmMax_,
                        coupling_.LMax(), dm_.data(), solutionOffset_.data(),
                        boundaryRows_.data(), forces, s);
}

void Viscoelastic3D::StrainOfSolutions(const double* solutions, double* spectral) {
  RequireOneRealisation("the strain of solutions");
  hipStream_t s = Stream();
  boundaryRows_.Resize(std::max<size_t>(fDoubles_, 1));
  LaunchFill(boundaryRows_.data(), boundaryRows_.size(), 0.0, s);
  LaunchSolutionsToRows(tasksDev_.data(), static_cast<int>(tasks_.size()), nmMax_,
                        coupling_.LMax(), dm_.data(), solutionOffset_.data(), solutions,
                        boundaryRows_.data(), s);
  const int L = coupling_.LMax();
  LaunchStrain(tasksDev_.data(), tasksDev_.data() + (L + 1), L, coupling_.Geometry(),
               dofsMem_.data(), boundaryRows_.data(), nullptr, nullptr, nullptr, nullptr,
               coupling_.View(spectral), s);
}

// OpenStep::AddMemory (include/sl3dpp/Viscoelastic3D.hpp:576-583): an
// increment of the decayed memory joins it and the fixed force
// (mu Analyse(increment), node by node); the next pass forms its force
// afresh.
void Viscoelastic3D::OpenAddMemory(const double* const increment[3]) {
  RequireOpen();
  hipStream_t s = Stream();
  DeviceState& state = *open_.state;
  double* B[3] = {fourierB_[0].data(), fourierB_[1].data(), fourierB_[2].data()};
  for (int b = 0; b < 3; ++b) {
    LaunchCombine(state.m[b].data(), 1.0, increment[b], 1.0, state.m[b].size(), s);
  }
  coupling_.Analyse(increment, sNext_.data(), B, s);
  LaunchScaleByNode(muNode_.data(), sNext_.data(), coupling_.SpectralDoubles() / 2,
                    coupling_.NodeCount(), s);
  LaunchCombine(fixed_.data(), 1.0, sNext_.data(), 1.0, coupling_.SpectralDoubles(), s);
  open_.startForce = false;
}

// OpenStep::AddForce's tensor (:585-598): an analysed tensor joins the
// fixed force; the next pass forms its force afresh.
void Viscoelastic3D::OpenAddForce(const double* analysed) {
  RequireOpen();
  LaunchCombine(fixed_.data(), 1.0, analysed, 1.0, coupling_.SpectralDoubles(), Stream());
  open_.startForce = false;
}

void Viscoelastic3D::LateralForce(const double* response, bool channels, double* force,
                                  Flavour& f) {
  hipStream_t s = Stream();
  const auto* sph = tasksDev_.data();
  const auto* tor = tasksDev_.data() + (coupling_.LMax() + 1);
  // the strain (the unit responses' pointers stay valid: the kernel reads
  // them times a zero amplitude when there are no channels)
  LaunchStrain(sph, tor, coupling_.LMax(), coupling_.Geometry(), dofsMem_.data(), response,
               f.unitsMem.data(), channels ? amp_.data() : nullptr,
               f.unitsMem.data() + dmHost_.size(), channels ? TideAmplitudes() : nullptr,
               coupling_.View(sNext_.data()), s);
  double* A[3] = {fourierA_[0].data(), fourierA_[1].data(), fourierA_[2].data()};
  double* B[3] = {fourierB_[0].data(), fourierB_[1].data(), fourierB_[2].data()};
  coupling_.SynthesiseToFourier(sNext_.data(), A, true, s);
  coupling_.RingPass(A, B, weight_.data(), s);
  coupling_.AnalyseFromFourier(B, analysed_.data(), nullptr, true, s);
  LaunchForce(tasksDev_.data(), coupling_.Geometry(), coupling_.View(analysed_.data()),
              dofsMem_.data(), forceWork_.data(), forceWorkCount_, force, s);
}

double Viscoelastic3D::Pair(const double* x, const double* force, const double* g,
                            const ChannelTerms& ch) {
  hipStream_t s = Stream();
  LaunchLateralPair(tasksDev_.data(), pairWork_.data(), pairWorkCount_, x, force, g, ch,
                    cgPartials_.data(), result_.data(), s);
  double v = 0.0;
  ReadBack(&v, result_.data(), 1);
  return v;
}

// detail::LateralConjugateGradients (include/sl3dpp/detail/LateralSolve.hpp:
// 110-183) with no start, on the memory dofs (lateral_cg.hpp): x = x0 +
// K^-1 Pᵀ c, kept as phi = f_m + c (its force) and y = G phi (its response
// without the channels), so the last pass's export solves phi as the fixed
// point's does its force.
StepInfo Viscoelastic3D::CloseByConjugateGradients(Flavour& f) {
  hipStream_t s = Stream();
  const size_t n = std::max<size_t>(fDoubles_, 1);
  for (auto* v : {&cgPhi_, &cgY_, &cgRho_, &cgGrho_, &cgPi_, &cgGpi_, &cgQ_}) {
    if (v->size() != n) {
      v->Resize(n);
      LaunchFill(v->data(), n, 0.0, s);  // the padding stays zero
    }
  }
  cgPartials_.Resize(std::max(pairWorkCount_, 1));
  const ChannelTerms none{};
  auto response = [&](const double* force, double* out) {
    LaunchResponse(tasksDev_.data(), work_.data(), workCount_, narrowCount_, f.response.data(),
                   force, out,
                   s);
  };
  // x0 = K^-1 f (OpenStep::Forces, :1448-1471): the fixed memory force's
  // response; the channels enter through their unit responses
  SolveAll(fixed_.data(), f);  // force_ = f_m, response_ = G f_m
  // <x0, f> = <G f_m + 2 chan, f_m> + <chan force, K^-1 chan force>, K^-1
  // symmetric (LateralSolve.hpp:118)
  double chanchan = 0.0;
  const int L = coupling_.LMax();
  for (int l = 0; l <= L; ++l) {
    const double* e = f.channelEnergy.data() + 4 * l;
    for (int m = 0; m <= l; ++m) {
      const size_t pair = static_cast<size_t>(l) * (l + 1) / 2 + m;
      for (int part = 0; part < 2; ++part) {
        const double a = ampHost_[2 * pair + part];
        const double b = hasTide_ ? ampHost_[ampDoubles_ + 2 * pair + part] : 0.0;
        chanchan += (m == 0 ? 1.0 : 2.0) * (a * a * e[0] + a * b * (e[1] + e[2]) + b * b * e[3]);
      }
    }
  }
  ChannelTerms twice{f.unitsMem.data(), amp_.data(), f.unitsMem.data() + dmHost_.size(),
                     TideAmplitudes(), 2.0};
  const double energy = Pair(response_.data(), force_.data(), nullptr, twice) + chanchan;
  solveEnergy.assign(1, energy);
  LaunchCopy(cgY_.data(), response_.data(), n, s);
  // r0 = f - A x0 = Pᵀ L P x0
  LateralForce(response_.data(), true, cgRho_.data(), f);
  LaunchCopy(cgPhi_.data(), force_.data(), n, s);
  response(cgRho_.data(), cgGrho_.data());
  double rz = Pair(cgGrho_.data(), cgRho_.data(), nullptr, none);
  auto relative = [&](double v) {
    return energy > 0.0 ? std::sqrt(std::abs(v) / energy) : std::sqrt(std::abs(v));
  };
  StepInfo info;
  info.residual = relative(rz);
  residualHistory.push_back(info.residual);
  LaunchCopy(cgPi_.data(), cgRho_.data(), n, s);
  LaunchCopy(cgGpi_.data(), cgGrho_.data(), n, s);
  const double target = options_.solveRtol;
  for (int it = 0; it < options_.solveMaxIterations; ++it) {
    if (info.residual <= target || !std::isfinite(info.residual)) break;
    LateralForce(cgGpi_.data(), false, cgQ_.data(), f);  // q = L G pi
    const double curvature = Pair(cgGpi_.data(), cgPi_.data(), cgQ_.data(), none);
    const double alpha = rz / curvature;
    LaunchCgAdvance(cgPhi_.data(), cgY_.data(), cgRho_.data(), cgPi_.data(), cgGpi_.data(),
                    cgQ_.data(), alpha, n, s);
    response(cgRho_.data(), cgGrho_.data());
    const double next = Pair(cgGrho_.data(), cgRho_.data(), nullptr, none);
    const double beta = next / rz;
    rz = next;
    LaunchCgDirection(cgPi_.data(), cgGpi_.data(), cgRho_.data(), cgGrho_.data(), beta, n, s);
    info.iterations = it + 1;
    info.residual = relative(rz);
    residualHistory.push_back(info.residual);
  }
  info.converged = info.residual <= target;
  info.solveIterations = info.iterations;
  // the solution: its force for the export's band solves, its strain
  LaunchCopy(force_.data(), cgPhi_.data(), n, s);
  rhsCurrent_ = false;
  LaunchCopy(response_.data(), cgY_.data(), n, s);
  LaunchStrain(tasksDev_.data(), tasksDev_.data() + (L + 1), L, coupling_.Geometry(),
               dofsMem_.data(), response_.data(), f.unitsMem.data(), amp_.data(),
               f.unitsMem.data() + dmHost_.size(), TideAmplitudes(), coupling_.View(s_.data()),
               s);
  return info;
}

std::vector<double> Viscoelastic3D::SolutionOfTask(int task) {
  RequireOneRealisation("the solutions' export");
  const auto& t = tasks_.at(task);
  if (lastStep_.scheme < 0) throw ArgumentError("no step has been taken");
  // the factors of the last step's own flavour (a first step's elastic
  // start uses another one after it)
  Flavour& f = FlavourOf(lastStep_.scheme, lastStep_.dt);
  hipStream_t s = Stream();
  // the band solves of the last pass's force on every dof, once per force
  // (every task's export then reads the same solves)
  if (!rhsCurrent_) {
    rhs_.Resize(std::max<size_t>(rhsDoubles_, 1));
    LaunchFill(rhs_.data(), rhs_.size(), 0.0, s);
    LaunchExpandForce(tasksDev_.data(), static_cast<int>(tasks_.size()), nmMax_, colsMax_,
                      dm_.data(), f.scale.data(), force_.data(), rhs_.data(), s);
    LaunchCholeskySolve(f.choleskyDev.data(), f.choleskyTasks.data(), f.choleskyTaskCount,
                        f.kdMax, f.factors.data(), rhs_.data(), s);
    int kvMax = 0;
    for (const auto& b : f.lu) kvMax = std::max(kvMax, b.kl + b.ku);
    LaunchLuSolve(f.luDev.data(), f.luTasks.data(), f.luTaskCount, kvMax, f.factors.data(),
                  f.pivots.data(), rhs_.data(), s);
    rhsCurrent_ = true;
  }
  std::vector<double> raw(static_cast<size_t>(t.ndof) * t.cols), scale(t.ndof),
      units(t.kind == 0 ? 2 * t.ndof : 0), amp(2 * (t.l + 1)), tideAmp(2 * (t.l + 1), 0.0);
  SL3_HIP(hipMemcpyAsync(raw.data(), rhs_.data() + t.rhsOff, sizeof(double) * raw.size(),
                         hipMemcpyDeviceToHost, s));
  SL3_HIP(hipMemcpyAsync(scale.data(), f.scale.data() + t.scaleOff,
                         sizeof(double) * t.ndof, hipMemcpyDeviceToHost, s));
  if (t.kind == 0) {
    SL3_HIP(hipMemcpyAsync(units.data(), f.units.data() + t.unitOff,
                           sizeof(double) * units.size(), hipMemcpyDeviceToHost, s));
    SL3_HIP(hipMemcpyAsync(amp.data(),
                           amp_.data() + static_cast<size_t>(t.l) * (t.l + 1),
                           sizeof(double) * amp.size(), hipMemcpyDeviceToHost, s));
    if (hasTide_) {
      SL3_HIP(hipMemcpyAsync(tideAmp.data(),
                             TideAmplitudes() + static_cast<size_t>(t.l) * (t.l + 1),
                             sizeof(double) * tideAmp.size(), hipMemcpyDeviceToHost, s));
    }
  }
  SL3_HIP(hipStreamSynchronize(s));
  // the reference's Solutions layout: ndof x (l + 1) complex, column-major
  std::vector<double> out(static_cast<size_t>(t.ndof) * (t.l + 1) * 2);
  for (int m = 0; m <= t.l; ++m) {
    for (int d = 0; d < t.ndof; ++d) {
      double re = raw[static_cast<size_t>(d) * t.cols + m] * scale[d];
      double im = raw[static_cast<size_t>(d) * t.cols + t.l + 1 + m] * scale[d];
      if (t.kind == 0) {
        re += amp[2 * m] * units[d] + tideAmp[2 * m] * units[t.ndof + d];
        im += amp[2 * m + 1] * units[d] + tideAmp[2 * m + 1] * units[t.ndof + d];
      }
      out[2 * (static_cast<size_t>(m) * t.ndof + d)] = re;
      out[2 * (static_cast<size_t>(m) * t.ndof + d) + 1] = im;
    }
  }
  return out;
}

}  // namespace sl3amd
