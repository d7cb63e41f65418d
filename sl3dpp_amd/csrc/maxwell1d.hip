// The 1D Maxwell evolution in the degree basis on the GPU (maxwell1d.hpp).
#include "maxwell1d.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "coupling.hpp"

namespace sl3amd {

namespace {

constexpr int kThreads = 256;

struct Geometry1D {
  int nodes, np, ngll, elements;  // memory nodes, points per element, GLL nodes, elements
  int g0, globals;                // first memory global node, memory global nodes
  const double* radius;           // per memory node
  const double* jac;              // per memory element
  const double* weight;           // per point
  const double* value;            // (np, ngll)
  const double* deriv;            // (np, ngll)
  const double* mu;               // per memory node: the branch's modulus
};

// The step's per-node weights: m~ = e m + alpha d0 (StartStep's decay of
// the memory), m+ = m~ + beta d.
struct Weights1D {
  const double* e;
  const double* alpha;
  const double* beta;
};

// The load channel: pattern (complex per (l, m) at l(l+1)/2 + m) times the
// ramp at step counter + offset, times each degree's unit response; with
// the flavour's scaling of the solved right-hand sides.
struct Channel1D {
  const double* pattern;
  const double* ramp;
  const int* counter;
  const double* units;
  const double* scale;
};

// The deviatoric scalars (M, R, S) of basis function i's U and V at a
// memory point: M = V, R = r U' - U + k2 V / 2, S = U + r V' - V.
__device__ inline void ScalarRows(const Geometry1D& g, int p, int elem, int i, double k2,
                                  double r, double rowU[3], double rowV[3]) {
  const double v = g.value[p * g.ngll + i];
  const double b = r * g.deriv[p * g.ngll + i] / g.jac[elem];
  rowU[0] = 0.0;
  rowU[1] = b - v;
  rowU[2] = v;
  rowV[0] = v;
  rowV[1] = 0.5 * k2 * v;
  rowV[2] = b - v;
}

__device__ inline double ChannelAmplitude(const Channel1D& c, const Degree1D& t, int col,
                                          int offset) {
  const int orders = t.l + 1;
  const int order = col < orders ? col : col - orders;
  const double* z = c.pattern + 2 * (static_cast<long long>(t.l) * (t.l + 1) / 2 + order);
  return (col < orders ? z[0] : z[1]) * c.ramp[c.counter[0] + offset];
}

// The right-hand side of every dof of every degree, scaled for the solve
// (BeginSolve): the memory force of mu m~ on U and V at the memory's global
// nodes (m~ = e m + alpha d formed here, the memory itself unchanged), zero
// elsewhere. One thread per (global mesh node, column); a node shared by
// two memory elements sums the lower element's points, then the upper's.
__global__ void __launch_bounds__(kThreads)
    MemoryForceKernel(const Degree1D* degrees, Geometry1D g, int meshNodes, Weights1D wts,
                      const int* __restrict__ dofs, const double* __restrict__ m,
                      const double* __restrict__ d, const double* __restrict__ scale,
                      double* rhs) {
  const Degree1D t = degrees[blockIdx.y];
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(meshNodes) * t.cols) return;
  const int node = static_cast<int>(idx / t.cols), col = static_cast<int>(idx % t.cols);
  const int* dn = dofs + t.dofOff + 3LL * node;
  const int dU = dn[0], dV = dn[1], dP = dn[2];
  double fU = 0.0, fV = 0.0;
  const int gl = node - g.g0;
  if (gl >= 0 && gl < g.globals) {
    const double k2 = t.l * (t.l + 1.0);
    const double D[3] = {k2 * (k2 - 2.0), 4.0 / 3.0, k2};
    const int n1 = g.ngll - 1;
    const int eLo = gl % n1 == 0 ? gl / n1 - 1 : gl / n1;
    const int eHi = gl / n1;
    for (int elem = max(eLo, 0); elem <= min(eHi, g.elements - 1); ++elem) {
      const int i = gl - elem * n1;
      for (int p = 0; p < g.np; ++p) {
        const int mn = elem * g.np + p;
        if (mn < t.firstNode) continue;
        const double w = g.jac[elem] * g.weight[p] * g.mu[mn];
        double rowU[3], rowV[3];
        ScalarRows(g, p, elem, i, k2, g.radius[mn], rowU, rowV);
        const double e = wts.e[mn], a = wts.alpha[mn];
        const long long at = t.memOff + static_cast<long long>(mn) * 3 * t.cols + col;
        double fu = 0.0, fv = 0.0;
        for (int s = 0; s < 3; ++s) {
          const long long x = at + static_cast<long long>(s) * t.cols;
          const double mt = e * m[x] + a * d[x];
          fu += rowU[s] * D[s] * mt;
          fv += rowV[s] * D[s] * mt;
        }
        fU += w * fu;
        fV += w * fv;
      }
    }
  }
  const long long row = t.rhsOff + col;
  if (dU >= 0) rhs[row + static_cast<long long>(dU) * t.cols] = scale[t.unitOff + dU] * fU;
  if (dV >= 0) rhs[row + static_cast<long long>(dV) * t.cols] = scale[t.unitOff + dV] * fV;
  if (dP >= 0) rhs[row + static_cast<long long>(dP) * t.cols] = 0.0;
}

// The end of a step: the solution on the element's dofs (unscaled, plus the
// load channel at the step's end), its deviatoric scalars d at the memory
// point, m+ = (e m + alpha d0) + beta d; d kept for the next step. The
// elastic start (no solve: the channel alone) passes useSolution = 0.
__global__ void __launch_bounds__(kThreads)
    MemoryUpdateKernel(const Degree1D* degrees, Geometry1D g, Weights1D wts, Channel1D ch,
                       int rampOffset, const int* __restrict__ dofs,
                       const double* __restrict__ rhs, double* m, double* d, int useSolution) {
  const Degree1D t = degrees[blockIdx.y];
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long long>(g.nodes) * t.cols) return;
  const int mn = static_cast<int>(idx / t.cols), col = static_cast<int>(idx % t.cols);
  if (mn < t.firstNode) return;
  const int elem = mn / g.np, p = mn % g.np;
  const double k2 = t.l * (t.l + 1.0);
  const double amp = ChannelAmplitude(ch, t, col, rampOffset);
  const double radius = g.radius[mn];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < g.ngll; ++i) {
    const int* dn = dofs + t.dofOff + 3LL * (g.g0 + elem * (g.ngll - 1) + i);
    double UV[2];
    for (int c = 0; c < 2; ++c) {
      const int dof = dn[c];
      double x = 0.0;
      if (dof >= 0) {
        const double solved =
            useSolution ? ch.scale[t.unitOff + dof] *
                              rhs[t.rhsOff + static_cast<long long>(dof) * t.cols + col]
                        : 0.0;
        x = solved + amp * ch.units[t.unitOff + dof];
      }
      UV[c] = x;
    }
    double rowU[3], rowV[3];
    ScalarRows(g, p, elem, i, k2, radius, rowU, rowV);
    for (int s = 0; s < 3; ++s) acc[s] += rowU[s] * UV[0] + rowV[s] * UV[1];
  }
  const double e = wts.e[mn], a = wts.alpha[mn], b = wts.beta[mn];
  for (int s = 0; s < 3; ++s) {
    const long long at = t.memOff + (static_cast<long long>(mn) * 3 + s) * t.cols + col;
    const double mt = e * m[at] + a * d[at];
    d[at] = acc[s];
    m[at] = mt + b * acc[s];
  }
}

// One component's surface value per (degree, column): the unscaled solution
// (none at the elastic start) plus the channel, at out[l (l + 1) + col].
__global__ void SurfaceKernel(const Degree1D* degrees, const int* __restrict__ surfaceDof,
                              Channel1D ch, const double* __restrict__ rhs, int useSolution,
                              double* out) {
  const Degree1D t = degrees[blockIdx.y];
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= t.cols) return;
  const int dof = surfaceDof[blockIdx.y];
  double v = 0.0;
  if (dof >= 0) {
    const double solved =
        useSolution
            ? ch.scale[t.unitOff + dof] * rhs[t.rhsOff + static_cast<long long>(dof) * t.cols + col]
            : 0.0;
    v = solved + ChannelAmplitude(ch, t, col, 0) * ch.units[t.unitOff + dof];
  }
  out[static_cast<long long>(t.l) * (t.l + 1) + col] = v;
}

__global__ void AdvanceKernel(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += 1;
}

dim3 GridOver(long long work, int degrees) {
  return dim3(static_cast<unsigned>(std::max(CeilDiv(work, kThreads), 1)), degrees);
}

}  // namespace

Maxwell1D::Maxwell1D(std::shared_ptr<const Earth> earth, int lmax, double eps)
    : earth_(std::move(earth)), lmax_(lmax), eps_(eps) {
  if (earth_ == nullptr) throw ArgumentError("no Earth");
  if (lmax < 0) throw ArgumentError("lmax must be non-negative");
  if (!(eps > 0.0 && eps < 1.0)) throw ArgumentError("eps must lie in (0, 1)");
  const Earth& e = *earth_;
  const RadialMesh& mesh = *e.mesh;
  const auto& pts = e.reference.points;
  ngll_ = mesh.NGll();
  np_ = pts.count;
  meshNodes_ = mesh.NodeCount();
  // the memory elements: solid, with an active branch, contiguous
  // (src/GridCoupling.cpp:35-52 selects the same ones)
  elem0_ = -1;
  for (int k = 0; k < mesh.ElementCount(); ++k) {
    if (!e.material.IsFluid(k) && e.rheology.ElementActive(mesh, k)) {
      if (elem0_ < 0) elem0_ = k;
      if (k != elem0_ + elements_) throw ArgumentError("the relaxing elements must be contiguous");
      ++elements_;
    }
  }
  if (elements_ == 0) throw ArgumentError("no relaxing element: nothing to evolve");
  nodes_ = elements_ * np_;
  std::vector<double> radius(nodes_), jac(elements_), mu(nodes_);
  itauRef_.resize(nodes_);
  for (int k = 0; k < elements_; ++k) {
    const int el = elem0_ + k;
    jac[k] = mesh.Jacobian(el);
    for (int p = 0; p < np_; ++p) {
      const int node = k * np_ + p;
      radius[node] = pts.RadiusAt(mesh, el, p);
      mu[node] = e.reference.pointMu[static_cast<size_t>(el) * np_ + p];
      itauRef_[node] = e.reference.pointItau[static_cast<size_t>(el) * np_ + p];
    }
  }
  hipStream_t s = Stream();
  radius_.Upload(radius.data(), radius.size(), s);
  jac_.Upload(jac.data(), jac.size(), s);
  mu_.Upload(mu.data(), mu.size(), s);
  pweight_.Upload(pts.weight.data(), pts.weight.size(), s);
  pvalue_.Upload(pts.value.data(), pts.value.size(), s);
  pderiv_.Upload(pts.derivative.data(), pts.derivative.size(), s);
  // the degrees: dof maps and sizes (the point shear changes neither)
  const PointShear elastic{&e.reference.points, &e.reference.elasticShear};
  std::vector<std::unique_ptr<DegreeOperator>> ops(lmax_ + 1);
  ParallelFor(lmax_ + 1, [&](int l) {
    ops[l] = std::make_unique<DegreeOperator>(e.reference.nodal, l, eps_, &elastic);
  });
  std::vector<int> dofs;
  dofs.reserve(static_cast<size_t>(lmax_ + 1) * meshNodes_ * 3);
  for (int l = 0; l <= lmax_; ++l) {
    const DegreeOperator& op = *ops[l];
    Degree1D d;
    d.l = l;
    d.ndof = op.DofCount();
    d.cols = 2 * (l + 1);
    const int fe = op.FirstElement();
    d.firstNode = fe <= elem0_ ? 0 : std::min(nodes_, (fe - elem0_) * np_);
    d.memOff = static_cast<long long>(memDoubles_);
    memDoubles_ += static_cast<size_t>(nodes_) * 3 * d.cols;
    d.rhsOff = static_cast<long long>(rhsDoubles_);
    rhsDoubles_ += static_cast<size_t>(d.ndof) * d.cols;
    d.unitOff = static_cast<long long>(unitDoubles_);
    unitDoubles_ += d.ndof;
    d.dofOff = static_cast<long long>(dofs.size());
    for (int node = 0; node < meshNodes_; ++node) {
      dofs.push_back(op.Dof(node, Component::U));
      dofs.push_back(op.Dof(node, Component::V));
      dofs.push_back(op.Dof(node, Component::Phi));
    }
    for (int c = 0; c < 3; ++c) {
      surfaceDof_[c].push_back(op.SurfaceDof(static_cast<Component>(c)));
    }
    colsMax_ = std::max(colsMax_, d.cols);
    degrees_.push_back(d);
  }
  degreesDev_.Upload(degrees_.data(), degrees_.size(), s);
  dofs_.Upload(dofs.data(), dofs.size(), s);
  for (int c = 0; c < 3; ++c) {
    surfaceDofDev_[c].Upload(surfaceDof_[c].data(), surfaceDof_[c].size(), s);
  }
  m_.Resize(memDoubles_);
  d_.Resize(memDoubles_);
  rhs_.Resize(std::max<size_t>(rhsDoubles_, 1));
  counter_.Resize(1);
  Reset();
}

Maxwell1D::~Maxwell1D() { DropGraph(); }

void Maxwell1D::DropGraph() {
  if (graphExec_ != nullptr) (void)hipGraphExecDestroy(graphExec_);
  if (graph_ != nullptr) (void)hipGraphDestroy(graph_);
  graphExec_ = nullptr;
  graph_ = nullptr;
  graphFlavour_ = nullptr;
}

void Maxwell1D::Reset() {
  hipStream_t s = Stream();
  LaunchFill(m_.data(), memDoubles_, 0.0, s);
  LaunchFill(d_.data(), memDoubles_, 0.0, s);
  LaunchFill(rhs_.data(), rhs_.size(), 0.0, s);
  const int zero = 0;
  SL3_HIP(hipMemcpyAsync(counter_.data(), &zero, sizeof(int), hipMemcpyHostToDevice, s));
  SL3_HIP(hipStreamSynchronize(s));
  stepIndex_ = 0;
  t_ = 0.0;
  hasStrain_ = false;
  last_ = nullptr;
}

void Maxwell1D::SetLoad(const std::vector<double>& pattern, const std::vector<double>& ramp) {
  const size_t need = static_cast<size_t>(lmax_ + 1) * (lmax_ + 2);
  if (pattern.size() != need) {
    throw ArgumentError("the load pattern holds (lmax + 1)(lmax + 2) doubles");
  }
  if (ramp.empty()) throw ArgumentError("the ramp holds the load's factor at every step time");
  DropGraph();  // a captured step holds the old buffers
  hipStream_t s = Stream();
  pattern_.Upload(pattern.data(), pattern.size(), s);
  ramp_.Upload(ramp.data(), ramp.size(), s);
  SL3_HIP(hipStreamSynchronize(s));
}

const Flavour1D& Maxwell1D::FlavourOf(int scheme, double dt) {
  if (dt == 0.0) scheme = kBackwardEuler;  // every scheme's zero-length step
  for (auto& f : flavours_) {
    if (f->scheme == scheme && f->dt == dt) return *f;
  }
  const Earth& e = *earth_;
  auto f = std::make_unique<Flavour1D>();
  f->scheme = scheme;
  f->dt = dt;
  const std::vector<double> shear = e.reference.PointShearModulus(scheme, dt);
  const PointShear ps{&e.reference.points, &shear};
  const int L = lmax_;
  std::vector<std::unique_ptr<DegreeOperator>> ops(L + 1);
  ParallelFor(L + 1, [&](int l) {
    ops[l] = std::make_unique<DegreeOperator>(e.reference.nodal, l, eps_, &ps);
  });
  hipStream_t s = Stream();
  // DegreeOperator::Factorise (src/DegreeOperator.cpp:775-852): Cholesky of
  // the scaled band where the diagonal is positive, LU of the full band
  // where it is not or the Cholesky fails
  std::vector<double> scale(unitDoubles_), bands;
  std::vector<BandSystem> chol;
  std::vector<int> cholDegree;
  for (int l = 0; l <= L; ++l) {
    const Degree1D& d = degrees_[l];
    std::copy(ops[l]->Scale().begin(), ops[l]->Scale().end(), scale.begin() + d.unitOff);
    if (!ops[l]->PositiveDiagonal()) continue;
    const auto band = ops[l]->ScaledUpperBand();
    BandSystem b;
    b.n = d.ndof;
    b.kl = ops[l]->Kd();
    b.nrhs = d.cols;
    b.ab_off = static_cast<long long>(bands.size());
    b.b_off = d.rhsOff;
    b.row_stride = d.cols;
    b.col_stride = 1;
    bands.insert(bands.end(), band.begin(), band.end());
    chol.push_back(b);
    cholDegree.push_back(l);
    f->kdMax = std::max(f->kdMax, b.kl);
  }
  std::vector<double> factors;
  std::vector<char> done(L + 1, 0);
  if (!chol.empty()) {
    DeviceArray<double> cholBands;
    cholBands.Upload(bands.data(), bands.size(), s);
    DeviceArray<BandSystem> cholDev;
    cholDev.Upload(chol.data(), chol.size(), s);
    DeviceArray<int> info(chol.size());
    LaunchCholeskyFactor(cholDev.data(), static_cast<int>(chol.size()), f->kdMax,
                         cholBands.data(), info.data(), s);
    std::vector<int> infoHost(chol.size());
    SL3_HIP(hipMemcpyAsync(infoHost.data(), info.data(), sizeof(int) * chol.size(),
                           hipMemcpyDeviceToHost, s));
    SL3_HIP(hipMemcpyAsync(bands.data(), cholBands.data(), sizeof(double) * bands.size(),
                           hipMemcpyDeviceToHost, s));
    SL3_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < chol.size(); ++k) {
      if (infoHost[k] != 0) continue;
      BandSystem b = chol[k];
      b.ab_off = static_cast<long long>(factors.size());
      factors.insert(factors.end(), bands.begin() + chol[k].ab_off,
                     bands.begin() + chol[k].ab_off + static_cast<long long>(b.kl + 1) * b.n);
      f->cholesky.push_back(b);
      done[cholDegree[k]] = 1;
    }
  }
  std::vector<double> luBands;
  std::vector<int> luDegree;
  long long pivots = 0;
  int klMax = 0;
  for (int l = 0; l <= L; ++l) {
    if (done[l]) continue;
    const Degree1D& d = degrees_[l];
    const auto band = ops[l]->ScaledLuBand();
    BandSystem b;
    b.n = d.ndof;
    b.kl = b.ku = ops[l]->Kd();
    b.nrhs = d.cols;
    b.ab_off = static_cast<long long>(luBands.size());
    b.piv_off = pivots;
    pivots += d.ndof;
    b.b_off = d.rhsOff;
    b.row_stride = d.cols;
    b.col_stride = 1;
    luBands.insert(luBands.end(), band.begin(), band.end());
    f->lu.push_back(b);
    luDegree.push_back(l);
    klMax = std::max(klMax, b.kl);
    f->kvMax = std::max(f->kvMax, b.kl + b.ku);
  }
  const long long luBase = static_cast<long long>(factors.size());
  for (auto& b : f->lu) b.ab_off += luBase;
  factors.insert(factors.end(), luBands.begin(), luBands.end());
  f->factors.Upload(factors.data(), std::max<size_t>(factors.size(), 1), s);
  f->pivots.Resize(std::max<long long>(pivots, 1));
  f->luDev.Upload(f->lu.data(), f->lu.size(), s);
  if (!f->lu.empty()) {
    DeviceArray<int> luInfo(f->lu.size());
    LaunchLuFactor(f->luDev.data(), static_cast<int>(f->lu.size()), klMax, klMax,
                   f->factors.data(), f->pivots.data(), luInfo.data(), s);
    std::vector<int> li(f->lu.size());
    SL3_HIP(hipMemcpyAsync(li.data(), luInfo.data(), sizeof(int) * li.size(),
                           hipMemcpyDeviceToHost, s));
    SL3_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < li.size(); ++k) {
      if (li[k] != 0) {
        throw NumericalError("degree " + std::to_string(luDegree[k]) +
                             ": the banded LU factorisation failed (info = " +
                             std::to_string(li[k]) + ")");
      }
    }
  }
  f->scale.Upload(scale.data(), scale.size(), s);
  f->choleskyDev.Upload(f->cholesky.data(), f->cholesky.size(), s);
  const auto ct = SolveTasks(f->cholesky);
  const auto lt = SolveTasks(f->lu);
  f->choleskyTasks.Upload(ct.data(), ct.size(), s);
  f->luTasks.Upload(lt.data(), lt.size(), s);
  f->choleskyTaskCount = static_cast<int>(ct.size());
  f->luTaskCount = static_cast<int>(lt.size());
  // the unit load responses (the load channel's unit solve): Load() scaled,
  // solved with this flavour's factors, unscaled
  {
    std::vector<double> unit(unitDoubles_);
    for (int l = 0; l <= L; ++l) {
      const auto load = ops[l]->Load();
      const Degree1D& d = degrees_[l];
      for (int i = 0; i < d.ndof; ++i) unit[d.unitOff + i] = load[i] * scale[d.unitOff + i];
    }
    f->units.Upload(unit.data(), unit.size(), s);
    auto single = [&](std::vector<BandSystem> sys) {
      for (auto& b : sys) {
        const auto it = std::find_if(degrees_.begin(), degrees_.end(),
                                     [&](const Degree1D& d) { return d.rhsOff == b.b_off; });
        b.nrhs = 1;
        b.b_off = it->unitOff;
        b.row_stride = 1;
        b.col_stride = b.n;
      }
      return sys;
    };
    const auto cu = single(f->cholesky), lu = single(f->lu);
    DeviceArray<BandSystem> cuDev, luDev;
    cuDev.Upload(cu.data(), cu.size(), s);
    luDev.Upload(lu.data(), lu.size(), s);
    const auto cut = SolveTasks(cu), lut = SolveTasks(lu);
    DeviceArray<int2> cutDev, lutDev;
    cutDev.Upload(cut.data(), cut.size(), s);
    lutDev.Upload(lut.data(), lut.size(), s);
    LaunchCholeskySolve(cuDev.data(), cutDev.data(), static_cast<int>(cut.size()), f->kdMax,
                        f->factors.data(), f->units.data(), s);
    LaunchLuSolve(luDev.data(), lutDev.data(), static_cast<int>(lut.size()), f->kvMax,
                  f->factors.data(), f->pivots.data(), f->units.data(), s);
    SL3_HIP(hipMemcpyAsync(unit.data(), f->units.data(), sizeof(double) * unit.size(),
                           hipMemcpyDeviceToHost, s));
    SL3_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < unit.size(); ++i) unit[i] *= scale[i];  // FinishSolve
    f->units.Upload(unit.data(), unit.size(), s);
  }
  // the step's band panel, timed on a scratch batch of the step's layout
  {
    DeviceArray<double> scratch(std::max<size_t>(rhsDoubles_, 1));
    LaunchFill(scratch.data(), scratch.size(), 0.0, s);
    f->panelRows = TuneCholeskySolvePanel(f->choleskyDev.data(), f->choleskyTasks.data(),
                                          f->choleskyTaskCount, f->kdMax, f->factors.data(),
                                          scratch.data(), s, f->panelMs);
    const int forced = BandPanelOverride();
    if (forced > 0) f->panelRows = forced;
    SL3_HIP(hipStreamSynchronize(s));  // the scratch is freed next
  }
  // the per-node step weights (include/sl3dpp/StepWeights.hpp:98-144)
  std::vector<double> ew(nodes_), aw(nodes_), bw(nodes_);
  for (int node = 0; node < nodes_; ++node) {
    const double h = dt * itauRef_[node];
    if (scheme == kBackwardEuler) {
      ew[node] = 1.0 / (1.0 + h);
      aw[node] = 0.0;
      bw[node] = 1.0 - ew[node];
    } else {
      const double ex = std::exp(-h);
      const double p = Phi1(h, ex);
      ew[node] = ex;
      aw[node] = p - ex;
      bw[node] = 1.0 - p;
    }
  }
  f->e.Upload(ew.data(), ew.size(), s);
  f->alpha.Upload(aw.data(), aw.size(), s);
  f->beta.Upload(bw.data(), bw.size(), s);
  SL3_HIP(hipStreamSynchronize(s));
  flavours_.push_back(std::move(f));
  return *flavours_.back();
}

namespace {

Geometry1D GeometryOf(int nodes, int np, int ngll, int elem0, int elements,
                      const DeviceArray<double>& radius, const DeviceArray<double>& jac,
                      const DeviceArray<double>& weight, const DeviceArray<double>& value,
                      const DeviceArray<double>& deriv, const DeviceArray<double>& mu) {
  return Geometry1D{nodes,         np,           ngll,         elements,
                    elem0 * (ngll - 1), elements * (ngll - 1) + 1, radius.data(),
                    jac.data(),    weight.data(), value.data(), deriv.data(),
                    mu.data()};
}

}  // namespace

void Maxwell1D::ElasticStart(const Flavour1D& f, hipStream_t s) {
  const Geometry1D g = GeometryOf(nodes_, np_, ngll_, elem0_, elements_, radius_, jac_, pweight_,
                                  pvalue_, pderiv_, mu_);
  const Weights1D w{f.e.data(), f.alpha.data(), f.beta.data()};
  const Channel1D ch{pattern_.data(), ramp_.data(), counter_.data(), f.units.data(),
                     f.scale.data()};
  ProfileScope prof("maxwell_update", s);
  hipLaunchKernelGGL(MemoryUpdateKernel,
                     GridOver(static_cast<long long>(nodes_) * colsMax_, lmax_ + 1),
                     dim3(kThreads), 0, s, degreesDev_.data(), g, w, ch, 0, dofs_.data(),
                     rhs_.data(), m_.data(), d_.data(), 0);
  SL3_HIP(hipGetLastError());
}

void Maxwell1D::EnqueueStep(const Flavour1D& f, hipStream_t s) {
  const Geometry1D g = GeometryOf(nodes_, np_, ngll_, elem0_, elements_, radius_, jac_, pweight_,
                                  pvalue_, pderiv_, mu_);
  const Weights1D w{f.e.data(), f.alpha.data(), f.beta.data()};
  const Channel1D ch{pattern_.data(), ramp_.data(), counter_.data(), f.units.data(),
                     f.scale.data()};
  const int degrees = lmax_ + 1;
  {
    ProfileScope prof("maxwell_force", s);
    hipLaunchKernelGGL(MemoryForceKernel,
                       GridOver(static_cast<long long>(meshNodes_) * colsMax_, degrees),
                       dim3(kThreads), 0, s, degreesDev_.data(), g, meshNodes_, w, dofs_.data(),
                       m_.data(), d_.data(), f.scale.data(), rhs_.data());
    SL3_HIP(hipGetLastError());
  }
  LaunchCholeskySolve(f.choleskyDev.data(), f.choleskyTasks.data(), f.choleskyTaskCount,
                      f.kdMax, f.factors.data(), rhs_.data(), s, f.panelRows);
  LaunchLuSolve(f.luDev.data(), f.luTasks.data(), f.luTaskCount, f.kvMax, f.factors.data(),
                f.pivots.data(), rhs_.data(), s);
  {
    ProfileScope prof("maxwell_update", s);
    hipLaunchKernelGGL(MemoryUpdateKernel,
                       GridOver(static_cast<long long>(nodes_) * colsMax_, degrees),
                       dim3(kThreads), 0, s, degreesDev_.data(), g, w, ch, 1, dofs_.data(),
                       rhs_.data(), m_.data(), d_.data(), 1);
    SL3_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(AdvanceKernel, dim3(1), dim3(64), 0, s, counter_.data());
  SL3_HIP(hipGetLastError());
}

void Maxwell1D::Evolve(double dt, int scheme, int steps, bool graph) {
  if (!(dt > 0.0)) throw ArgumentError("the step length must be positive");
  if (scheme != kExponentialTrapezoid && scheme != kBackwardEuler) {
    throw ArgumentError("the 1D evolution takes the exponential trapezoid or backward Euler");
  }
  if (steps < 0) throw ArgumentError("steps must be non-negative");
  if (pattern_.size() == 0) throw ArgumentError("no load set");
  if (static_cast<size_t>(stepIndex_) + static_cast<size_t>(steps) >= ramp_.size()) {
    throw ArgumentError("the ramp holds fewer step times than the steps asked for");
  }
  hipStream_t s = Stream();
  if (!hasStrain_) {
    // the first step from rest: the elastic start under the load at t0
    // (a zero-length step: the unrelaxed systems, no memory)
    const Flavour1D& elastic = FlavourOf(kBackwardEuler, 0.0);
    ElasticStart(elastic, s);
    SL3_HIP(hipStreamSynchronize(s));
    hasStrain_ = true;
    last_ = &elastic;
  }
  const Flavour1D& f = FlavourOf(scheme, dt);
  tuning_.assign(1, f.panelRows);
  tuning_.insert(tuning_.end(), f.panelMs, f.panelMs + kPanelChoices);
  if (steps == 0) return;
  if (graph) {
    if (graphFlavour_ != &f) {
      DropGraph();
      SL3_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      EnqueueStep(f, s);
      SL3_HIP(hipStreamEndCapture(s, &graph_));
      SL3_HIP(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0));
      graphFlavour_ = &f;
    }
    for (int k = 0; k < steps; ++k) SL3_HIP(hipGraphLaunch(graphExec_, s));
  } else {
    for (int k = 0; k < steps; ++k) EnqueueStep(f, s);
  }
  SL3_HIP(hipStreamSynchronize(s));
  stepIndex_ += steps;
  t_ += steps * dt;
  last_ = &f;
}

std::vector<double> Maxwell1D::Memory() const {
  std::vector<double> out(memDoubles_);
  SL3_HIP(hipMemcpy(out.data(), m_.data(), sizeof(double) * out.size(), hipMemcpyDeviceToHost));
  return out;
}

std::vector<double> Maxwell1D::Strain() const {
  std::vector<double> out(memDoubles_);
  SL3_HIP(hipMemcpy(out.data(), d_.data(), sizeof(double) * out.size(), hipMemcpyDeviceToHost));
  return out;
}

std::vector<double> Maxwell1D::Solutions() const {
  if (last_ == nullptr) throw ArgumentError("no step has been taken");
  std::vector<double> y(rhsDoubles_), scale(unitDoubles_), units(unitDoubles_);
  std::vector<double> pattern(pattern_.size()), ramp(ramp_.size());
  SL3_HIP(hipMemcpy(y.data(), rhs_.data(), sizeof(double) * y.size(), hipMemcpyDeviceToHost));
  SL3_HIP(hipMemcpy(scale.data(), last_->scale.data(), sizeof(double) * scale.size(),
                    hipMemcpyDeviceToHost));
  SL3_HIP(hipMemcpy(units.data(), last_->units.data(), sizeof(double) * units.size(),
                    hipMemcpyDeviceToHost));
  SL3_HIP(hipMemcpy(pattern.data(), pattern_.data(), sizeof(double) * pattern.size(),
                    hipMemcpyDeviceToHost));
  SL3_HIP(hipMemcpy(ramp.data(), ramp_.data(), sizeof(double) * ramp.size(),
                    hipMemcpyDeviceToHost));
  const double r = ramp[stepIndex_];
  const bool solved = last_->dt > 0.0;  // the elastic start solves nothing
  for (const auto& d : degrees_) {
    for (int c = 0; c < d.cols; ++c) {
      const int order = c <= d.l ? c : c - d.l - 1;
      const double* z = pattern.data() + 2 * (static_cast<size_t>(d.l) * (d.l + 1) / 2 + order);
      const double amp = (c <= d.l ? z[0] : z[1]) * r;
      for (int i = 0; i < d.ndof; ++i) {
        double& v = y[d.rhsOff + static_cast<size_t>(i) * d.cols + c];
        v = (solved ? scale[d.unitOff + i] * v : 0.0) + amp * units[d.unitOff + i];
      }
    }
  }
  return y;
}

void Maxwell1D::Surface(int component, double* out_dev, hipStream_t s) const {
  if (component < 0 || component > 2) throw ArgumentError("component: 0 U, 1 V, 2 phi");
  if (last_ == nullptr) throw ArgumentError("no step has been taken");
  const Channel1D ch{pattern_.data(), ramp_.data(), counter_.data(), last_->units.data(),
                     last_->scale.data()};
  hipLaunchKernelGGL(SurfaceKernel, dim3(CeilDiv(colsMax_, 64), lmax_ + 1), dim3(64), 0, s,
                     degreesDev_.data(), surfaceDofDev_[component].data(), ch, rhs_.data(),
                     last_->dt > 0.0 ? 1 : 0, out_dev);
  SL3_HIP(hipGetLastError());
}

}  // namespace sl3amd
