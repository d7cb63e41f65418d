// Host set-up: mesh, material, rheology, memory points, 1D reference point
// moduli (see model.hpp).
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "common.hpp"

namespace sl3amd {

namespace {
constexpr double kPi = 3.14159265358979323846;

double Lerp(double a, double b, double t) { return a + t * (b - a); }
}  // namespace

double LegendreP(int n, double x) {
  if (n == 0) return 1.0;
  double p0 = 1.0, p1 = x;
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

void GaussLegendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      const double pn = LegendreP(n, z), pm = LegendreP(n - 1, z);
      const double dp = n * (z * pn - pm) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (std::abs(dz) < 1e-16) break;
    }
    const double pn = LegendreP(n, z), pm = LegendreP(n - 1, z);
    const double dp = n * (z * pn - pm) / (z * z - 1.0);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

void GaussLobattoLegendre(int n, std::vector<double>& x,
                          std::vector<double>& w) {
  // Nodes: +-1 and the roots of P'_N, N = n - 1, by Newton on
  // x P_N - P_{N-1} from the Chebyshev-Lobatto points.
  const int N = n - 1;
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = -std::cos(kPi * i / N);
    if (i > 0 && i < N) {
      for (int it = 0; it < 100; ++it) {
        const double pn = LegendreP(N, z), pm = LegendreP(N - 1, z);
        const double dz = (z * pn - pm) / (n * pn);
        z -= dz;
        if (std::abs(dz) < 1e-16) break;
      }
    }
    x[i] = z;
    const double pn = LegendreP(N, z);
    w[i] = 2.0 / (N * (N + 1.0) * pn * pn);
  }
}

RadialMesh::RadialMesh(const std::vector<double>& breakpoints,
                       const std::vector<int>& elementsPerSpan, int ngll)
    : ngll_(ngll) {
  if (ngll < 2) throw ArgumentError("ngll must be at least 2");
  for (size_t s = 0; s + 1 < breakpoints.size(); ++s) {
    const double lo = breakpoints[s], hi = breakpoints[s + 1];
    const int nel = elementsPerSpan[s];
    if (nel < 1 || !(hi > lo)) throw ArgumentError("bad mesh span");
    for (int e = 0; e < nel; ++e) {
      const double a = Lerp(lo, hi, static_cast<double>(e) / nel);
      const double b =
          e + 1 == nel ? hi : Lerp(lo, hi, static_cast<double>(e + 1) / nel);
      left_.push_back(a);
      right_.push_back(b);
      jac_.push_back(0.5 * (b - a));
      layer_.push_back(static_cast<int>(s));
    }
  }
  GaussLobattoLegendre(ngll, xi_, w_);
  const int n = ngll, N = n - 1;
  std::vector<double> P(n);
  for (int i = 0; i < n; ++i) P[i] = LegendreP(N, xi_[i]);
  deriv_.assign(static_cast<size_t>(n) * n, 0.0);
  for (int k = 0; k < n; ++k) {
    for (int i = 0; i < n; ++i) {
      if (k != i) deriv_[k * n + i] = P[k] / (P[i] * (xi_[k] - xi_[i]));
    }
  }
  deriv_[0] = -0.25 * N * (N + 1);
  deriv_[static_cast<size_t>(n) * n - 1] = 0.25 * N * (N + 1);
  radius_.resize(left_.size() * n);
  for (size_t e = 0; e < left_.size(); ++e) {
    for (int i = 0; i < n; ++i) {
      radius_[e * n + i] = left_[e] + (xi_[i] + 1.0) * jac_[e];
    }
    radius_[e * n] = left_[e];
    radius_[e * n + N] = right_[e];
  }
}

int RadialMesh::ElementOf(double r) const {
  const auto it = std::upper_bound(left_.begin(), left_.end(), r);
  const long i = static_cast<long>(it - left_.begin()) - 1;
  return static_cast<int>(
      std::clamp<long>(i, 0, static_cast<long>(left_.size()) - 1));
}

double RadialMesh::TruncationRadius(int l, double eps) const {
  const double rmax = right_.back();
  const double r = rmax * std::pow(eps, 1.0 / (l + 1));
  return std::clamp(r, left_.front(), rmax);
}

int RadialMesh::StartElement(int l, double eps) const {
  return ElementOf(TruncationRadius(l, eps));
}

Material Material::WithShear(const std::vector<double>& newMu) const {
  Material out = *this;
  out.mu = newMu;
  for (size_t i = 0; i < mu.size(); ++i) {
    out.A[i] = out.C[i] = kappa[i] + 4.0 * newMu[i] / 3.0;
    out.F[i] = kappa[i] - 2.0 * newMu[i] / 3.0;
    out.L[i] = out.N[i] = newMu[i];
  }
  return out;
}

bool Rheology::ElementActive(const RadialMesh& mesh, int e) const {
  for (int i = 0; i < mesh.NGll(); ++i) {
    if (active[e * mesh.NGll() + i]) return true;
  }
  return false;
}

double MemoryPoints::RadiusAt(const RadialMesh& mesh, int e, int p) const {
  double r = 0.0;
  for (int i = 0; i < ngll; ++i) r += Value(p, i) * mesh.Radius(e, i);
  return r;
}

MemoryPoints MemoryPoints::At(const RadialMesh& mesh, const std::vector<double>& x,
                              const std::vector<double>& w) {
  MemoryPoints out;
  out.ngll = mesh.NGll();
  out.count = static_cast<int>(x.size());
  out.xi = x;
  out.weight = w;
  const auto& nodes = mesh.Points();
  const int n = out.ngll;
  out.value.resize(static_cast<size_t>(out.count) * n);
  out.derivative.resize(static_cast<size_t>(out.count) * n);
  for (int p = 0; p < out.count; ++p) {
    for (int i = 0; i < n; ++i) {
      // l_i and l_i' at xi_p by the product rule, one factor at a time
      double value = 1.0, slope = 0.0;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        const double scale = 1.0 / (nodes[i] - nodes[j]);
        slope = slope * (x[p] - nodes[j]) * scale + value * scale;
        value *= (x[p] - nodes[j]) * scale;
      }
      out.value[p * n + i] = value;
      out.derivative[p * n + i] = slope;
    }
  }
  return out;
}

MemoryPoints MemoryPoints::Gauss(const RadialMesh& mesh) {
  std::vector<double> x, w;
  GaussLegendre(mesh.NGll() - 1, x, w);
  // ascending on [-1, 1]
  std::reverse(x.begin(), x.end());
  std::reverse(w.begin(), w.end());
  return At(mesh, x, w);
}

MemoryPoints MemoryPoints::Gll(const RadialMesh& mesh) {
  return At(mesh, mesh.Points(), mesh.Weights());
}

std::vector<double> Reference1D::PointShearModulus(int scheme, double dt) const {
  if (dt == 0.0) return elasticShear;
  std::vector<double> s(pointMu.size(), 0.0);
  for (size_t i = 0; i < s.size(); ++i) {
    const double h = dt * pointItau[i];
    double unrelaxed;
    if (scheme == kBackwardEuler) {
      unrelaxed = 1.0 / (1.0 + h);
    } else {
      const double e = std::exp(-h);
      const double beta = 1.0 - Phi1(h, e);
      unrelaxed = 1.0 - beta;
    }
    s[i] += pointMu[i] * unrelaxed;
  }
  return s;
}

namespace {

// The 1D reference's point quantities and the long-term material, from the
// nodal material and the Maxwell branch.
void BuildReference(Earth& earth) {
  const RadialMesh& mesh = *earth.mesh;
  const int n = mesh.NGll();
  const int E = mesh.ElementCount();
  const size_t size = static_cast<size_t>(E) * n;
  const Material& mat = earth.material;
  const Rheology& rh = earth.rheology;
  Reference1D& ref = earth.reference;
  ref.points = MemoryPoints::Gauss(mesh);
  const int np = ref.points.count;
  auto interpolate = [&](const std::vector<double>& nodal) {
    std::vector<double> out(static_cast<size_t>(E) * np, 0.0);
    for (int e = 0; e < E; ++e) {
      for (int p = 0; p < np; ++p) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
          s += ref.points.Value(p, i) * nodal[static_cast<size_t>(e) * n + i];
        }
        out[static_cast<size_t>(e) * np + p] = std::max(s, 0.0);
      }
    }
    return out;
  };
  ref.pointMu = interpolate(rh.branchMu);
  ref.pointItau = interpolate(rh.branchItau);
  // Rheology::LongTermMaterial (src/Rheology.cpp:460-485): each branch's
  // shear removed through the isotropic perturbation rule.
  ref.nodal = mat;
  ref.nodal.mu = rh.muLongTerm;
  for (size_t i = 0; i < size; ++i) {
    const double delta = (0.0 - 1.0) * rh.branchMu[i];
    ref.nodal.A[i] += (4.0 / 3.0) * delta;
    ref.nodal.C[i] += (4.0 / 3.0) * delta;
    ref.nodal.F[i] -= (2.0 / 3.0) * delta;
    ref.nodal.L[i] += delta;
    ref.nodal.N[i] += delta;
  }
  ref.elasticShear = ref.pointMu;
}

}  // namespace

std::shared_ptr<const Earth> BuildEarth(const EarthSpec& spec) {
  auto earth = std::make_shared<Earth>();
  earth->spec = spec;
  const Scales& sc = earth->scales;
  const double L = sc.length;
  const auto& layers = spec.layers;
  if (layers.empty()) throw ArgumentError("the Earth needs a layer");
  if (layers.front().r0 != 0.0) throw ArgumentError("the first layer starts at r = 0");
  std::vector<double> breaks{0.0};
  std::vector<int> counts;
  for (size_t i = 0; i < layers.size(); ++i) {
    if (i > 0 && layers[i].r0 != layers[i - 1].r1) {
      throw ArgumentError("layers must be contiguous");
    }
    breaks.push_back(layers[i].r1 / L);
    counts.push_back(layers[i].elements);
  }
  earth->mesh = std::make_shared<RadialMesh>(breaks, counts, spec.ngll);
  const auto& mesh = *earth->mesh;
  const int n = mesh.NGll();
  const int E = mesh.ElementCount();
  const size_t size = static_cast<size_t>(E) * n;

  // Non-dimensional linear profiles per layer: value = a + b r.
  struct Linear {
    double a, b;
  };
  auto linear = [&](double r0, double r1, double v0, double v1, double scale) {
    const double x0 = r0 / L, x1 = r1 / L;
    const double b = (v1 - v0) / scale / (x1 - x0);
    return Linear{v0 / scale - b * x0, b};
  };
  std::vector<Linear> rho, kappa, mu;
  std::vector<double> massBelow{0.0};
  for (const auto& ly : layers) {
    rho.push_back(linear(ly.r0, ly.r1, ly.rho0, ly.rho1, sc.density));
    kappa.push_back(linear(ly.r0, ly.r1, ly.kappa0, ly.kappa1, sc.Modulus()));
    mu.push_back(linear(ly.r0, ly.r1, ly.mu0, ly.mu1, sc.Modulus()));
    const auto& d = rho.back();
    const double x0 = ly.r0 / L, x1 = ly.r1 / L;
    massBelow.push_back(massBelow.back() +
                        4.0 * kPi *
                            (d.a * (x1 * x1 * x1 - x0 * x0 * x0) / 3.0 +
                             d.b * (x1 * x1 * x1 * x1 - x0 * x0 * x0 * x0) / 4.0));
  }
  auto gravity = [&](double r) {
    if (r <= 0.0) return 0.0;
    size_t k = 0;
    while (k + 1 < layers.size() && r > layers[k].r1 / L) ++k;
    const auto& d = rho[k];
    const double x0 = layers[k].r0 / L;
    const double m =
        massBelow[k] + 4.0 * kPi *
                           (d.a * (r * r * r - x0 * x0 * x0) / 3.0 +
                            d.b * (r * r * r * r - x0 * x0 * x0 * x0) / 4.0);
    return m / (r * r);
  };

  Material& mat = earth->material;
  mat.mesh = earth->mesh;
  mat.rho.resize(size);
  mat.drho.resize(size);
  mat.g.resize(size);
  mat.kappa.resize(size);
  mat.mu.resize(size);
  mat.fluid.assign(E, 0);
  for (int e = 0; e < E; ++e) {
    const int k = mesh.LayerOf(e);
    mat.fluid[e] = (layers[k].mu0 == 0.0 && layers[k].mu1 == 0.0) ? 1 : 0;
    for (int i = 0; i < n; ++i) {
      const double r = mesh.Radius(e, i);
      const size_t at = static_cast<size_t>(e) * n + i;
      mat.rho[at] = rho[k].a + rho[k].b * r;
      mat.drho[at] = rho[k].b;
      mat.g[at] = gravity(r);
      mat.kappa[at] = kappa[k].a + kappa[k].b * r;
      mat.mu[at] = mu[k].a + mu[k].b * r;
    }
  }
  if (mat.fluid[E - 1]) throw ArgumentError("the surface layer must be solid");
  mat.A.resize(size);
  mat.C.resize(size);
  mat.F.resize(size);
  mat.L.resize(size);
  mat.N.resize(size);
  mat = mat.WithShear(mat.mu);
  mat.radius = mesh.Right(E - 1);
  mat.surfaceGravity = mat.g.back();

  // Maxwell: mu_1 = the Voigt shear (isotropic: mu), itau = mu / eta.
  Rheology& rh = earth->rheology;
  rh.branchMu.assign(size, 0.0);
  rh.branchItau.assign(size, 0.0);
  rh.muLongTerm = mat.mu;
  rh.active.assign(size, 0);
  const double etaScale = sc.Viscosity();
  for (int e = 0; e < E; ++e) {
    if (mesh.LayerOf(e) != spec.viscousLayer) continue;
    if (mat.IsFluid(e)) throw ArgumentError("a fluid layer cannot relax");
    for (int i = 0; i < n; ++i) {
      const size_t at = static_cast<size_t>(e) * n + i;
      const double r = mesh.Radius(e, i);
      const double eta = (r * L > spec.viscosityBoundary ? spec.viscosityUpper
                                                          : spec.viscosityLower) /
                         etaScale;
      rh.branchMu[at] = mat.mu[at];
      rh.branchItau[at] = mat.mu[at] / eta;
      rh.muLongTerm[at] -= rh.branchMu[at];
      if (rh.branchMu[at] > 0.0) rh.active[at] = 1;
    }
  }

  BuildReference(*earth);
  return earth;
}

std::shared_ptr<const Earth> BuildEarth(const NodalEarthSpec& spec) {
  const int n = spec.ngll;
  const int E = static_cast<int>(spec.edges.size()) - 1;
  if (E < 1) throw ArgumentError("at least one element");
  if (n < 2 || n > 16) throw ArgumentError("ngll in [2, 16]");
  const size_t size = static_cast<size_t>(E) * n;
  for (const auto* v : {&spec.rho, &spec.drho, &spec.g, &spec.A, &spec.C, &spec.F, &spec.L, &spec.N}) {
    if (v->size() != size) throw ArgumentError("nodal arrays hold elements x ngll values");
  }
  if (spec.fluid.size() != static_cast<size_t>(E)) throw ArgumentError("one fluid flag per element");
  const bool maxwell = !spec.branchMu.empty();
  if (maxwell && (spec.branchMu.size() != size || spec.branchItau.size() != size)) {
    throw ArgumentError("the Maxwell branch holds elements x ngll values");
  }
  if (!(spec.length > 0.0 && spec.density > 0.0)) throw ArgumentError("positive scales");
  auto earth = std::make_shared<Earth>();
  Scales& sc = earth->scales;
  sc.length = spec.length;
  sc.density = spec.density;
  const double L = sc.length;
  std::vector<double> edges(spec.edges);
  for (auto& r : edges) r /= L;
  earth->mesh = std::make_shared<RadialMesh>(edges, n);
  const RadialMesh& mesh = *earth->mesh;
  Material& mat = earth->material;
  mat.mesh = earth->mesh;
  const double gScale = L / (sc.Time() * sc.Time());
  const double modulus = sc.Modulus();
  mat.rho.resize(size);
  mat.drho.resize(size);
  mat.g.resize(size);
  mat.A.resize(size);
  mat.C.resize(size);
  mat.F.resize(size);
  mat.L.resize(size);
  mat.N.resize(size);
  mat.mu.resize(size);
  mat.kappa.resize(size);
  for (size_t i = 0; i < size; ++i) {
    mat.rho[i] = spec.rho[i] / sc.density;
    mat.drho[i] = spec.drho[i] * L / sc.density;
    mat.g[i] = spec.g[i] / gScale;
    mat.A[i] = spec.A[i] / modulus;
    mat.C[i] = spec.C[i] / modulus;
    mat.F[i] = spec.F[i] / modulus;
    mat.L[i] = spec.L[i] / modulus;
    mat.N[i] = spec.N[i] / modulus;
    // the Voigt shear and bulk moduli of the transversely isotropic ones
    mat.mu[i] = (mat.C[i] + mat.A[i] - 2.0 * mat.F[i] + 5.0 * mat.N[i] + 6.0 * mat.L[i]) / 15.0;
    mat.kappa[i] = (mat.C[i] + 4.0 * mat.A[i] + 4.0 * mat.F[i] - 4.0 * mat.N[i]) / 9.0;
  }
  mat.fluid.assign(spec.fluid.begin(), spec.fluid.end());
  if (mat.fluid[E - 1]) throw ArgumentError("the surface element must be solid");
  mat.radius = mesh.Right(E - 1);
  mat.surfaceGravity = mat.g.back();
  Rheology& rh = earth->rheology;
  rh.branchMu.assign(size, 0.0);
  rh.branchItau.assign(size, 0.0);
  rh.active.assign(size, 0);
  if (maxwell) {
    for (size_t i = 0; i < size; ++i) {
      rh.branchMu[i] = spec.branchMu[i] / modulus;
      rh.branchItau[i] = spec.branchItau[i] * sc.Time();
      if (rh.branchMu[i] > 0.0 && rh.branchItau[i] > 0.0) rh.active[i] = 1;
    }
  }
  for (int e = 0; e < E; ++e) {
    if (mat.IsFluid(e) && rh.ElementActive(mesh, e)) throw ArgumentError("a fluid element cannot relax");
  }
  rh.muLongTerm = mat.mu;
  for (size_t i = 0; i < size; ++i) rh.muLongTerm[i] -= rh.branchMu[i];
  BuildReference(*earth);
  return earth;
}

}  // namespace sl3amd
