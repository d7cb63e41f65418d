// Degree systems, host side (see degree_systems.hpp).
//
// The weak form of the gravitating, self-consistent elastic problem of one
// spherical-harmonic degree l (k2 = l(l + 1)) reduces, at every integration
// point, to a quadratic form in a handful of field "measures":
//   spheroidal: U, V, phi and their r d/dr (six measures);
//   toroidal:   W and r dW/dr (two measures).
// With the transversely isotropic moduli (C, A, F, L, N) the elastic energy
// density is
//   C x^2 + (A - N) y^2 + 2 F x y + k2 L s^2 + k2 (k2 - 2) N V^2,
//   x = r U',  y = 2U - k2 V,  s = U + r V' - V,
// (toroidal: k2 L (r W' - W)^2 + k2 (k2 - 2) N W^2), gravity adds
//   4 rho (pi G rho r - g) r U^2 + 2 k2 rho g r U V + 2 k2 rho r V phi
//   + 2 rho r U (r phi'),
// and the potential's own energy is (r^2 phi'^2 + k2 phi^2) / (4 pi G), with
// r^2 (d rho/dr) / g phi^2 in a fluid (the reference's formulation:
// src/DegreeOperator.cpp:186-474, src/ToroidalOperator.cpp:80-147). The
// point shear s_p at a memory point adds s_p (k2 (k2 - 2) M^2 + 4/3 R^2 +
// k2 S^2) in the deviatoric scalars M = V, R = r U' - U + k2 V / 2,
// S = U + r V' - V (toroidal: s_p (k2 (r W' - W)^2 + k2 (k2 - 2) W^2)).
//
// Each integration point therefore contributes weight * B^T Q B to its
// element's matrix, where B maps the element's nodal values to the
// measures (Lagrange values and r d/dr of the basis at the point) and Q is
// the 6 x 6 (or 2 x 2) energy matrix. The quadrature is the GLL rule at the
// nodes; memory points use the same code with their own basis values.
#include "degree_systems.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "common.hpp"

namespace sl3amd {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxNodes = 16;

// ---- the element basis at one integration point -----------------------------

struct PointBasis {
  double r = 0.0;       // radius of the point
  double weight = 0.0;  // quadrature weight with the element's jacobian
  double value[kMaxNodes] = {};
  double rdr[kMaxNodes] = {};  // r d/dr of each basis function
};

// The GLL node q of element e: basis values are the Kronecker delta.
PointBasis AtNode(const RadialMesh& mesh, int e, int q) {
  PointBasis b;
  const double jac = mesh.Jacobian(e);
  b.r = mesh.Radius(e, q);
  b.weight = mesh.Weights()[q] * jac;
  for (int i = 0; i < mesh.NGll(); ++i) {
    b.value[i] = i == q ? 1.0 : 0.0;
    b.rdr[i] = b.r * mesh.Derivative(q, i) / jac;
  }
  return b;
}

// Memory point p of element e.
PointBasis AtMemoryPoint(const RadialMesh& mesh, const MemoryPoints& pts, int e, int p) {
  PointBasis b;
  const double jac = mesh.Jacobian(e);
  for (int i = 0; i < mesh.NGll(); ++i) b.r += pts.Value(p, i) * mesh.Radius(e, i);
  b.weight = pts.weight[p] * jac;
  for (int i = 0; i < mesh.NGll(); ++i) {
    b.value[i] = pts.Value(p, i);
    b.rdr[i] = b.r * pts.Derivative(p, i) / jac;
  }
  return b;
}

// ---- energy matrices over the measures ---------------------------------------

// A symmetric quadratic form over M measures, built from combinations.
template <int M>
struct Energy {
  double q[M][M] = {};
  // c a a^T for the combination a.
  void Square(double c, const double (&a)[M]) {
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j) q[i][j] += c * a[i] * a[j];
  }
  // c (a b^T + b a^T).
  void Cross(double c, const double (&a)[M], const double (&b)[M]) {
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j) q[i][j] += c * (a[i] * b[j] + b[i] * a[j]);
  }
  // c at (i, j) and (j, i): the symmetric pair of one cross term.
  void Pair(int i, int j, double c) {
    q[i][j] += c;
    if (i != j) q[j][i] += c;
  }
};

// Spheroidal measures.
enum { mU = 0, mV = 1, mP = 2, mrU = 3, mrV = 4, mrP = 5 };
using Energy6 = Energy<6>;

// Toroidal measures.
enum { mW = 0, mrW = 1 };
using Energy2 = Energy<2>;

struct NodalMaterial {
  double C, A, F, L, N, rho, drho, g;
};

NodalMaterial MaterialAt(const Material& m, size_t at) {
  return {m.C[at], m.A[at], m.F[at], m.L[at], m.N[at], m.rho[at], m.drho[at], m.g[at]};
}

// The elastic (solid) part of the spheroidal energy.
void AddElastic(Energy6& E, const NodalMaterial& m, double k2) {
  const double x[6] = {0, 0, 0, 1, 0, 0};
  const double y[6] = {2, -k2, 0, 0, 0, 0};
  const double s[6] = {1, -1, 0, 0, 1, 0};
  const double v[6] = {0, 1, 0, 0, 0, 0};
  E.Square(m.C, x);
  E.Square(m.A - m.N, y);
  E.Cross(m.F, x, y);
  E.Square(k2 * m.L, s);
  E.Square(k2 * (k2 - 2.0) * m.N, v);
}

// The gravitational couplings of a solid point (the potential's own energy
// excluded).
void AddGravity(Energy6& E, const NodalMaterial& m, double r, double k2, double G) {
  E.Pair(mU, mU, 4.0 * m.rho * (kPi * G * m.rho * r - m.g) * r);
  E.Pair(mU, mV, k2 * m.rho * m.g * r);
  E.Pair(mV, mP, k2 * m.rho * r);
  E.Pair(mU, mrP, m.rho * r);
}

// A fluid's stratification r^2 rho' / g, zero where g vanishes.
double Stratification(double r, const NodalMaterial& m) {
  return m.g > 0.0 ? r * r * m.drho / m.g : 0.0;
}

// The point shear s in the deviatoric scalars.
void AddPointShear(Energy6& E, double s, double k2) {
  const double dM[6] = {0, 1, 0, 0, 0, 0};
  const double dR[6] = {-1, 0.5 * k2, 0, 1, 0, 0};
  const double dS[6] = {1, -1, 0, 0, 1, 0};
  E.Square(s * k2 * (k2 - 2.0), dM);
  E.Square(s * (4.0 / 3.0), dR);
  E.Square(s * k2, dS);
}

// ---- element matrices ------------------------------------------------------

// A dense element matrix over nodes x comps local dofs (node-major).
struct ElementMatrix {
  int nodes, comps;
  std::vector<double> k;
  ElementMatrix(int n, int c) : nodes(n), comps(c), k(static_cast<size_t>(n * c) * n * c, 0.0) {}
  double& operator()(int a, int b) { return k[static_cast<size_t>(a) * nodes * comps + b]; }
  // weight B^T Q B with the M = 2 comps measures: measure c is the value
  // of component c, measure comps + c its r d/dr.
  template <int M>
  void AddPoint(const PointBasis& p, const Energy<M>& E) {
    const int size = nodes * comps;
    for (int a = 0; a < size; ++a) {
      const int i = a / comps, c = a % comps;
      const double ba[2] = {p.value[i], p.rdr[i]};
      if (ba[0] == 0.0 && ba[1] == 0.0) continue;
      for (int b = a; b < size; ++b) {
        const int j = b / comps, d = b % comps;
        const double bb[2] = {p.value[j], p.rdr[j]};
        double sum = 0.0;
        for (int u = 0; u < 2; ++u)
          for (int v = 0; v < 2; ++v) sum += ba[u] * E.q[c + u * comps][d + v * comps] * bb[v];
        (*this)(a, b) += p.weight * sum;
      }
    }
  }
};

}  // namespace

// ---- SymmetricBand ------------------------------------------------------------

SymmetricBand::SymmetricBand(int n, int kd)
    : n_(n), kd_(kd), upper_(static_cast<size_t>(kd + 1) * n, 0.0) {}

void SymmetricBand::Add(int i, int j, double v) {
  if (i > j) std::swap(i, j);
  if (j - i > kd_) throw ArgumentError("an entry outside the band");
  upper_[Offset(i, j)] += v;
}

double SymmetricBand::Entry(int i, int j) const {
  if (i > j) std::swap(i, j);
  return j - i > kd_ ? 0.0 : upper_[Offset(i, j)];
}

std::vector<double> SymmetricBand::Multiply(const std::vector<double>& x, int nrhs) const {
  if (static_cast<long long>(x.size()) != static_cast<long long>(n_) * nrhs) {
    throw ArgumentError("expected ndof x nrhs values");
  }
  std::vector<double> y(x.size(), 0.0);
  for (int k = 0; k < nrhs; ++k) {
    const double* xk = x.data() + static_cast<size_t>(k) * n_;
    double* yk = y.data() + static_cast<size_t>(k) * n_;
    for (int i = 0; i < n_; ++i) {
      double sum = upper_[Offset(i, i)] * xk[i];
      for (int j = std::max(0, i - kd_); j < i; ++j) sum += upper_[Offset(j, i)] * xk[j];
      for (int j = i + 1; j <= std::min(n_ - 1, i + kd_); ++j) sum += upper_[Offset(i, j)] * xk[j];
      yk[i] = sum;
    }
  }
  return y;
}

std::vector<double> SymmetricBand::ScaledUpper(const std::vector<double>& scale) const {
  std::vector<double> out(upper_.size());
  for (int j = 0; j < n_; ++j)
    for (int i = std::max(0, j - kd_); i <= j; ++i)
      out[Offset(i, j)] = upper_[Offset(i, j)] * scale[i] * scale[j];
  return out;
}

std::vector<double> SymmetricBand::ScaledLu(const std::vector<double>& scale) const {
  const int rows = 3 * kd_ + 1;
  std::vector<double> out(static_cast<size_t>(rows) * n_, 0.0);
  for (int j = 0; j < n_; ++j) {
    for (int i = std::max(0, j - kd_); i <= std::min(n_ - 1, j + kd_); ++i) {
      const int lo = std::min(i, j), hi = std::max(i, j);
      out[static_cast<size_t>(j) * rows + (2 * kd_ + i - j)] =
          upper_[Offset(lo, hi)] * scale[lo] * scale[hi];
    }
  }
  return out;
}

std::vector<double> DiagonalScaling(const SymmetricBand& a, bool* positive) {
  std::vector<double> s(a.Size());
  bool allPositive = true;
  for (int j = 0; j < a.Size(); ++j) {
    const double d = a.Diagonal(j);
    allPositive = allPositive && d > 0.0;
    s[j] = d != 0.0 ? 1.0 / std::sqrt(std::abs(d)) : 1.0;
  }
  if (positive != nullptr) *positive = allPositive;
  return s;
}

// ---- spheroidal ---------------------------------------------------------------

namespace {

// The element above the last fluid element: where the non-gravitating
// problem starts (one fluid region at most, a core below the solid).
int SolidBase(const Material& mat) {
  const int E = mat.mesh->ElementCount();
  int base = 0, regions = 0;
  for (int e = 0; e < E; ++e) {
    if (!mat.IsFluid(e)) continue;
    if (e == 0 || !mat.IsFluid(e - 1)) ++regions;
    base = e + 1;
  }
  if (regions > 1) {
    throw ArgumentError("the non-gravitating problem takes at most one fluid region, a core");
  }
  return base;
}

}  // namespace

DegreeOperator::DegreeOperator(const Material& material, int l, const DegreeOptions& options)
    : material_(&material), l_(l), gravitating_(options.gravitating) {
  if (l < 0) throw ArgumentError("degree must be non-negative");
  const RadialMesh& mesh = *material.mesh;
  const int n = mesh.NGll();
  if (n > kMaxNodes) throw ArgumentError("at most 16 GLL nodes per element");
  const int E = mesh.ElementCount();
  firstElement_ = l > 0 ? mesh.StartElement(l, options.eps) : 0;
  if (!gravitating_ && l > 0) firstElement_ = std::max(firstElement_, SolidBase(material));
  if (!gravitating_ && l == 1) {
    throw ArgumentError("the non-gravitating degree 1 (rigid translation) is not supported");
  }
  if (firstElement_ >= E) throw ArgumentError("the degree's element subset is empty");

  // Numbering: upwards through the global nodes from the first element,
  // components U, V, phi in that order, each present where the node and
  // the problem carry it.
  const int nodes = mesh.NodeCount();
  const int g0 = mesh.GlobalIndex(firstElement_, 0);
  std::vector<char> solidNode(nodes, 0);
  for (int e = firstElement_; e < E; ++e) {
    if (l == 0 || !material.IsFluid(e)) {
      for (int i = 0; i < n; ++i) solidNode[mesh.GlobalIndex(e, i)] = 1;
    }
  }
  auto carries = [&](int node, Component c) {
    if (node < g0) return false;
    const bool centre = l == 0 && node == g0 && mesh.Left(firstElement_) == 0.0;
    const bool surface = node == nodes - 1;
    switch (c) {
      case Component::U: return solidNode[node] && !centre;
      case Component::V: return solidNode[node] && l > 0;
      case Component::Phi: return gravitating_ && !(surface && l == 1);
    }
    return false;
  };
  dof_.assign(static_cast<size_t>(nodes) * 3, -1);
  int ndof = 0;
  for (int node = g0; node < nodes; ++node) {
    for (int c = 0; c < 3; ++c) {
      if (carries(node, Component(c))) dof_[static_cast<size_t>(node) * 3 + c] = ndof++;
    }
  }
  // A node's dofs and its successor's interleave: the band half-width is
  // (components - 1) + components (n - 1).
  const int comps = gravitating_ ? 3 : (l == 0 ? 1 : 2);
  const int kd = comps * (n - 1) + comps - 1;
  matrix_ = SymmetricBand(ndof, kd);

  const double k2 = l * (l + 1.0);
  const double G = material.G;
  const double ifpg = 1.0 / (4.0 * kPi * G);
  const PointShear* shear = options.pointShear;
  for (int e = firstElement_; e < E; ++e) {
    const bool fluid = material.IsFluid(e) && l > 0;
    ElementMatrix K(n, 3);
    for (int q = 0; q < n; ++q) {
      const size_t at = static_cast<size_t>(e) * n + q;
      const PointBasis p = AtNode(mesh, e, q);
      const NodalMaterial m = MaterialAt(material, at);
      Energy6 Q;
      if (gravitating_) {
        Q.Pair(mrP, mrP, ifpg);
        Q.Pair(mP, mP, ifpg * k2);
      }
      if (fluid) {
        if (gravitating_) Q.Pair(mP, mP, Stratification(p.r, m));
      } else {
        AddElastic(Q, m, k2);
        if (gravitating_) AddGravity(Q, m, p.r, k2, G);
      }
      K.AddPoint(p, Q);
    }
    if (shear != nullptr && !fluid) {
      const MemoryPoints& pts = *shear->points;
      for (int k = 0; k < pts.count; ++k) {
        const double s = (*shear->modulus)[static_cast<size_t>(e) * pts.count + k];
        if (s == 0.0) continue;
        Energy6 Q;
        AddPointShear(Q, s, k2);
        K.AddPoint(AtMemoryPoint(mesh, pts, e, k), Q);
      }
    }
    // scatter the upper triangle
    for (int a = 0; a < 3 * n; ++a) {
      const int ga = Dof(mesh.GlobalIndex(e, a / 3), Component(a % 3));
      if (ga < 0) continue;
      for (int b = a; b < 3 * n; ++b) {
        const int gb = Dof(mesh.GlobalIndex(e, b / 3), Component(b % 3));
        if (gb < 0) continue;
        const double v = K(a, b);
        if (v != 0.0) matrix_.Add(ga, gb, v);
      }
    }
  }

  auto add = [&](int i, int j, double v) {
    if (i >= 0 && j >= 0) matrix_.Add(i, j, v);
  };
  if (gravitating_ && l > 0) {
    // Fluid-solid boundaries: the surface terms of the gravitational
    // energy, +-rho_f (g U^2 + 2 U phi) r^2 with the fluid's density, +
    // for a fluid below the boundary.
    for (int e = firstElement_; e + 1 < E; ++e) {
      if (material.IsFluid(e) == material.IsFluid(e + 1)) continue;
      const int node = mesh.GlobalIndex(e, n - 1);
      const double r = mesh.Radius(e, n - 1);
      const double g = material.g[static_cast<size_t>(e) * n + n - 1];
      const bool fluidBelow = material.IsFluid(e);
      const double rhoF = fluidBelow ? material.rho[static_cast<size_t>(e) * n + n - 1]
                                     : material.rho[static_cast<size_t>(e + 1) * n];
      const double sign = fluidBelow ? 1.0 : -1.0;
      const int u = Dof(node, Component::U), phi = Dof(node, Component::Phi);
      add(u, u, sign * rhoF * g * r * r);
      add(u, phi, sign * rhoF * r * r);
    }
    // The exterior and (for a truncated subset) interior potential fields
    // as Dirichlet-to-Neumann terms.
  }
  if (gravitating_) {
    const int top = mesh.GlobalIndex(E - 1, n - 1);
    add(Dof(top, Component::Phi), Dof(top, Component::Phi), (l + 1) * material.radius * ifpg);
    const double rb = mesh.Left(firstElement_);
    if (rb > 0.0) {
      const int bottom = mesh.GlobalIndex(firstElement_, 0);
      add(Dof(bottom, Component::Phi), Dof(bottom, Component::Phi), l * rb * ifpg);
    }
  }
  scale_ = DiagonalScaling(matrix_, &positive_);
}

int DegreeOperator::SurfaceDof(Component c) const {
  return Dof(material_->mesh->NodeCount() - 1, c);
}

std::vector<double> DegreeOperator::Load() const {
  std::vector<double> b(DofCount(), 0.0);
  const double a = material_->radius;
  b[SurfaceDof(Component::U)] = -material_->surfaceGravity * a * a;
  if (SurfaceDof(Component::Phi) >= 0) b[SurfaceDof(Component::Phi)] = -a * a;
  return b;
}

std::vector<double> DegreeOperator::LoadForce() const {
  std::vector<double> b(DofCount(), 0.0);
  const double a = material_->radius;
  b[SurfaceDof(Component::U)] = -material_->surfaceGravity * a * a;
  return b;
}

std::vector<double> DegreeOperator::LoadPotential() const {
  std::vector<double> b(DofCount(), 0.0);
  const double a = material_->radius;
  if (SurfaceDof(Component::Phi) >= 0) b[SurfaceDof(Component::Phi)] = -a * a;
  return b;
}

std::vector<double> DegreeOperator::Tangential() const {
  std::vector<double> b(DofCount(), 0.0);
  const double a = material_->radius;
  if (SurfaceDof(Component::V) >= 0) {
    b[SurfaceDof(Component::V)] = -material_->surfaceGravity * a * a * l_ * (l_ + 1.0);
  }
  return b;
}

std::vector<double> DegreeOperator::Traction() const {
  std::vector<double> b(DofCount(), 0.0);
  const double a = material_->radius;
  b[SurfaceDof(Component::U)] = -a * a;
  return b;
}

std::vector<double> DegreeOperator::TractionTangential() const {
  std::vector<double> b(DofCount(), 0.0);
  const double a = material_->radius;
  if (SurfaceDof(Component::V) >= 0) b[SurfaceDof(Component::V)] = -a * a * l_ * (l_ + 1.0);
  return b;
}

std::vector<double> DegreeOperator::Tide(int power) const {
  const int p = power < 0 ? l_ : power;
  std::vector<double> b(DofCount(), 0.0);
  if (p == 0 || !gravitating_) return b;
  const Material& mat = *material_;
  const RadialMesh& mesh = *mat.mesh;
  const int n = mesh.NGll();
  const double a = mat.radius;
  const double k2 = l_ * (l_ + 1.0);
  for (int e = firstElement_; e < mesh.ElementCount(); ++e) {
    const bool fluid = mat.IsFluid(e) && l_ > 0;
    for (int q = 0; q < n; ++q) {
      const PointBasis pt = AtNode(mesh, e, q);
      const NodalMaterial m = MaterialAt(mat, static_cast<size_t>(e) * n + q);
      // psi and r psi' at the point, exactly
      double psiM[6] = {};
      psiM[mP] = std::pow(pt.r / a, p);
      psiM[mrP] = pt.r * (p / a) * std::pow(pt.r / a, p - 1);
      Energy6 Q;
      if (fluid) {
        Q.Pair(mP, mP, Stratification(pt.r, m));
      } else {
        AddGravity(Q, m, pt.r, k2, mat.G);
      }
      // rows: the point's value measures of U, V (solid) or phi (fluid)
      const int node = mesh.GlobalIndex(e, q);
      for (int c = 0; c < 3; ++c) {
        const int d = Dof(node, Component(c));
        if (d < 0) continue;
        if (fluid != (c == static_cast<int>(Component::Phi))) continue;
        double f = 0.0;
        for (int k = 0; k < 6; ++k) f += Q.q[c][k] * psiM[k];
        b[d] -= pt.weight * f;
      }
    }
  }
  if (l_ > 0) {
    for (int e = firstElement_; e + 1 < mesh.ElementCount(); ++e) {
      if (mat.IsFluid(e) == mat.IsFluid(e + 1)) continue;
      const int node = mesh.GlobalIndex(e, n - 1);
      const double r = mesh.Radius(e, n - 1);
      const bool fluidBelow = mat.IsFluid(e);
      const double rhoF = fluidBelow ? mat.rho[static_cast<size_t>(e) * n + n - 1]
                                     : mat.rho[static_cast<size_t>(e + 1) * n];
      b[Dof(node, Component::U)] -= (fluidBelow ? 1.0 : -1.0) * rhoF * r * r * std::pow(r / a, p);
    }
  }
  return b;
}

std::vector<double> DegreeOperator::Apply(const std::vector<double>& x, int nrhs) const {
  return matrix_.Multiply(x, nrhs);
}

// ---- toroidal -------------------------------------------------------------------

ToroidalOperator::ToroidalOperator(const Material& material, int l, double eps,
                                   const PointShear* pointShear)
    : material_(&material), l_(l) {
  if (l < 1) throw ArgumentError("the toroidal system exists for degrees one and above");
  const RadialMesh& mesh = *material.mesh;
  const int n = mesh.NGll();
  if (n > kMaxNodes) throw ArgumentError("at most 16 GLL nodes per element");
  const int E = mesh.ElementCount();
  firstElement_ = mesh.StartElement(l, eps);
  // W on the nodes of solid elements, upwards; each run of solid elements
  // (a region) is a contiguous dof range.
  dof_.assign(mesh.NodeCount(), -1);
  int ndof = 0;
  std::vector<int> regionTop;
  for (int e = firstElement_; e < E; ++e) {
    if (material.IsFluid(e)) continue;
    for (int i = 0; i < n; ++i) {
      int& d = dof_[mesh.GlobalIndex(e, i)];
      if (d < 0) d = ndof++;
    }
    if (e + 1 == E || material.IsFluid(e + 1)) regionTop.push_back(ndof - 1);
  }
  if (ndof == 0) throw ArgumentError("the degree's subset holds no solid element");
  if (l == 1) pinned_ = regionTop;  // rigid rotations: the top of each region
  matrix_ = SymmetricBand(ndof, n - 1);

  const double k2 = l * (l + 1.0);
  for (int e = firstElement_; e < E; ++e) {
    if (material.IsFluid(e)) continue;
    ElementMatrix K(n, 1);
    const double t[2] = {-1.0, 1.0};  // r W' - W
    const double w[2] = {1.0, 0.0};
    for (int q = 0; q < n; ++q) {
      const size_t at = static_cast<size_t>(e) * n + q;
      Energy2 Q;
      Q.Square(k2 * material.L[at], t);
      Q.Square(k2 * (k2 - 2.0) * material.N[at], w);
      K.AddPoint(AtNode(mesh, e, q), Q);
    }
    if (pointShear != nullptr) {
      const MemoryPoints& pts = *pointShear->points;
      for (int k = 0; k < pts.count; ++k) {
        const double s = (*pointShear->modulus)[static_cast<size_t>(e) * pts.count + k];
        if (s == 0.0) continue;
        Energy2 Q;
        Q.Square(s * k2, t);
        Q.Square(s * k2 * (k2 - 2.0), w);
        K.AddPoint(AtMemoryPoint(mesh, pts, e, k), Q);
      }
    }
    for (int a = 0; a < n; ++a) {
      const int ga = dof_[mesh.GlobalIndex(e, a)];
      for (int b = a; b < n; ++b) {
        const double v = K(a, b);
        if (v != 0.0) matrix_.Add(ga, dof_[mesh.GlobalIndex(e, b)], v);
      }
    }
  }
  bool positive = true;
  scale_ = DiagonalScaling(matrix_, &positive);
  if (!positive) {
    for (int j = 0; j < ndof; ++j) {
      if (!(matrix_.Diagonal(j) > 0.0)) {
        throw NumericalError("toroidal degree " + std::to_string(l) +
                             ": a non-positive diagonal at dof " + std::to_string(j));
      }
    }
  }
}

std::vector<double> ToroidalOperator::ScaledUpperBand() const {
  std::vector<double> out = matrix_.ScaledUpper(scale_);
  const int n = matrix_.Size(), kd = matrix_.Kd();
  for (const int p : pinned_) {
    // row and column p of the identity
    for (int j = std::max(0, p - kd); j <= std::min(n - 1, p + kd); ++j) {
      const int lo = std::min(p, j), hi = std::max(p, j);
      out[static_cast<size_t>(hi) * (kd + 1) + (kd + lo - hi)] = lo == hi ? 1.0 : 0.0;
    }
  }
  return out;
}

std::vector<double> ToroidalOperator::Apply(const std::vector<double>& x, int nrhs) const {
  return matrix_.Multiply(x, nrhs);
}

}  // namespace sl3amd
