// Degree systems, host side: dof maps, band assembly, scaling.
#include "operators.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "common.hpp"

namespace sl3amd {

namespace {
constexpr double kPi = 3.14159265358979323846;

// r^2 (d rho / dr) / g, zero where g vanishes (src/DegreeOperator.cpp:18-21).
double Stratification(double r, double drho, double g) {
  return g > 0.0 ? r * r * drho / g : 0.0;
}
}  // namespace

// ---- spheroidal -------------------------------------------------------------

DegreeOperator::DegreeOperator(const Material& material, int l, double eps,
                               const PointShear* pointShear)
    : material_(&material), l_(l) {
  if (l < 0) throw ArgumentError("degree must be non-negative");
  const auto& mesh = *material.mesh;
  firstElement_ = l > 0 ? mesh.StartElement(l, eps) : 0;
  kd_ = 3 * (mesh.NGll() - 1) + 2;  // three components, gravitating
  BuildDofMap();
  Assemble(pointShear);
  // Factorise's scaling (src/DegreeOperator.cpp:781-788).
  scale_.resize(ndof_);
  positive_ = true;
  for (int j = 0; j < ndof_; ++j) {
    const double d = band_[static_cast<size_t>(j) * (kd_ + 1) + kd_];
    positive_ = positive_ && d > 0.0;
    scale_[j] = 1.0 / std::sqrt(std::abs(d) > 0.0 ? std::abs(d) : 1.0);
  }
}

// src/DegreeOperator.cpp:77-127 (gravitating): U, V, phi on solid nodes (no
// V at l = 0), phi alone on fluid ones, upwards from the first element; no
// surface phi at l = 1; no U at a centre at l = 0.
void DegreeOperator::BuildDofMap() {
  const auto& mesh = *material_->mesh;
  const int n = mesh.NGll();
  const int nglob = mesh.NodeCount();
  const int g0 = mesh.GlobalIndex(firstElement_, 0);
  std::vector<char> solid(nglob, 0);
  for (int e = firstElement_; e < mesh.ElementCount(); ++e) {
    if (l_ == 0 || !material_->IsFluid(e)) {
      for (int i = 0; i < n; ++i) solid[mesh.GlobalIndex(e, i)] = 1;
    }
  }
  dof_.assign(static_cast<size_t>(nglob) * 3, -1);
  ndof_ = 0;
  for (int gn = g0; gn < nglob; ++gn) {
    if (solid[gn]) {
      dof_[gn * 3 + 0] = ndof_++;
      if (l_ != 0) dof_[gn * 3 + 1] = ndof_++;
      dof_[gn * 3 + 2] = ndof_++;
    } else {
      dof_[gn * 3 + 2] = ndof_++;
    }
  }
  auto drop = [&](size_t position) {
    const int removed = dof_[position];
    dof_[position] = -1;
    for (auto& d : dof_) {
      if (d > removed) --d;
    }
    --ndof_;
  };
  if (l_ == 1) drop(static_cast<size_t>(nglob - 1) * 3 + 2);
  if (l_ == 0 && mesh.Left(firstElement_) == 0.0) drop(static_cast<size_t>(g0) * 3);
}

int DegreeOperator::SurfaceDof(Component c) const {
  const auto& mesh = *material_->mesh;
  return Dof(mesh.GlobalIndex(mesh.ElementCount() - 1, mesh.NGll() - 1), c);
}

// src/DegreeOperator.cpp:186-474 with the base coefficients (:197-231), the
// point shear without a point bulk, and the closures.
void DegreeOperator::Assemble(const PointShear* pointShear) {
  const auto& mat = *material_;
  const auto& mesh = *mat.mesh;
  const int n = mesh.NGll();
  const int n3 = 3 * n;
  const double k2 = l_ * (l_ + 1.0);
  const double G = mat.G;
  const double ifpg = 1.0 / (4.0 * kPi * G);
  const auto& w = mesh.Weights();
  const size_t size = mat.rho.size();

  // The per-node coefficients.
  std::vector<double> AN(size), UU(size), UV(size), VP(size), UPhi(size),
      strat(size);
  for (size_t i = 0; i < size; ++i) {
    const double r = mesh.Radii()[i];
    const double rho = mat.rho[i], g = mat.g[i];
    AN[i] = mat.A[i] - mat.N[i];
    UU[i] = 4.0 * rho * (kPi * G * rho * r - g) * r;
    UV[i] = k2 * rho * g * r;
    VP[i] = k2 * rho * r;
    UPhi[i] = rho * r * r;
    strat[i] = Stratification(r, mat.drho[i], g);
  }

  band_.assign(static_cast<size_t>(kd_ + 1) * ndof_, 0.0);
  std::vector<double> local(static_cast<size_t>(n3) * n3);
  std::vector<double> B(static_cast<size_t>(n) * n);
  std::vector<double> X(n3), Y(n3), S(n3), dM(n3), dR(n3), dS(n3);
  auto at = [&](int a, int b) -> double& {
    return local[static_cast<size_t>(a) * n3 + b];
  };

  for (int e = firstElement_; e < mesh.ElementCount(); ++e) {
    std::fill(local.begin(), local.end(), 0.0);
    const size_t base = static_cast<size_t>(e) * n;
    const double jac = mesh.Jacobian(e);
    const bool fluid = mat.IsFluid(e) && l_ > 0;
    for (int q = 0; q < n; ++q) {
      for (int i = 0; i < n; ++i) {
        B[q * n + i] = mesh.Radius(e, q) * mesh.Derivative(q, i) / jac;
      }
    }
    // phi' phi' stiffness: int r^2 phi' psi' dr / (4 pi G)
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        double kp = 0.0;
        for (int q = 0; q < n; ++q) {
          const double rq = mesh.Radius(e, q);
          kp += mesh.Derivative(q, i) * w[q] * rq * rq * mesh.Derivative(q, j);
        }
        at(3 * i + 2, 3 * j + 2) += ifpg * kp / jac;
      }
    }
    for (int q = 0; q < n; ++q) {
      const double Wq = w[q] * jac;
      at(3 * q + 2, 3 * q + 2) += ifpg * k2 * Wq;
      if (fluid) {
        at(3 * q + 2, 3 * q + 2) += strat[base + q] * Wq;
        continue;
      }
      std::fill(X.begin(), X.end(), 0.0);
      std::fill(Y.begin(), Y.end(), 0.0);
      std::fill(S.begin(), S.end(), 0.0);
      for (int i = 0; i < n; ++i) {
        const double b = B[q * n + i];
        X[3 * i] = b;
        S[3 * i + 1] = b;
      }
      Y[3 * q] = 2.0;
      S[3 * q] = 1.0;
      Y[3 * q + 1] = -k2;
      S[3 * q + 1] -= 1.0;
      const double cC = mat.C[base + q], cAN = AN[base + q],
                   cF = mat.F[base + q], cL = mat.L[base + q];
      for (int a = 0; a < n3; ++a) {
        const double xa = X[a], ya = Y[a], sa = S[a];
        if (xa == 0.0 && ya == 0.0 && sa == 0.0) continue;
        for (int b = 0; b < n3; ++b) {
          at(a, b) += Wq * (cC * xa * X[b] + cAN * ya * Y[b] +
                            cF * (xa * Y[b] + ya * X[b]) + k2 * cL * sa * S[b]);
        }
      }
      at(3 * q + 1, 3 * q + 1) += k2 * (k2 - 2.0) * mat.N[base + q] * Wq;
      at(3 * q, 3 * q) += UU[base + q] * Wq;
      const double cUV = UV[base + q] * Wq;
      at(3 * q, 3 * q + 1) += cUV;
      at(3 * q + 1, 3 * q) += cUV;
      const double cVP = VP[base + q] * Wq;
      at(3 * q + 1, 3 * q + 2) += cVP;
      at(3 * q + 2, 3 * q + 1) += cVP;
    }
    if (!fluid) {
      // int rho r^2 (U phi' + phi U') dr
      for (int i = 0; i < n; ++i) {
        const double mi = UPhi[base + i] * w[i];
        for (int j = 0; j < n; ++j) {
          const double m = mi * mesh.Derivative(i, j);
          at(3 * i, 3 * j + 2) += m;
          at(3 * j + 2, 3 * i) += m;
        }
      }
    }
    if (pointShear != nullptr && !fluid) {
      const auto& pts = *pointShear->points;
      const auto& modulus = *pointShear->modulus;
      for (int p = 0; p < pts.count; ++p) {
        const double s = modulus[static_cast<size_t>(e) * pts.count + p];
        if (s == 0.0) continue;
        double rp = 0.0;
        for (int i = 0; i < n; ++i) rp += pts.Value(p, i) * mesh.Radius(e, i);
        std::fill(dM.begin(), dM.end(), 0.0);
        std::fill(dR.begin(), dR.end(), 0.0);
        std::fill(dS.begin(), dS.end(), 0.0);
        for (int i = 0; i < n; ++i) {
          const double v = pts.Value(p, i);
          const double b = rp * pts.Derivative(p, i) / jac;
          dR[3 * i] = b - v;
          dR[3 * i + 1] = 0.5 * k2 * v;
          dS[3 * i] = v;
          dS[3 * i + 1] = b - v;
          dM[3 * i + 1] = v;
        }
        const double weight = pts.weight[p] * jac * s;
        for (int a = 0; a < n3; ++a) {
          if (dM[a] == 0.0 && dR[a] == 0.0 && dS[a] == 0.0) continue;
          for (int b = 0; b < n3; ++b) {
            at(a, b) += weight * (k2 * (k2 - 2.0) * dM[a] * dM[b] +
                                  (4.0 / 3.0) * dR[a] * dR[b] +
                                  k2 * dS[a] * dS[b]);
          }
        }
      }
    }
    for (int a = 0; a < n3; ++a) {
      const int ga = Dof(mesh.GlobalIndex(e, a / 3), Component(a % 3));
      if (ga < 0) continue;
      for (int b = 0; b < n3; ++b) {
        const int gb = Dof(mesh.GlobalIndex(e, b / 3), Component(b % 3));
        if (gb < ga) continue;
        const double value = at(a, b);
        if (value != 0.0) BandAt(ga, gb) += value;
      }
    }
  }
  auto point = [&](int i, int j, double value) {
    if (i < 0 || j < 0) return;
    if (i > j) std::swap(i, j);
    BandAt(i, j) += value;
  };
  // Fluid-solid interfaces (src/DegreeOperator.cpp:163-184, 446-460).
  if (l_ > 0) {
    for (int e = firstElement_; e + 1 < mesh.ElementCount(); ++e) {
      if (mat.IsFluid(e) == mat.IsFluid(e + 1)) continue;
      const int gn = mesh.GlobalIndex(e, n - 1);
      const double rt = mesh.Radius(e, n - 1);
      const double gt = mat.g[static_cast<size_t>(e) * n + (n - 1)];
      double rhoF, sign;
      if (mat.IsFluid(e)) {
        rhoF = mat.rho[static_cast<size_t>(e) * n + (n - 1)];
        sign = 1.0;
      } else {
        rhoF = mat.rho[static_cast<size_t>(e + 1) * n];
        sign = -1.0;
      }
      const int iU = Dof(gn, Component::U), iP = Dof(gn, Component::Phi);
      point(iU, iU, sign * rhoF * gt * rt * rt);
      point(iU, iP, sign * rhoF * rt * rt);
    }
  }
  // Dirichlet-to-Neumann closures.
  const int top = mesh.GlobalIndex(mesh.ElementCount() - 1, n - 1);
  if (l_ != 1) {
    point(Dof(top, Component::Phi), Dof(top, Component::Phi),
          (l_ + 1) * mat.radius * ifpg);
  }
  if (mesh.Left(firstElement_) > 0.0) {
    const int bottom = mesh.GlobalIndex(firstElement_, 0);
    point(Dof(bottom, Component::Phi), Dof(bottom, Component::Phi),
          l_ * mesh.Left(firstElement_) * ifpg);
  }
}

std::vector<double> DegreeOperator::ScaledUpperBand() const {
  const int ld = kd_ + 1;
  std::vector<double> out(static_cast<size_t>(ld) * ndof_, 0.0);
  for (int j = 0; j < ndof_; ++j) {
    for (int i = std::max(0, j - kd_); i <= j; ++i) {
      out[static_cast<size_t>(j) * ld + (kd_ + i - j)] =
          band_[static_cast<size_t>(j) * ld + (kd_ + i - j)] * scale_[i] *
          scale_[j];
    }
  }
  return out;
}

std::vector<double> DegreeOperator::ScaledLuBand() const {
  const int ld = kd_ + 1, ldlu = 3 * kd_ + 1;
  std::vector<double> out(static_cast<size_t>(ldlu) * ndof_, 0.0);
  auto scaled = [&](int i, int j) {
    return band_[static_cast<size_t>(j) * ld + (kd_ + i - j)] * scale_[i] *
           scale_[j];
  };
  for (int j = 0; j < ndof_; ++j) {
    for (int i = std::max(0, j - kd_); i <= std::min(ndof_ - 1, j + kd_); ++i) {
      out[static_cast<size_t>(j) * ldlu + (2 * kd_ + i - j)] =
          i <= j ? scaled(i, j) : scaled(j, i);
    }
  }
  return out;
}

std::vector<double> DegreeOperator::Load() const {
  std::vector<double> b(ndof_, 0.0);
  const double a = material_->radius;
  b[SurfaceDof(Component::U)] = -material_->surfaceGravity * a * a;
  const int phi = SurfaceDof(Component::Phi);
  if (phi >= 0) b[phi] = -a * a;
  return b;
}

// The force of an external potential psi = (r/a)^p on the degree system
// (src/DegreeOperator.cpp:999-1038): minus the gravitational couplings of
// the assembly applied to psi. Solid nodes take the U row (rho r^2 psi')
// and the V row (l(l+1) rho r psi); fluid nodes the stratification on the
// potential; each fluid-solid interface node the point coupling
// sign rho_F r^2 psi. p = 0 is no force.
std::vector<double> DegreeOperator::Tide(int power) const {
  const int p = power < 0 ? l_ : power;
  std::vector<double> b(ndof_, 0.0);
  if (p == 0) return b;
  const auto& mat = *material_;
  const auto& mesh = *mat.mesh;
  const int n = mesh.NGll();
  const double a = mat.radius;
  const double k2 = l_ * (l_ + 1.0);
  const auto& w = mesh.Weights();
  for (int e = firstElement_; e < mesh.ElementCount(); ++e) {
    const bool fluid = mat.IsFluid(e) && l_ > 0;
    const double jac = mesh.Jacobian(e);
    for (int q = 0; q < n; ++q) {
      const size_t at = static_cast<size_t>(e) * n + q;
      const int gn = mesh.GlobalIndex(e, q);
      const double r = mesh.Radius(e, q);
      const double wq = w[q] * jac;
      const double psi = std::pow(r / a, p);
      if (fluid) {
        b[Dof(gn, Component::Phi)] -= wq * Stratification(r, mat.drho[at], mat.g[at]) * psi;
        continue;
      }
      const double dpsi = (p / a) * std::pow(r / a, p - 1);
      const int iU = Dof(gn, Component::U);
      if (iU >= 0) b[iU] -= wq * mat.rho[at] * r * r * dpsi;
      if (l_ > 0) b[Dof(gn, Component::V)] -= wq * k2 * mat.rho[at] * r * psi;
    }
  }
  if (l_ > 0) {
    for (int e = firstElement_; e + 1 < mesh.ElementCount(); ++e) {
      if (mat.IsFluid(e) == mat.IsFluid(e + 1)) continue;
      const int gn = mesh.GlobalIndex(e, n - 1);
      const double rt = mesh.Radius(e, n - 1);
      const bool below = mat.IsFluid(e);  // the fluid side is the lower element
      const double rhoF = below ? mat.rho[static_cast<size_t>(e) * n + (n - 1)]
                                : mat.rho[static_cast<size_t>(e + 1) * n];
      const double sign = below ? 1.0 : -1.0;
      b[Dof(gn, Component::U)] -= sign * rhoF * rt * rt * std::pow(rt / a, p);
    }
  }
  return b;
}

static std::vector<double> ApplyUpperBand(const std::vector<double>& band,
                                          int ndof, int kd,
                                          const std::vector<double>& x,
                                          int nrhs) {
  const int ldab = kd + 1;
  std::vector<double> y(x.size(), 0.0);
  for (int k = 0; k < nrhs; ++k) {
    const double* xk = x.data() + static_cast<size_t>(k) * ndof;
    double* yk = y.data() + static_cast<size_t>(k) * ndof;
    for (int j = 0; j < ndof; ++j) {
      for (int i = std::max(0, j - kd); i <= j; ++i) {
        const double a = band[static_cast<size_t>(j) * ldab + (kd + i - j)];
        yk[i] += a * xk[j];
        if (i != j) yk[j] += a * xk[i];
      }
    }
  }
  return y;
}

std::vector<double> DegreeOperator::Apply(const std::vector<double>& x,
                                          int nrhs) const {
  if (static_cast<int>(x.size()) != ndof_ * nrhs) {
    throw ArgumentError("expected ndof x nrhs values");
  }
  return ApplyUpperBand(band_, ndof_, kd_, x, nrhs);
}

// ---- toroidal ---------------------------------------------------------------

// src/ToroidalOperator.cpp:14-239.
ToroidalOperator::ToroidalOperator(const Material& material, int l, double eps,
                                   const PointShear* pointShear)
    : material_(&material), l_(l) {
  if (l < 1) {
    throw ArgumentError("the toroidal system exists for degrees one and above");
  }
  const auto& mesh = *material.mesh;
  const int n = mesh.NGll();
  firstElement_ = mesh.StartElement(l, eps);
  kd_ = n - 1;
  dof_.assign(mesh.NodeCount(), -1);
  bool inRegion = false;
  for (int e = firstElement_; e < mesh.ElementCount(); ++e) {
    if (material.IsFluid(e)) {
      if (inRegion) regions_.back().second = ndof_;
      inRegion = false;
      continue;
    }
    if (!inRegion) regions_.push_back({ndof_, ndof_});
    inRegion = true;
    for (int i = 0; i < n; ++i) {
      const int gn = mesh.GlobalIndex(e, i);
      if (dof_[gn] < 0) dof_[gn] = ndof_++;
    }
  }
  if (inRegion) regions_.back().second = ndof_;
  if (ndof_ == 0) throw ArgumentError("the degree's subset holds no solid element");
  if (l_ == 1) {
    for (const auto& [first, end] : regions_) pinned_.push_back(end - 1);
  }
  // Assembly: k^2 L (r d_r W - W)^2 + k^2 (k^2 - 2) N W^2 at the GLL nodes,
  // plus the point shear.
  const double k2 = l * (l + 1.0);
  const auto& w = mesh.Weights();
  band_.assign(static_cast<size_t>(kd_ + 1) * ndof_, 0.0);
  auto bandAt = [&](int i, int j) -> double& {
    return band_[static_cast<size_t>(j) * (kd_ + 1) + (kd_ + i - j)];
  };
  for (int e = firstElement_; e < mesh.ElementCount(); ++e) {
    if (material.IsFluid(e)) continue;
    const size_t base = static_cast<size_t>(e) * n;
    const double jac = mesh.Jacobian(e);
    for (int i = 0; i < n; ++i) {
      const int gi = dof_[mesh.GlobalIndex(e, i)];
      for (int j = i; j < n; ++j) {
        const int gj = dof_[mesh.GlobalIndex(e, j)];
        double value = 0.0;
        for (int q = 0; q < n; ++q) {
          const double rq = mesh.Radius(e, q);
          const double si = rq * mesh.Derivative(q, i) / jac - (q == i ? 1.0 : 0.0);
          const double sj = rq * mesh.Derivative(q, j) / jac - (q == j ? 1.0 : 0.0);
          value += w[q] * jac * k2 * material.L[base + q] * si * sj;
        }
        if (i == j) value += w[i] * jac * k2 * (k2 - 2.0) * material.N[base + i];
        bandAt(std::min(gi, gj), std::max(gi, gj)) += value;
      }
    }
  }
  if (pointShear != nullptr) {
    const auto& pts = *pointShear->points;
    const auto& modulus = *pointShear->modulus;
    for (int e = firstElement_; e < mesh.ElementCount(); ++e) {
      if (material.IsFluid(e)) continue;
      const double jac = mesh.Jacobian(e);
      for (int p = 0; p < pts.count; ++p) {
        const double s = modulus[static_cast<size_t>(e) * pts.count + p];
        if (s == 0.0) continue;
        double rp = 0.0;
        for (int i = 0; i < n; ++i) rp += pts.Value(p, i) * mesh.Radius(e, i);
        const double weight = pts.weight[p] * jac * s;
        for (int i = 0; i < n; ++i) {
          const int gi = dof_[mesh.GlobalIndex(e, i)];
          const double ti = rp * pts.Derivative(p, i) / jac - pts.Value(p, i);
          const double wi = pts.Value(p, i);
          for (int j = i; j < n; ++j) {
            const int gj = dof_[mesh.GlobalIndex(e, j)];
            const double tj = rp * pts.Derivative(p, j) / jac - pts.Value(p, j);
            const double wj = pts.Value(p, j);
            const double value = weight * (k2 * ti * tj + k2 * (k2 - 2.0) * wi * wj);
            if (value == 0.0) continue;
            bandAt(std::min(gi, gj), std::max(gi, gj)) += value;
          }
        }
      }
    }
  }
  scale_.resize(ndof_);
  for (int j = 0; j < ndof_; ++j) {
    const double d = band_[static_cast<size_t>(j) * (kd_ + 1) + kd_];
    if (!(d > 0.0)) {
      throw NumericalError("toroidal degree " + std::to_string(l_) +
                           ": a non-positive diagonal at dof " + std::to_string(j));
    }
    scale_[j] = 1.0 / std::sqrt(d);
  }
}

std::vector<double> ToroidalOperator::ScaledUpperBand() const {
  const int ld = kd_ + 1;
  std::vector<double> out(static_cast<size_t>(ld) * ndof_, 0.0);
  for (int j = 0; j < ndof_; ++j) {
    for (int i = std::max(0, j - kd_); i <= j; ++i) {
      out[static_cast<size_t>(j) * ld + (kd_ + i - j)] =
          band_[static_cast<size_t>(j) * ld + (kd_ + i - j)] * scale_[i] *
          scale_[j];
    }
  }
  for (const int p : pinned_) {
    for (int j = std::max(0, p - kd_); j <= std::min(ndof_ - 1, p + kd_); ++j) {
      const int i = std::min(p, j), k = std::max(p, j);
      out[static_cast<size_t>(k) * ld + (kd_ + i - k)] = i == k ? 1.0 : 0.0;
    }
  }
  return out;
}

std::vector<double> ToroidalOperator::Apply(const std::vector<double>& x,
                                            int nrhs) const {
  if (static_cast<int>(x.size()) != ndof_ * nrhs) {
    throw ArgumentError("expected ndof x nrhs values");
  }
  return ApplyUpperBand(band_, ndof_, kd_, x, nrhs);
}

}  // namespace sl3amd
