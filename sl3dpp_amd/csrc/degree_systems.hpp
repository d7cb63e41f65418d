// The per-degree radial systems of the step, host side: dof numbering, the
// Galerkin assembly of the weak form into a symmetric band, and the diagonal
// scaling that precedes the GPU factorisation (band.hip). The class names and
// the queries they answer are the reference's interface
// (DegreeOperator, include/sl3dpp/DegreeOperator.hpp:94-384;
// ToroidalOperator, include/sl3dpp/ToroidalOperator.hpp:58-176); the
// assembly is organised differently: every integration point contributes
// w B^T Q B, with B the point's "measures" of the element's fields (values
// and r d/dr of U, V, phi, or of W) and Q a small symmetric energy matrix
// built from the material at that point. GLL nodes and memory points go
// through the same path; only their basis values differ.
#pragma once

#include <utility>
#include <vector>

#include "model.hpp"

namespace sl3amd {

// A shear modulus added at the memory points (the point shear of
// include/sl3dpp/MemoryPoints.hpp:144-154), indexed (element, point).
struct PointShear {
  const MemoryPoints* points = nullptr;
  const std::vector<double>* modulus = nullptr;
};

enum class Component { U = 0, V = 1, Phi = 2 };

// DegreeOperator::Options' fields this build honours
// (include/sl3dpp/DegreeOperator.hpp:101-131).
struct DegreeOptions {
  double eps = 1e-8;        // graded truncation level
  bool gravitating = true;  // self-gravitation (phi dofs, gravity terms)
  const PointShear* pointShear = nullptr;
};

// A real symmetric band matrix kept as its upper band in LAPACK 'U' storage,
// (kd + 1) x n column-major: A(i, j), i <= j, at (kd + i - j) + j (kd + 1).
class SymmetricBand {
 public:
  SymmetricBand() = default;
  SymmetricBand(int n, int kd);
  int Size() const { return n_; }
  int Kd() const { return kd_; }
  // Adds v to A(i, j) and, off the diagonal, to A(j, i); either order.
  void Add(int i, int j, double v);
  double Diagonal(int j) const { return upper_[Offset(j, j)]; }
  double Entry(int i, int j) const;  // zero outside the band
  const std::vector<double>& Upper() const { return upper_; }
  // y = A x for nrhs column-major columns of length n.
  std::vector<double> Multiply(const std::vector<double>& x, int nrhs) const;
  // S A S with S = diag(scale), as an upper band ((kd + 1) x n) or as the
  // full band in LU storage (2 kd extra rows above: 3 kd + 1 rows, A(i, j)
  // at (2 kd + i - j) + j (3 kd + 1)).
  std::vector<double> ScaledUpper(const std::vector<double>& scale) const;
  std::vector<double> ScaledLu(const std::vector<double>& scale) const;

 private:
  size_t Offset(int i, int j) const {
    return static_cast<size_t>(j) * (kd_ + 1) + (kd_ + i - j);
  }
  int n_ = 0, kd_ = 0;
  std::vector<double> upper_;
};

// The symmetric scaling of DegreeOperator::Factorise
// (src/DegreeOperator.cpp:781-788): s_j = 1 / sqrt|A_jj| (1 for a zero
// diagonal); `positive` says whether every diagonal is > 0, the condition
// for trying Cholesky first.
std::vector<double> DiagonalScaling(const SymmetricBand& a, bool* positive);

class DegreeOperator {
 public:
  DegreeOperator(const Material& material, int l, const DegreeOptions& options);
  // The gravitating operator with the given point shear (the step's).
  DegreeOperator(const Material& material, int l, double eps,
                 const PointShear* pointShear)
      : DegreeOperator(material, l, DegreeOptions{eps, true, pointShear}) {}

  int Degree() const { return l_; }
  bool Gravitating() const { return gravitating_; }
  int FirstElement() const { return firstElement_; }
  int DofCount() const { return matrix_.Size(); }
  int Kd() const { return matrix_.Kd(); }
  // The dof of a component at a global mesh node, -1 where absent.
  int Dof(int node, Component c) const {
    return dof_[static_cast<size_t>(node) * 3 + static_cast<int>(c)];
  }
  int SurfaceDof(Component c) const;
  const std::vector<double>& Band() const { return matrix_.Upper(); }
  const std::vector<double>& Scale() const { return scale_; }
  bool PositiveDiagonal() const { return positive_; }
  std::vector<double> ScaledUpperBand() const { return matrix_.ScaledUpper(scale_); }
  std::vector<double> ScaledLuBand() const { return matrix_.ScaledLu(scale_); }
  // Forces of unit surface sources (src/DegreeOperator.cpp:960-997): the
  // load (a surface mass: pressure g_s and the potential's jump), a unit
  // normal traction and a unit tangential traction.
  std::vector<double> Load() const;
  // Its two parts (LoadPart::Force, LoadPart::Potential) and the unit
  // tangential load g a^2 l(l + 1) at V(a) (src/DegreeOperator.cpp:960-980).
  std::vector<double> LoadForce() const;
  std::vector<double> LoadPotential() const;
  std::vector<double> Tangential() const;
  std::vector<double> Traction() const;
  std::vector<double> TractionTangential() const;
  // The force of the external potential psi = (r/a)^p (p = l for a
  // negative power): minus the operator's gravitational couplings applied
  // to psi, the potential's own energy left out
  // (src/DegreeOperator.cpp:999-1038).
  std::vector<double> Tide(int power) const;
  // A x with the unscaled operator, nrhs column-major columns.
  std::vector<double> Apply(const std::vector<double>& x, int nrhs) const;

 private:
  const Material* material_;
  int l_;
  bool gravitating_;
  int firstElement_ = 0;
  std::vector<int> dof_;  // (global node, component)
  SymmetricBand matrix_;
  std::vector<double> scale_;
  bool positive_ = true;
};

class ToroidalOperator {
 public:
  ToroidalOperator(const Material& material, int l, double eps,
                   const PointShear* pointShear);
  int Degree() const { return l_; }
  int FirstElement() const { return firstElement_; }
  int DofCount() const { return matrix_.Size(); }
  int Kd() const { return matrix_.Kd(); }
  int Dof(int node) const { return dof_[node]; }
  // Degree 1's rigid rotations: the top dof of each solid region is pinned
  // (src/ToroidalOperator.cpp:190-239).
  const std::vector<int>& Pinned() const { return pinned_; }
  const std::vector<double>& Band() const { return matrix_.Upper(); }
  const std::vector<double>& Scale() const { return scale_; }
  // The scaled band with the pinned rows and columns those of the identity.
  std::vector<double> ScaledUpperBand() const;
  std::vector<double> Apply(const std::vector<double>& x, int nrhs) const;

 private:
  const Material* material_;
  int l_;
  int firstElement_ = 0;
  std::vector<int> dof_;  // per global node
  std::vector<int> pinned_;
  SymmetricBand matrix_;
  std::vector<double> scale_;
};

}  // namespace sl3amd
