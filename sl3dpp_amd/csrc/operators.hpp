// The per-degree radial systems of the 3D step, host side: dof maps, band
// assembly and the symmetric diagonal scaling that precedes the GPU
// factorisation. The factor and the solves run on the GPU (band.hip); the
// class names and what they hold follow the reference's DegreeOperator
// (include/sl3dpp/DegreeOperator.hpp:94-384) and ToroidalOperator
// (include/sl3dpp/ToroidalOperator.hpp:58-176), gravitating, undistributed.
#pragma once

#include <vector>

#include "model.hpp"

namespace sl3amd {

// A shear modulus at the memory points the assembly integrates there
// (include/sl3dpp/MemoryPoints.hpp:144-154), (element, point).
struct PointShear {
  const MemoryPoints* points = nullptr;
  const std::vector<double>* modulus = nullptr;
};

enum class Component { U = 0, V = 1, Phi = 2 };

class DegreeOperator {
 public:
  DegreeOperator(const Material& material, int l, double eps,
                 const PointShear* pointShear);
  int Degree() const { return l_; }
  int FirstElement() const { return firstElement_; }
  int DofCount() const { return ndof_; }
  int Kd() const { return kd_; }
  int Dof(int node, Component c) const {
    return dof_[static_cast<size_t>(node) * 3 + static_cast<int>(c)];
  }
  int SurfaceDof(Component c) const;
  // The assembled (unscaled) upper band, (kd + 1) x ndof column-major.
  const std::vector<double>& Band() const { return band_; }
  // Factorise's diagonal scaling s_j = 1 / sqrt|A_jj| and whether every
  // diagonal is positive (the Cholesky attempt).
  const std::vector<double>& Scale() const { return scale_; }
  bool PositiveDiagonal() const { return positive_; }
  // The scaled upper band for the Cholesky attempt, and the scaled full
  // band in LU storage (3 kd + 1 rows) for the fallback
  // (src/DegreeOperator.cpp:781-846).
  std::vector<double> ScaledUpperBand() const;
  std::vector<double> ScaledLuBand() const;
  // The unit surface load's force, LoadPart::Both
  // (src/DegreeOperator.cpp:960-970).
  std::vector<double> Load() const;
  // The force of the external potential (r/a)^p, p = power, or p = l for a
  // negative power (src/DegreeOperator.cpp:999-1038).
  std::vector<double> Tide(int power) const;
  // A x with the unscaled band, nrhs column-major columns
  // (src/DegreeOperator.cpp:480-505).
  std::vector<double> Apply(const std::vector<double>& x, int nrhs) const;

 private:
  void BuildDofMap();
  void Assemble(const PointShear* pointShear);
  double& BandAt(int i, int j) {
    return band_[static_cast<size_t>(j) * (kd_ + 1) + (kd_ + i - j)];
  }

  const Material* material_;
  int l_;
  int firstElement_ = 0;
  int kd_ = 0;
  int ndof_ = 0;
  std::vector<int> dof_;
  std::vector<double> band_;
  std::vector<double> scale_;
  bool positive_ = true;
};

class ToroidalOperator {
 public:
  ToroidalOperator(const Material& material, int l, double eps,
                   const PointShear* pointShear);
  int Degree() const { return l_; }
  int FirstElement() const { return firstElement_; }
  int DofCount() const { return ndof_; }
  int Kd() const { return kd_; }
  int Dof(int node) const { return dof_[node]; }
  const std::vector<int>& Pinned() const { return pinned_; }
  const std::vector<double>& Band() const { return band_; }
  const std::vector<double>& Scale() const { return scale_; }
  // The scaled band with the pinned rows and columns the identity's
  // (src/ToroidalOperator.cpp:190-239).
  std::vector<double> ScaledUpperBand() const;
  std::vector<double> Apply(const std::vector<double>& x, int nrhs) const;

 private:
  const Material* material_;
  int l_;
  int firstElement_ = 0;
  int kd_ = 0;
  int ndof_ = 0;
  std::vector<int> dof_;
  std::vector<std::pair<int, int>> regions_;
  std::vector<int> pinned_;
  std::vector<double> band_;
  std::vector<double> scale_;
};

}  // namespace sl3amd
