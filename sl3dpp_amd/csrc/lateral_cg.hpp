// Preconditioned conjugate gradients on the step's laterally weighted
// degree systems (detail::LateralConjugateGradients,
// include/sl3dpp/detail/LateralSolve.hpp:110-183, driven by
// Viscoelastic3D::OpenStep::Forces/Solve, src/Viscoelastic3D.cpp:1448-1496),
// taken on the memory dofs.
//
// The system is (K - L) x = f with K the reference degree systems and
// L = Pᵀ F A W S P the lateral force of the strain at the memory dofs P.
// Started from x0 = K⁻¹ f, every residual r = f - A x is Pᵀ rho with rho on
// the memory dofs, and every search direction p = K⁻¹ Pᵀ pi. So with
// G = P K⁻¹ Pᵀ (the memory response) the iteration needs only memory-dof
// vectors: z = K⁻¹ r enters as P z = G rho and <z, r> = <G rho, rho>, and
// <p, A p> = <G pi, pi - L G pi>. One iteration is one response GEMM and one
// application of L (strain, synthesis, the weight, analysis, force).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "coupling.hpp"

namespace sl3amd {

// One workgroup of the pairing: a 64-row x 16-column tile of one task's
// memory-dof block.
struct PairWork {
  int task = 0;
  int r0 = 0;
  int c0 = 0;
  int pad = 0;
};

std::vector<PairWork> PairWorkList(const std::vector<DegreeTask>& tasks);

// The channels' unit responses on the memory dofs with their amplitudes
// (the strain kernel's layout), times scale; amp null: none.
struct ChannelTerms {
  const double* units = nullptr;
  const double* amp = nullptr;
  const double* tideUnits = nullptr;
  const double* tideAmp = nullptr;
  double scale = 1.0;
};

// LateralPair (include/sl3dpp/detail/LateralSolve.hpp:66-91) of a solution
// on the memory dofs with a force there: sum over tasks, rows < nm and
// columns < cols of w(m) x[c][r] (f - g)[r][c], w = 1 at m = 0 and 2 above
// (x: the response layout, column-major; f, g: the force layout,
// row-major; g null for none). The channel terms, when given, are added to
// x of the spheroidal tasks (times their scale). The workgroups' partial
// sums are reduced in a fixed order into result[0] (device): the same value
// on every run.
void LaunchLateralPair(const DegreeTask* tasks_dev, const PairWork* work_dev, int workCount,
                       const double* x, const double* f, const double* g,
                       const ChannelTerms& channels, double* partials, double* result,
                       hipStream_t stream);

// After the step length alpha: phi += alpha pi, y += alpha Gpi,
// rho += alpha (q - pi), over n doubles of each.
void LaunchCgAdvance(double* phi, double* y, double* rho, const double* pi,
                     const double* Gpi, const double* q, double alpha, size_t n,
                     hipStream_t stream);

// The next direction: pi = rho + beta pi, Gpi = Grho + beta Gpi.
void LaunchCgDirection(double* pi, double* Gpi, const double* rho, const double* Grho,
                       double beta, size_t n, hipStream_t stream);

}  // namespace sl3amd
