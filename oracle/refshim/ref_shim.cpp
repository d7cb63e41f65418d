// ref_shim.cpp -- extern "C" entry points over the reference's own
// self-contained sources (src/BandSolvers.cpp, include/sl3dpp/StepWeights.hpp,
// include/sl3dpp/SpinComponents.hpp), compiled from where they lie under the
// reference checkout by oracle/Makefile into oracle/_ref/. Test
// infrastructure only: it lets the tests check the oracle's restatement
// against the reference's compiled code on the same inputs.
#include <complex>
#include <span>

#include "sl3dpp/SpinComponents.hpp"
#include "sl3dpp/StepWeights.hpp"
#include "sl3dpp/detail/BandSolvers.hpp"

using namespace sl3dpp;

extern "C" {

int sl3ref_band_cholesky_factor(double* ab, int n, int kd) {
  return detail::BandCholeskyFactor(
      std::span<double>(ab, static_cast<std::size_t>(kd + 1) * n), n, kd);
}
void sl3ref_band_cholesky_solve(const double* ab, int n, int kd, double* b,
                                int nrhs) {
  detail::BandCholeskySolve(
      std::span<const double>(ab, static_cast<std::size_t>(kd + 1) * n), n, kd,
      std::span<double>(b, static_cast<std::size_t>(n) * nrhs), nrhs);
}
int sl3ref_band_lu_factor(double* ab, int n, int kl, int ku, int* pivots) {
  return detail::BandLuFactor(
      std::span<double>(ab, static_cast<std::size_t>(2 * kl + ku + 1) * n), n,
      kl, ku, std::span<int>(pivots, static_cast<std::size_t>(n)));
}
void sl3ref_band_lu_solve(const double* ab, int n, int kl, int ku,
                          const int* pivots, double* b, int nrhs) {
  detail::BandLuSolve(
      std::span<const double>(ab, static_cast<std::size_t>(2 * kl + ku + 1) * n),
      n, kl, ku, std::span<const int>(pivots, static_cast<std::size_t>(n)),
      std::span<double>(b, static_cast<std::size_t>(n) * nrhs), nrhs);
}
void sl3ref_trapezoid_weights(double h, double* e, double* alpha,
                              double* beta) {
  const auto w = ExponentialTrapezoidWeights(h);
  *e = w.e;
  *alpha = w.alpha;
  *beta = w.beta;
}
double sl3ref_backward_euler_weight(double h) { return BackwardEulerWeight(h); }

void sl3ref_deviatoric_components(int l, const double* v, double r,
                                  double* A) {
  DegreeNodeValues values;
  values.U = {v[0], v[1]};
  values.V = {v[2], v[3]};
  values.W = {v[4], v[5]};
  values.rdU = {v[6], v[7]};
  values.rdV = {v[8], v[9]};
  values.rdW = {v[10], v[11]};
  const auto out = DeviatoricComponents(l, values, r);
  for (int c = 0; c < 5; ++c) {
    A[2 * c] = out[c].real();
    A[2 * c + 1] = out[c].imag();
  }
}
void sl3ref_force_rows(int l, const double* a, double r, double* rows) {
  std::array<std::complex<double>, 5> A;
  for (int c = 0; c < 5; ++c) A[c] = {a[2 * c], a[2 * c + 1]};
  const auto out = ForceRowsFromComponents(l, A, r);
  const std::complex<double> all[6] = {out.u, out.v, out.w,
                                       out.du, out.dv, out.dw};
  for (int c = 0; c < 6; ++c) {
    rows[2 * c] = all[c].real();
    rows[2 * c + 1] = all[c].imag();
  }
}
}
