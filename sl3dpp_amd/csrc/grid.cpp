// Surface grid, spin layouts and Wigner tables (see grid.hpp).
#include "grid.hpp"

#include <cmath>
#include <cstdlib>

#include "common.hpp"
#include "model.hpp"

namespace sl3amd {

namespace {
constexpr double kPi = 3.14159265358979323846;

// d^j_{ab}(beta) at j = max(|a|,|b|), the single term of the Wigner sum,
// in logs so that it underflows cleanly to zero near the poles.
double WignerStart(int a, int b, double beta) {
  const int j = std::max(std::abs(a), std::abs(b));
  int p, q;
  double sign = 1.0;
  if (b == j) {
    p = j + a;
    q = j - a;
  } else if (b == -j) {
    p = j - a;
    q = j + a;
    if ((a + j) & 1) sign = -1.0;
  } else if (a == j) {
    p = j + b;
    q = j - b;
    if ((j - b) & 1) sign = -1.0;
  } else {
    p = j - b;
    q = j + b;
  }
  double logv = 0.5 * (std::lgamma(2.0 * j + 1.0) - std::lgamma(p + 1.0) -
                       std::lgamma(q + 1.0));
  const double c = std::cos(0.5 * beta), s = std::sin(0.5 * beta);
  if (p > 0) {
    if (c == 0.0) return 0.0;
    logv += p * std::log(std::abs(c));
    if (c < 0.0 && (p & 1)) sign = -sign;
  }
  if (q > 0) {
    if (s == 0.0) return 0.0;
    logv += q * std::log(std::abs(s));
    if (s < 0.0 && (q & 1)) sign = -sign;
  }
  if (logv < -700.0) return 0.0;
  return sign * std::exp(logv);
}
}  // namespace

void WignerColumn(int lmax, int N, int m, double theta, double* out) {
  const int lmin = std::max(std::abs(m), std::abs(N));
  if (lmin > lmax) return;
  const double x = std::cos(theta);
  double prev = 0.0, cur;
  int j = lmin;
  if (lmin == 0) {
    out[0] = 1.0;
    if (lmax == 0) return;
    out[1] = x;
    prev = 1.0;
    cur = x;
    j = 1;
  } else {
    cur = WignerStart(m, N, theta);
    out[0] = cur;
  }
  for (; j < lmax; ++j) {
    const double jj = j;
    const double left = jj * std::sqrt((jj + 1) * (jj + 1) - N * N) *
                        std::sqrt((jj + 1) * (jj + 1) - m * m);
    const double a = (2 * jj + 1) * (jj * (jj + 1) * x - static_cast<double>(m) * N);
    const double b = (jj + 1) * std::sqrt(jj * jj - N * N) * std::sqrt(jj * jj - m * m);
    const double next = (a * cur - b * prev) / left;
    prev = cur;
    cur = next;
    out[j + 1 - lmin] = cur;
  }
}

SpinLayout SpinLayout::Make(int spin, int lmax) {
  SpinLayout s;
  s.spin = spin;
  s.lmax = lmax;
  s.mlo = spin == 0 ? 0 : -lmax;
  s.rowOffset.assign(1, 0);
  for (int m = s.mlo; m <= lmax; ++m) {
    const int lmin = s.LMin(m);
    s.rowOffset.push_back(s.rowOffset.back() + std::max(0, lmax - lmin + 1));
  }
  return s;
}

SurfaceGrid::SurfaceGrid(int lmax, double oversampling) : lmax_(lmax) {
  if (lmax < 2) throw ArgumentError("the coupling grid needs lmax >= 2");
  ntheta_ = static_cast<int>(std::ceil(oversampling * (lmax + 1) - 1e-9));
  nphi_ = 2 * ntheta_;
  std::vector<double> x, w;
  GaussLegendre(ntheta_, x, w);  // x descending: colatitudes ascend
  theta_.resize(ntheta_);
  for (int i = 0; i < ntheta_; ++i) theta_[i] = std::acos(x[i]);
  wtheta_ = w;
  for (int b = 0; b < 3; ++b) layouts_[b] = SpinLayout::Make(SpinOf(b), lmax);
}

std::vector<double> SurfaceGrid::WignerTableIn(int block, const SpinLayout& lay) const {
  if (lay.spin != layouts_[block].spin || lay.lmax < lmax_) {
    throw ArgumentError("a Wigner table's rows: the block's spin, degree >= the grid's");
  }
  const int N = lay.spin;
  std::vector<double> table(static_cast<size_t>(lay.Rows()) * ntheta_, 0.0);
  std::vector<double> column(lmax_ + 1);
  for (int m = layouts_[block].mlo; m <= lmax_; ++m) {
    const int lmin = lay.LMin(m);
    if (lmin > lmax_) continue;
    for (int i = 0; i < ntheta_; ++i) {
      WignerColumn(lmax_, N, m, theta_[i], column.data());
      for (int l = lmin; l <= lmax_; ++l) {
        const double c = std::sqrt((2.0 * l + 1.0) / (4.0 * kPi));
        table[static_cast<size_t>(lay.Row(l, m)) * ntheta_ + i] =
            c * column[l - lmin];
      }
    }
  }
  return table;
}

}  // namespace sl3amd
