// The oversampled Gauss-Legendre surface grid of the 3D coupling and the
// device layouts of its fields (host side).
//
// Grid (include/sl3dpp/SurfaceGrid.hpp:9-14,59-75): ntheta Gauss-Legendre
// colatitudes ascending, nphi = 2 ntheta uniform longitudes from zero,
// ntheta = ceil(oversampling (lmax + 1)) (194 x 388 at lmax 128, 3/2 rule).
//
// Basis: Y^N_lm = sqrt((2l+1)/4pi) d^l_{mN}(theta) e^{i m phi} (standard
// Wigner small-d, first index the order): Condon-Shortley at N = 0, and the
// Dahlen-Tromp canonical components for N != 0.
//
// Device layouts (the transforms' GEMM-friendly ones; the reference's
// l-major layouts exist only at the C-ABI boundary):
//  * spectral block of spin N: rows (m, l), m = mlo..lmax (mlo = -lmax for
//    the complex spins -2, -1; 0 for the real spin-0 block), l =
//    max(|m|,|N|)..lmax, row = rowOffset[m - mlo] + l - lmin(m); each row
//    holds `nodes` complex values (node fastest);
//  * Fourier block of spin N: F_m(theta) at ((m - mlo) * nodes + node) *
//    ntheta + theta, complex;
//  * grid fields: the reference's [node][theta][phi] (complex for spins
//    -2, -1, real for spin 0);
//  * Wigner table of spin N: c_l d^l_{mN}(theta_i) at (rowOffset[m - mlo] +
//    l - lmin(m)) * ntheta + i.
#pragma once

#include <cstddef>
#include <vector>

namespace sl3amd {

struct SpinLayout {
  int spin = 0;   // -2, -1 or 0
  int mlo = 0;    // -lmax, or 0 for the real block
  int lmax = 0;
  std::vector<int> rowOffset;  // per m - mlo, plus the total at the end
  int Rows() const { return rowOffset.back(); }
  int Orders() const { return lmax - mlo + 1; }
  int LMin(int m) const {
    const int am = m < 0 ? -m : m;
    const int an = spin < 0 ? -spin : spin;
    return am > an ? am : an;
  }
  int Row(int l, int m) const { return rowOffset[m - mlo] + l - LMin(m); }
  static SpinLayout Make(int spin, int lmax);
};

class SurfaceGrid {
 public:
  SurfaceGrid(int lmax, double oversampling);
  int LMax() const { return lmax_; }
  int NTheta() const { return ntheta_; }
  int NPhi() const { return nphi_; }
  size_t FieldSize() const { return static_cast<size_t>(ntheta_) * nphi_; }
  const std::vector<double>& CoLatitudes() const { return theta_; }
  const std::vector<double>& CoLatitudeWeights() const { return wtheta_; }
  const SpinLayout& Layout(int block) const { return layouts_[block]; }
  // block 0: spin -2, 1: spin -1, 2: spin 0 (real)
  static int SpinOf(int block) { return block == 0 ? -2 : (block == 1 ? -1 : 0); }
  // The Wigner table of a block (see the file comment).
  std::vector<double> WignerTable(int block) const { return WignerTableIn(block, layouts_[block]); }
  // The same table on this grid's colatitudes with the rows of `rows` (the
  // layout of a degree >= this grid's): the rows of degrees and orders up to
  // this grid's degree filled, the others zero.
  std::vector<double> WignerTableIn(int block, const SpinLayout& rows) const;

 private:
  int lmax_, ntheta_, nphi_;
  std::vector<double> theta_, wtheta_;
  SpinLayout layouts_[3];
};

// d^l_{mN}(theta) for l = lmin..lmax into out[l - lmin]: the three-term
// recurrence in l from the single-term start at lmin = max(|m|,|N|).
void WignerColumn(int lmax, int N, int m, double theta, double* out);

// The reference's l-major spectral indices (SurfaceGrid.hpp:116-118,197-200).
inline size_t SpinIndex(int N, int l, int m) {
  return static_cast<size_t>(l) * l - static_cast<size_t>(N * N) +
         static_cast<size_t>(m + l);
}
inline size_t RealIndex(int l, int m) {
  return static_cast<size_t>(l) * (l + 1) / 2 + static_cast<size_t>(m);
}

}  // namespace sl3amd
