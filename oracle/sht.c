/*
 * sht.c -- the oracle's generalised spherical-harmonic transforms (FP64).
 *
 * Basis (include/sl3dpp/SurfaceGrid.hpp:9-14,191-200; docs/design.md:
 * 246-252): Y^N_lm(theta, phi) = sqrt((2l+1)/4pi) d^l_{mN}(theta) e^{i m phi}
 * with d the standard (Wigner / Varshalovich) small-d matrix, first index
 * the order m. For N = 0 this is the complex orthonormal Condon-Shortley
 * harmonic the reference names; for N != 0 it makes the canonical
 * components of the gradient of a spin-0 field, in the Dahlen-Tromp basis
 * e_- = (theta^ - i phi^)/sqrt2, e_+ = -(theta^ + i phi^)/sqrt2, equal to
 * (zeta/sqrt2) Y^{-1}_lm and (zeta/sqrt2) Y^{+1}_lm, the identity the
 * reference's force rows and strain components assume
 * (include/sl3dpp/SpinComponents.hpp:53-72).
 * Forward transform: f_lm = int f conj(Y^N_lm) dOmega by Gauss-Legendre in
 * colatitude and the uniform rule in longitude.
 *
 * Direct sums with a precomputed table: a restatement for checking, not a
 * fast transform. Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* d^{j0}_{a b}(beta) at j0 = max(|a|, |b|): the single-term case of the
 * Wigner sum, evaluated in logs (underflow to zero near the poles). */
static double wigner_start(int a, int b, double beta) {
  const int j = abs(a) > abs(b) ? abs(a) : abs(b);
  double sign = 1.0;
  int p;  /* power of cos(beta/2) */
  int q;  /* power of sin(beta/2) */
  double lognorm;
  if (b == j) {  /* d^j_{a j} = sqrt(C(2j, j+a)) c^{j+a} s^{j-a} */
    p = j + a;
    q = j - a;
  } else if (b == -j) {  /* d^j_{a,-j} = (-1)^{a+j} d^j_{-a, j} */
    p = j - a;
    q = j + a;
    if ((a + j) & 1) sign = -1.0;
  } else if (a == j) {  /* d^j_{j b} = (-1)^{j-b} d^j_{b j} */
    p = j + b;
    q = j - b;
    if ((j - b) & 1) sign = -1.0;
  } else { /* a == -j: d^j_{-j, b} = d^j_{-b, j} */
    p = j - b;
    q = j + b;
  }
  lognorm = 0.5 * (lgamma(2.0 * j + 1.0) - lgamma((double)p + 1.0) -
                   lgamma((double)q + 1.0));
  const double c = cos(0.5 * beta), s = sin(0.5 * beta);
  double logv = lognorm;
  if (p > 0) {
    if (c == 0.0) return 0.0;
    logv += p * log(fabs(c));
    if (c < 0.0 && (p & 1)) sign = -sign;
  }
  if (q > 0) {
    if (s == 0.0) return 0.0;
    logv += q * log(fabs(s));
    if (s < 0.0 && (q & 1)) sign = -sign;
  }
  if (logv < -700.0) return 0.0;
  return sign * exp(logv);
}

/* d^l_{m N}(theta), l = lmin..lmax, lmin = max(|m|,|N|), by the three-term
 * recurrence in l (Edmonds 1957, eq. 4.1.?; any standard text):
 *   j sqrt((j+1)^2-N^2) sqrt((j+1)^2-m^2) d^{j+1}
 *     = (2j+1)(j(j+1) cos - m N) d^j - (j+1) sqrt(j^2-N^2) sqrt(j^2-m^2) d^{j-1}.
 */
void sl3o_wigner_d_column(int lmax, int N, int m, double theta, double* out) {
  const int lmin = abs(m) > abs(N) ? abs(m) : abs(N);
  if (lmin > lmax) return;
  const double x = cos(theta);
  double prev = 0.0;
  double cur;
  int j = lmin;
  if (lmin == 0) {
    out[0] = 1.0;
    if (lmax >= 1) out[1] = x;
    prev = 1.0;
    cur = x;
    j = 1;
    if (lmax < 1) return;
  } else {
    cur = wigner_start(m, N, theta);
    out[0] = cur;
  }
  for (; j < lmax; ++j) {
    const double jj = (double)j;
    const double left = jj * sqrt(((jj + 1) * (jj + 1) - N * N)) *
                        sqrt(((jj + 1) * (jj + 1) - m * m));
    const double a = (2 * jj + 1) * (jj * (jj + 1) * x - (double)m * N);
    const double b = (jj + 1) * sqrt(jj * jj - N * N) * sqrt(jj * jj - m * m);
    const double next = (a * cur - b * prev) / left;
    prev = cur;
    cur = next;
    out[j + 1 - lmin] = cur;
  }
}

/* Gauss-Legendre nodes (descending in x, so ascending colatitude) and
 * weights by Newton's method on P_n. */
void sl3o_gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) {
        p1 = z;
        p0 = 1.0;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) {
        /* one more evaluation for the weight at the converged node */
        p0 = 1.0;
        p1 = z;
        for (int k = 2; k <= n; ++k) {
          const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        break;
      }
    }
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

static size_t spin_index(int N, int l, int m) {
  return (size_t)l * l - (size_t)(N * N) + (size_t)(m + l);
}
static size_t real_index(int l, int m) {
  return (size_t)l * (l + 1) / 2 + (size_t)m;
}

/* One field's synthesis: F_m(theta_i) = sum_l sqrt((2l+1)/4pi) d^l_{mN} f_lm,
 * then f(theta_i, phi_j) = sum_m F_m e^{i m phi_j} (SurfaceGrid.hpp:176-185,
 * 242-253). */
void sl3o_synthesise(int lmax, int N, int real_field, int ntheta, int nphi,
                     const double* theta, const double* coef, double* field) {
  const int mlo = real_field ? 0 : -lmax;
  double* d = (double*)malloc(sizeof(double) * (lmax + 1));
  double* F = (double*)malloc(sizeof(double) * 2 * (2 * lmax + 1));
  for (int i = 0; i < ntheta; ++i) {
    for (int m = mlo; m <= lmax; ++m) {
      double re = 0.0, im = 0.0;
      const int lmin = abs(m) > abs(N) ? abs(m) : abs(N);
      if (lmin <= lmax) {
        sl3o_wigner_d_column(lmax, N, m, theta[i], d);
        for (int l = lmin; l <= lmax; ++l) {
          const double c = sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) * d[l - lmin];
          const size_t k = real_field ? real_index(l, m) : spin_index(N, l, m);
          re += c * coef[2 * k];
          im += c * coef[2 * k + 1];
        }
      }
      F[2 * (m + lmax)] = re;
      F[2 * (m + lmax) + 1] = im;
    }
    for (int j = 0; j < nphi; ++j) {
      const double phi = 2.0 * M_PI * j / nphi;
      double re = 0.0, im = 0.0;
      for (int m = mlo; m <= lmax; ++m) {
        const double c = cos(m * phi), s = sin(m * phi);
        const double fr = F[2 * (m + lmax)], fi = F[2 * (m + lmax) + 1];
        double tr = fr * c - fi * s, ti = fr * s + fi * c;
        if (real_field && m > 0) {
          tr *= 2.0;
          ti = 0.0;
        }
        re += tr;
        im += ti;
      }
      if (real_field) {
        field[(size_t)i * nphi + j] = re;
      } else {
        field[2 * ((size_t)i * nphi + j)] = re;
        field[2 * ((size_t)i * nphi + j) + 1] = im;
      }
    }
  }
  free(d);
  free(F);
}

/* One field's analysis: F_m(theta_i) = (2pi/nphi) sum_j f e^{-i m phi_j},
 * f_lm = sum_i w_i sqrt((2l+1)/4pi) d^l_{mN}(theta_i) F_m(theta_i)
 * (SurfaceGrid.hpp:164-173, 217-227). */
void sl3o_analyse(int lmax, int N, int real_field, int ntheta, int nphi,
                  const double* theta, const double* wtheta,
                  const double* field, double* coef) {
  const int mlo = real_field ? 0 : -lmax;
  const size_t ncoef = real_field ? (size_t)(lmax + 1) * (lmax + 2) / 2
                                  : (size_t)(lmax + 1) * (lmax + 1) - N * N;
  memset(coef, 0, sizeof(double) * 2 * ncoef);
  double* d = (double*)malloc(sizeof(double) * (lmax + 1));
  for (int i = 0; i < ntheta; ++i) {
    for (int m = mlo; m <= lmax; ++m) {
      const int lmin = abs(m) > abs(N) ? abs(m) : abs(N);
      if (lmin > lmax) continue;
      double re = 0.0, im = 0.0;
      for (int j = 0; j < nphi; ++j) {
        const double phi = 2.0 * M_PI * j / nphi;
        const double c = cos(m * phi), s = -sin(m * phi);
        if (real_field) {
          const double f = field[(size_t)i * nphi + j];
          re += f * c;
          im += f * s;
        } else {
          const double fr = field[2 * ((size_t)i * nphi + j)];
          const double fi = field[2 * ((size_t)i * nphi + j) + 1];
          re += fr * c - fi * s;
          im += fr * s + fi * c;
        }
      }
      const double scale = 2.0 * M_PI / nphi * wtheta[i];
      re *= scale;
      im *= scale;
      sl3o_wigner_d_column(lmax, N, m, theta[i], d);
      for (int l = lmin; l <= lmax; ++l) {
        const double c = sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) * d[l - lmin];
        const size_t k = real_field ? real_index(l, m) : spin_index(N, l, m);
        coef[2 * k] += c * re;
        coef[2 * k + 1] += c * im;
      }
    }
  }
  free(d);
}
