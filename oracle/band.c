/*
 * band.c -- banded Cholesky / LU factorisations and solves (FP64), the
 * oracle's restatement of src/BandSolvers.cpp:23-139 (LAPACK dpbtf2/dpbtrs,
 * dgbtf2/dgbtrs unblocked algorithms, LAPACK band storage, 0-based pivots).
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stddef.h>

#include "oracle.h"

/* A(i, j) of a band whose diagonal sits in row `diag` of columns `ld` long. */
#define BAND(ab, ld, diag, i, j) \
  (ab)[(size_t)(j) * (size_t)(ld) + (size_t)((diag) + (i) - (j))]

/* src/BandSolvers.cpp:23-45: A = U^T U column by column; returns 0 or the
 * 1-based index of the first non-positive (or NaN) pivot. */
int sl3o_band_cholesky_factor(double* ab, int n, int kd) {
  const int ld = kd + 1;
  double row[64];
  for (int j = 0; j < n; ++j) {
    const double d = BAND(ab, ld, kd, j, j);
    if (!(d > 0.0)) return j + 1;
    const double ujj = sqrt(d);
    BAND(ab, ld, kd, j, j) = ujj;
    const int kn = kd < n - 1 - j ? kd : n - 1 - j;
    const double inv = 1.0 / ujj;
    for (int c = 1; c <= kn; ++c) {
      BAND(ab, ld, kd, j, j + c) *= inv;
      row[c] = BAND(ab, ld, kd, j, j + c);
    }
    /* trailing update within the band: A(j+r, j+c) -= row[r] row[c] */
    for (int c = 1; c <= kn; ++c) {
      const double uc = row[c];
      for (int r = 1; r <= c; ++r) {
        BAND(ab, ld, kd, j + r, j + c) -= row[r] * uc;
      }
    }
  }
  return 0;
}

/* src/BandSolvers.cpp:47-70: U^T y = b forward, U x = y backward. */
void sl3o_band_cholesky_solve(const double* ab, int n, int kd, double* b,
                              int nrhs) {
  const int ld = kd + 1;
  for (int k = 0; k < nrhs; ++k) {
    double* x = b + (size_t)k * (size_t)n;
    for (int j = 0; j < n; ++j) {
      double s = x[j];
      for (int i = (j - kd > 0 ? j - kd : 0); i < j; ++i) {
        s -= BAND(ab, ld, kd, i, j) * x[i];
      }
      x[j] = s / BAND(ab, ld, kd, j, j);
    }
    for (int j = n - 1; j >= 0; --j) {
      x[j] /= BAND(ab, ld, kd, j, j);
      const double xj = x[j];
      for (int i = (j - kd > 0 ? j - kd : 0); i < j; ++i) {
        x[i] -= BAND(ab, ld, kd, i, j) * xj;
      }
    }
  }
}

/* src/BandSolvers.cpp:72-115: P L U with partial pivoting on the band
 * (kl sub-, ku superdiagonals, 2kl+ku+1 rows with kl rows of fill room). */
int sl3o_band_lu_factor(double* ab, int n, int kl, int ku, int* pivots) {
  const int kv = kl + ku;
  const int ld = 2 * kl + ku + 1;
  const int lim = kv < n ? kv : n;
  for (int j = ku + 1; j < lim; ++j) {
    for (int i = 0; i < j - ku; ++i) BAND(ab, ld, kv, i, j) = 0.0;
  }
  int info = 0;
  int ju = 0;
  for (int j = 0; j < n; ++j) {
    if (j + kv < n) {
      for (int i = j; i < j + kl; ++i) BAND(ab, ld, kv, i, j + kv) = 0.0;
    }
    const int km = kl < n - 1 - j ? kl : n - 1 - j;
    int p = j;
    for (int i = j + 1; i <= j + km; ++i) {
      if (fabs(BAND(ab, ld, kv, i, j)) > fabs(BAND(ab, ld, kv, p, j))) p = i;
    }
    pivots[j] = p;
    if (BAND(ab, ld, kv, p, j) == 0.0) {
      if (info == 0) info = j + 1;
      continue;
    }
    {
      const int reach = p + ku < n - 1 ? p + ku : n - 1;
      if (reach > ju) ju = reach;
    }
    if (p != j) {
      for (int c = j; c <= ju; ++c) {
        const double t = BAND(ab, ld, kv, p, c);
        BAND(ab, ld, kv, p, c) = BAND(ab, ld, kv, j, c);
        BAND(ab, ld, kv, j, c) = t;
      }
    }
    if (km > 0) {
      const double inv = 1.0 / BAND(ab, ld, kv, j, j);
      for (int i = j + 1; i <= j + km; ++i) BAND(ab, ld, kv, i, j) *= inv;
      for (int c = j + 1; c <= ju; ++c) {
        const double ujc = BAND(ab, ld, kv, j, c);
        for (int i = j + 1; i <= j + km; ++i) {
          BAND(ab, ld, kv, i, c) -= BAND(ab, ld, kv, i, j) * ujc;
        }
      }
    }
  }
  return info;
}

/* src/BandSolvers.cpp:117-139: interchanges and L eliminations, then U
 * (kl + ku superdiagonals) backward. */
void sl3o_band_lu_solve(const double* ab, int n, int kl, int ku,
                        const int* pivots, double* b, int nrhs) {
  const int kv = kl + ku;
  const int ld = 2 * kl + ku + 1;
  for (int k = 0; k < nrhs; ++k) {
    double* x = b + (size_t)k * (size_t)n;
    for (int j = 0; j + 1 < n; ++j) {
      const int p = pivots[j];
      if (p != j) {
        const double t = x[p];
        x[p] = x[j];
        x[j] = t;
      }
      const double xj = x[j];
      const int lm = kl < n - 1 - j ? kl : n - 1 - j;
      for (int i = j + 1; i <= j + lm; ++i) x[i] -= BAND(ab, ld, kv, i, j) * xj;
    }
    for (int j = n - 1; j >= 0; --j) {
      x[j] /= BAND(ab, ld, kv, j, j);
      const double xj = x[j];
      for (int i = (j - kv > 0 ? j - kv : 0); i < j; ++i) {
        x[i] -= BAND(ab, ld, kv, i, j) * xj;
      }
    }
  }
}
