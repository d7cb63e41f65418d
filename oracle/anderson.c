/*
 * anderson.c -- the oracle's Anderson mixer (Walker-Ni, FP64), restating
 * include/sl3dpp/detail/AndersonMixer.hpp:45-229. The small least-squares
 * solve dgelss(rcond = -1) is restated as the pseudo-inverse of the
 * symmetric Gram matrix through a cyclic Jacobi eigen-decomposition:
 * for a symmetric matrix the singular values are the |eigenvalues| and the
 * minimum-norm least-squares solution drops those at or below
 * eps * max|eigenvalue| (LAPACK's rank rule for rcond < 0).
 * Test infrastructure only (see oracle.h).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* m <= 16 */
int sl3o_gram_least_squares(int m, const double* gram, const double* rhs,
                            double* gamma) {
  double a[256], v[256];
  if (m < 1 || m > 16) return -1;
  for (int i = 0; i < m * m; ++i) a[i] = gram[i];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) v[i * m + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += a[p * m + q] * a[p * m + q];
    if (off == 0.0) break;
    for (int p = 0; p < m; ++p) {
      for (int q = p + 1; q < m; ++q) {
        const double apq = a[p * m + q];
        if (apq == 0.0) continue;
        const double tau = (a[q * m + q] - a[p * m + p]) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) /
                         (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = a[k * m + p], akq = a[k * m + q];
          a[k * m + p] = c * akp - s * akq;
          a[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = a[p * m + k], aqk = a[q * m + k];
          a[p * m + k] = c * apk - s * aqk;
          a[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = v[k * m + p], vkq = v[k * m + q];
          v[k * m + p] = c * vkp - s * vkq;
          v[k * m + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  double smax = 0.0;
  for (int i = 0; i < m; ++i)
    if (fabs(a[i * m + i]) > smax) smax = fabs(a[i * m + i]);
  const double cut = DBL_EPSILON * smax;
  int rank = 0;
  for (int i = 0; i < m; ++i) gamma[i] = 0.0;
  for (int k = 0; k < m; ++k) {
    const double lam = a[k * m + k];
    if (!(fabs(lam) > cut)) continue;
    ++rank;
    double proj = 0.0;
    for (int i = 0; i < m; ++i) proj += v[i * m + k] * rhs[i];
    proj /= lam;
    for (int i = 0; i < m; ++i) gamma[i] += v[i * m + k] * proj;
  }
  return rank;
}

/* The mixer state (AndersonMixer.hpp:220-228). */
typedef struct sl3o_mixer {
  int memory, count, slot;
  size_t n;
  double *xprev, *fprev, *gram;
  double **dG, **dF;
} sl3o_mixer;

sl3o_mixer* sl3o_mixer_create(int memory) {
  sl3o_mixer* mx = (sl3o_mixer*)calloc(1, sizeof(sl3o_mixer));
  mx->memory = memory;
  return mx;
}

void sl3o_mixer_destroy(sl3o_mixer* mx) {
  if (!mx) return;
  if (mx->xprev) {
    for (int i = 0; i < mx->memory; ++i) {
      free(mx->dG[i]);
      free(mx->dF[i]);
    }
    free(mx->dG);
    free(mx->dF);
    free(mx->xprev);
    free(mx->fprev);
    free(mx->gram);
  }
  free(mx);
}

/* AndersonMixer.hpp:65-219 on one contiguous vector: x <- mixed iterate
 * from x and its image gx. Returns 1 when mixed, 0 on the first call or a
 * Picard fallback. */
int sl3o_mixer_update(sl3o_mixer* mx, size_t n, double* x, const double* gx) {
  if (mx->memory <= 0) {
    memcpy(x, gx, sizeof(double) * n);
    return 0;
  }
  if (!mx->xprev) {
    mx->n = n;
    mx->dG = (double**)calloc(mx->memory, sizeof(double*));
    mx->dF = (double**)calloc(mx->memory, sizeof(double*));
    for (int i = 0; i < mx->memory; ++i) {
      mx->dG[i] = (double*)calloc(n, sizeof(double));
      mx->dF[i] = (double*)calloc(n, sizeof(double));
    }
    mx->gram = (double*)calloc((size_t)mx->memory * mx->memory, sizeof(double));
    mx->xprev = (double*)malloc(sizeof(double) * n);
    mx->fprev = (double*)malloc(sizeof(double) * n);
    for (size_t j = 0; j < n; ++j) {
      mx->xprev[j] = x[j];
      mx->fprev[j] = gx[j] - x[j];
      x[j] = gx[j];
    }
    return 0;
  }
  const int slot = mx->slot;
  const int m = mx->count + 1 < mx->memory ? mx->count + 1 : mx->memory;
  double column[16] = {0}, rhs[16] = {0};
  double* dGs = mx->dG[slot];
  double* dFs = mx->dF[slot];
  for (size_t j = 0; j < n; ++j) {
    const double f = gx[j] - x[j];
    const double dx = x[j] - mx->xprev[j];
    dFs[j] = f - mx->fprev[j];
    dGs[j] = dx + dFs[j];
    mx->xprev[j] = x[j];
    mx->fprev[j] = f;
    for (int i = 0; i < m; ++i) {
      column[i] += mx->dF[i][j] * dFs[j];
      rhs[i] += mx->dF[i][j] * f;
    }
  }
  mx->count = m;
  mx->slot = (slot + 1) % mx->memory;
  for (int i = 0; i < m; ++i) {
    mx->gram[slot * mx->memory + i] = column[i];
    mx->gram[i * mx->memory + slot] = column[i];
  }
  double g[256], gamma[16];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) g[i * m + j] = mx->gram[i * mx->memory + j];
  sl3o_gram_least_squares(m, g, rhs, gamma);
  for (size_t j = 0; j < n; ++j) {
    double value = gx[j];
    for (int i = 0; i < m; ++i) value -= gamma[i] * mx->dG[i][j];
    x[j] = value;
  }
  return 1;
}
