/*
 * pointwise.c -- the oracle's step weights and spin algebra (FP64).
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>

#include "oracle.h"

/* include/sl3dpp/StepWeights.hpp:98-108: (1 - e) / h, with the 5-term
 * alternating series below |h| = 1e-3. */
double sl3o_phi1(double h, double e) {
  if (fabs(h) < 1e-3) {
    return 1.0 - h / 2.0 + h * h / 6.0 - h * h * h / 24.0 +
           h * h * h * h / 120.0;
  }
  return (1.0 - e) / h;
}

/* include/sl3dpp/StepWeights.hpp:132-137: e = exp(-h), alpha = phi1 - e,
 * beta = 1 - phi1. */
void sl3o_trapezoid_weights(double h, double* e, double* alpha,
                            double* beta) {
  const double ee = exp(-h);
  const double p = sl3o_phi1(h, ee);
  *e = ee;
  *alpha = p - ee;
  *beta = 1.0 - p;
}

/* include/sl3dpp/StepWeights.hpp:141-144 */
double sl3o_backward_euler_weight(double h) { return 1.0 / (1.0 + h); }

/* The lateral inverse relaxation time of Viscoelastic3D's constructor
 * (src/Viscoelastic3D.cpp:250-283): exp(-delta) sampled on each memory
 * node's grid scales the node's reference inverse time,
 * itau(x) = itau_ref(node) exp(-delta(x)); per_node values per node, delta
 * NULL for none. */
void sl3o_lateral_itau(int64_t nodes, int64_t per_node, const double* itau_ref_node,
                       const double* delta, double* itau) {
  for (int64_t node = 0; node < nodes; ++node) {
    for (int64_t k = 0; k < per_node; ++k) {
      const int64_t x = node * per_node + k;
      itau[x] = itau_ref_node[node] * (delta != NULL ? exp(-delta[x]) : 1.0);
    }
  }
}

/* complex helpers on (re, im) pairs */
typedef struct {
  double re, im;
} cplx;
static cplx C(double re, double im) {
  cplx z = {re, im};
  return z;
}
static cplx cadd(cplx a, cplx b) { return C(a.re + b.re, a.im + b.im); }
static cplx csub(cplx a, cplx b) { return C(a.re - b.re, a.im - b.im); }
static cplx cscale(double s, cplx a) { return C(s * a.re, s * a.im); }
static cplx ctimes_i(cplx a) { return C(-a.im, a.re); }
static cplx cdiv(cplx a, double d) { return C(a.re / d, a.im / d); }

/* include/sl3dpp/SpinComponents.hpp:53-72 */
void sl3o_deviatoric_components(int l, const double* v, double r, double* A) {
  const cplx U = C(v[0], v[1]), V = C(v[2], v[3]), W = C(v[4], v[5]);
  const cplx rdU = C(v[6], v[7]), rdV = C(v[8], v[9]), rdW = C(v[10], v[11]);
  const double zeta2 = l * (l + 1.0);
  const double zeta = sqrt(zeta2);
  cplx out[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
  /* complex / real divides each part, as std::complex does */
  out[2] = cdiv(cadd(csub(rdU, U), cscale(0.5 * zeta2, V)), 3.0 * r);
  if (l >= 1) {
    const cplx tangential = cadd(csub(rdV, V), U);
    const cplx twist = csub(W, rdW);
    const double den = 2.0 * sqrt(2.0) * r;
    out[1] = cdiv(cscale(zeta, cadd(tangential, ctimes_i(twist))), den);
    out[3] = cdiv(cscale(zeta, csub(tangential, ctimes_i(twist))), den);
  }
  if (l >= 2) {
    const double f = 0.5 * zeta * sqrt(zeta2 - 2.0) / r;
    out[0] = cscale(f, csub(V, ctimes_i(W)));
    out[4] = cscale(f, cadd(V, ctimes_i(W)));
  }
  for (int c = 0; c < 5; ++c) {
    A[2 * c] = out[c].re;
    A[2 * c + 1] = out[c].im;
  }
}

/* include/sl3dpp/SpinComponents.hpp:124-145 */
void sl3o_force_rows(int l, const double* a, double r, double* rows) {
  cplx A[5];
  for (int c = 0; c < 5; ++c) A[c] = C(a[2 * c], a[2 * c + 1]);
  const double zeta2 = l * (l + 1.0);
  const double zeta = sqrt(zeta2);
  const double root = sqrt(zeta2 - 2.0 > 0.0 ? zeta2 - 2.0 : 0.0);
  const double s2 = sqrt(2.0);
  const cplx p13 = cadd(A[1], A[3]), m13 = csub(A[1], A[3]);
  const cplx p04 = cadd(A[0], A[4]), m04 = csub(A[0], A[4]);
  cplx out[6];
  out[0] = csub(cscale(s2 * zeta, p13), cscale(4.0, A[2]));
  out[1] = csub(cadd(cscale(zeta * root, p04), cscale(2.0 * zeta2, A[2])),
                cscale(s2 * zeta, p13));
  out[2] = ctimes_i(csub(cscale(zeta * root, m04), cscale(s2 * zeta, m13)));
  out[3] = cscale(4.0 * r, A[2]);
  out[4] = cscale(s2 * zeta * r, p13);
  out[5] = ctimes_i(cscale(s2 * zeta * r, m13));
  for (int c = 0; c < 6; ++c) {
    rows[2 * c] = out[c].re;
    rows[2 * c + 1] = out[c].im;
  }
}
