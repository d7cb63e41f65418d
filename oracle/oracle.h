/*
 * oracle.h -- CPU restatement (FP64, plain C11 + OpenMP) of the 3D
 * lateral-viscosity step's data-parallel path of sl3dpp.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is
 * compared against and the timed CPU baseline; nothing in the product
 * (sl3dpp_amd/) links, loads or calls it. Only tests/, the smoke check of
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it.
 *
 * Every function names the reference file:line whose algorithm it restates.
 * Pins (what checks each function against something independent):
 *   band solvers        golden LAPACK vectors (scipy dpbtrf/dpbtrs/dgbtrf/
 *                       dgbtrs, tests/golden) and oracle/_ref (the
 *                       reference's own src/BandSolvers.cpp compiled here)
 *   step weights        oracle/_ref (StepWeights.hpp) and mpmath closed form
 *   spin algebra        oracle/_ref (SpinComponents.hpp)
 *   Wigner d / GSH      sympy Wigner d at small l, scipy sph_harm (N = 0),
 *                       round trips, the D&T gradient identity on the grid
 *   Anderson LS         scipy dgelss on random Gram systems
 *   step                delta = 0 reduces to the slaved 1D solve; converged
 *                       fixed point satisfies its own equation
 */
#ifndef SL3_ORACLE_H
#define SL3_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- band solvers: src/BandSolvers.cpp:23-139 ------------------------- */
int sl3o_band_cholesky_factor(double* ab, int n, int kd);
void sl3o_band_cholesky_solve(const double* ab, int n, int kd, double* b,
                              int nrhs);
int sl3o_band_lu_factor(double* ab, int n, int kl, int ku, int* pivots);
void sl3o_band_lu_solve(const double* ab, int n, int kl, int ku,
                        const int* pivots, double* b, int nrhs);

/* ---- step weights: include/sl3dpp/StepWeights.hpp:98-144 --------------- */
double sl3o_phi1(double h, double e);
void sl3o_trapezoid_weights(double h, double* e, double* alpha, double* beta);
double sl3o_backward_euler_weight(double h);
void sl3o_lateral_itau(int64_t nodes, int64_t per_node, const double* itau_ref_node,
                       const double* delta, double* itau);

/* ---- spin algebra: include/sl3dpp/SpinComponents.hpp:53-145 ------------ */
/* values: U, V, W, rdU, rdV, rdW as (re, im) pairs (12 doubles);
 * A: five components as (re, im) pairs (10 doubles). */
void sl3o_deviatoric_components(int l, const double* values, double r,
                                double* A);
/* rows: u, v, w, du, dv, dw as (re, im) pairs (12 doubles). */
void sl3o_force_rows(int l, const double* A, double r, double* rows);

/* ---- Wigner d and the GSH transforms ------------------------------------ */
/* d^l_{m N}(theta) (standard Wigner convention, first index the order m) for
 * l = max(|m|,|N|)..lmax into out[l - lmin]. */
void sl3o_wigner_d_column(int lmax, int N, int m, double theta, double* out);
/* Gauss-Legendre nodes on [-1,1] (descending, so colatitudes ascend) and
 * weights. */
void sl3o_gauss_legendre(int n, double* x, double* w);
/* Spin-N synthesis / analysis on the (ntheta, nphi) grid, one field.
 * Complex spin fields: coefficients at l*l - N*N + m + l (l >= |N|).
 * Real spin-0 fields (real_field = 1): coefficients at l(l+1)/2 + m, m >= 0;
 * field values real (stride 1 doubles). Complex field values are (re, im).
 * SurfaceGrid.hpp:141-252 layouts. */
void sl3o_synthesise(int lmax, int N, int real_field, int ntheta, int nphi,
                     const double* theta, const double* coef, double* field);
void sl3o_analyse(int lmax, int N, int real_field, int ntheta, int nphi,
                  const double* theta, const double* wtheta,
                  const double* field, double* coef);

/* ---- Anderson least squares: AndersonMixer.hpp:186-209 ------------------ */
/* gamma = pinv(G) rhs for the symmetric m x m Gram matrix, singular values
 * below machine precision times the largest dropped (dgelss, rcond -1).
 * Returns the rank. */
int sl3o_gram_least_squares(int m, const double* gram, const double* rhs,
                            double* gamma);

#ifdef __cplusplus
}
#endif
#endif
