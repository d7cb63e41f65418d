/*
 * sl3dpp_amd.h -- the C ABI of the MI355X (gfx950) implementation of the
 * data-parallel core of sl3dpp's 3D lateral-viscosity step.
 *
 * Plain C: pointers, sizes and status codes; no HIP or torch types. Device
 * pointers are marked `_dev`; everything else is host memory borrowed for
 * the duration of the call. Calls run on the library's stream (see
 * sl3_synchronize) and are asynchronous unless they return host data.
 *
 * Status codes: 0 success; < 0 argument or runtime failure (message from
 * sl3_last_error); > 0 a numerical failure the reference reports by
 * exception (a factorisation that failed outright).
 */
#ifndef SL3DPP_AMD_H
#define SL3DPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SL3_OK = 0,
  SL3_EINVAL = -1,   /* std::invalid_argument in the reference */
  SL3_EHIP = -2,     /* a HIP runtime failure */
  SL3_ENOMEM = -3,
  SL3_EINTERNAL = -4,
  SL3_ENUMERIC = 1   /* std::runtime_error: a factorisation failed */
};

/* The message of the last failed call on the calling thread. */
const char* sl3_last_error(void);
/* Library version string. */
const char* sl3_version(void);
/* Waits for every call issued so far on the library's stream. */
int sl3_synchronize(void);

/* ===================================================================== *
 * Banded Cholesky and LU, FP64, LAPACK band storage (column-major bands, *
 * 0-based pivots). Replace sl3dpp::detail::BandCholeskyFactor,          *
 * BandCholeskySolve, BandLuFactor and BandLuSolve                       *
 * (include/sl3dpp/detail/BandSolvers.hpp:24-49, src/BandSolvers.cpp).   *
 * ===================================================================== */

/* One system, as BandCholeskyFactor (BandSolvers.hpp:24-32): factorises
 * the (kd+1) x n upper band at ab_dev in place; *info = 0 or the 1-based
 * index of the first non-positive leading minor. */
int sl3_band_cholesky_factor(double* ab_dev, int n, int kd, int* info);
/* As BandCholeskySolve (BandSolvers.hpp:34-37): B (n x nrhs column-major)
 * at b_dev overwritten by X. */
int sl3_band_cholesky_solve(const double* ab_dev, int n, int kd,
                            double* b_dev, int nrhs);
/* As BandLuFactor (BandSolvers.hpp:39-45): (2kl+ku+1) x n band, pivots
 * (0-based) to piv_dev; *info = 0 or the 1-based first zero pivot. */
int sl3_band_lu_factor(double* ab_dev, int n, int kl, int ku, int* piv_dev,
                       int* info);
/* As BandLuSolve (BandSolvers.hpp:47-49). */
int sl3_band_lu_solve(const double* ab_dev, int n, int kl, int ku,
                      const int* piv_dev, double* b_dev, int nrhs);

/* Batched forms: `count` independent systems in one launch each (the
 * per-degree loops of DegreeOperator::Factorise and Viscoelastic3D::
 * SolveForces, src/DegreeOperator.cpp:775-901, src/Viscoelastic3D.cpp:
 * 876-978). Per system s (host arrays): n[s], kd[s] (Cholesky) or kl[s],
 * ku[s] (LU), the offset in doubles of its band in ab_dev, of its pivots
 * in piv_dev, of its right-hand sides in b_dev, their count nrhs[s] and
 * their layout: element (i, k) at b_dev[b_off[s] + i*row_stride[s] +
 * k*col_stride[s]] (column-major: row_stride 1, col_stride n). info_dev
 * receives each system's info. */
int sl3_band_cholesky_factor_batched(int count, const int* n, const int* kd,
                                     const int64_t* ab_off, double* ab_dev,
                                     int* info_dev);
int sl3_band_cholesky_solve_batched(int count, const int* n, const int* kd,
                                    const int64_t* ab_off,
                                    const double* ab_dev, const int* nrhs,
                                    const int64_t* b_off,
                                    const int64_t* row_stride,
                                    const int64_t* col_stride, double* b_dev);
int sl3_band_lu_factor_batched(int count, const int* n, const int* kl,
                               const int* ku, const int64_t* ab_off,
                               double* ab_dev, const int64_t* piv_off,
                               int* piv_dev, int* info_dev);
int sl3_band_lu_solve_batched(int count, const int* n, const int* kl,
                              const int* ku, const int64_t* ab_off,
                              const double* ab_dev, const int64_t* piv_off,
                              const int* piv_dev, const int* nrhs,
                              const int64_t* b_off, const int64_t* row_stride,
                              const int64_t* col_stride, double* b_dev);

#ifdef __cplusplus
}
#endif
#endif
