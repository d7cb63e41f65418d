// C ABI of the banded solvers (include/sl3dpp_amd.h), the drop-in for
// sl3dpp::detail::BandCholeskyFactor / BandCholeskySolve / BandLuFactor /
// BandLuSolve (include/sl3dpp/detail/BandSolvers.hpp:24-49).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace sl3amd {

namespace {

constexpr int kMaxBand = 95;  // kl + ku (or kd) per system: LDS-sized

void Check(bool ok, const std::string& what) {
  if (!ok) throw ArgumentError(what);
}

struct Batch {
  std::vector<BandSystem> systems;
  int klMax = 0, kuMax = 0;
};

Batch Describe(int count, const int* n, const int* kl, const int* ku,
               const int64_t* ab_off, const int64_t* piv_off, const int* nrhs,
               const int64_t* b_off, const int64_t* row_stride,
               const int64_t* col_stride) {
  Check(count >= 0, "the system count must be non-negative");
  Batch batch;
  batch.systems.resize(count);
  for (int s = 0; s < count; ++s) {
    auto& sys = batch.systems[s];
    sys.n = n[s];
    sys.kl = kl[s];
    sys.ku = ku ? ku[s] : 0;
    Check(sys.n >= 1, "a band system has at least one row");
    Check(sys.kl >= 0 && sys.ku >= 0, "band widths must be non-negative");
    Check(sys.kl + sys.ku <= kMaxBand,
          "band wider than the kernels' LDS staging (kl + ku <= 95)");
    sys.ab_off = ab_off[s];
    sys.piv_off = piv_off ? piv_off[s] : 0;
    if (nrhs) {
      sys.nrhs = nrhs[s];
      Check(sys.nrhs >= 0, "the right-hand side count must be non-negative");
      sys.b_off = b_off[s];
      sys.row_stride = row_stride ? row_stride[s] : 1;
      sys.col_stride = col_stride ? col_stride[s] : sys.n;
    }
    batch.klMax = std::max(batch.klMax, sys.kl);
    batch.kuMax = std::max(batch.kuMax, sys.ku);
  }
  return batch;
}

void FactorCholesky(const Batch& batch, double* ab, int* info) {
  hipStream_t stream = Stream();
  DeviceArray<BandSystem> systems;
  systems.Upload(batch.systems.data(), batch.systems.size(), stream);
  LaunchCholeskyFactor(systems.data(), static_cast<int>(batch.systems.size()),
                       batch.klMax, ab, info, stream);
  SL3_HIP(hipStreamSynchronize(stream));
}

void FactorLu(const Batch& batch, double* ab, int* piv, int* info) {
  hipStream_t stream = Stream();
  DeviceArray<BandSystem> systems;
  systems.Upload(batch.systems.data(), batch.systems.size(), stream);
  LaunchLuFactor(systems.data(), static_cast<int>(batch.systems.size()),
                 batch.klMax, batch.kuMax, ab, piv, info, stream);
  SL3_HIP(hipStreamSynchronize(stream));
}

void Solve(const Batch& batch, const double* ab, const int* piv, double* b,
           bool lu) {
  hipStream_t stream = Stream();
  DeviceArray<BandSystem> systems;
  systems.Upload(batch.systems.data(), batch.systems.size(), stream);
  const auto tasks = SolveTasks(batch.systems);
  DeviceArray<int2> taskArray;
  taskArray.Upload(tasks.data(), tasks.size(), stream);
  if (lu) {
    LaunchLuSolve(systems.data(), taskArray.data(),
                  static_cast<int>(tasks.size()), batch.klMax + batch.kuMax, ab,
                  piv, b, stream);
  } else {
    LaunchCholeskySolve(systems.data(), taskArray.data(),
                        static_cast<int>(tasks.size()), batch.klMax, ab, b,
                        stream);
  }
  SL3_HIP(hipStreamSynchronize(stream));
}

}  // namespace
}  // namespace sl3amd

using namespace sl3amd;

extern "C" {

int sl3_band_cholesky_factor(double* ab_dev, int n, int kd, int* info) {
  return Guard([&] {
    Check(ab_dev != nullptr && info != nullptr, "null argument");
    const int64_t off = 0;
    const auto batch = Describe(1, &n, &kd, nullptr, &off, nullptr, nullptr,
                                nullptr, nullptr, nullptr);
    DeviceArray<int> infoDev(1);
    FactorCholesky(batch, ab_dev, infoDev.data());
    SL3_HIP(hipMemcpy(info, infoDev.data(), sizeof(int), hipMemcpyDeviceToHost));
  });
}

int sl3_band_cholesky_solve(const double* ab_dev, int n, int kd, double* b_dev,
                            int nrhs) {
  return Guard([&] {
    Check(ab_dev != nullptr && b_dev != nullptr, "null argument");
    const int64_t off = 0, rs = 1, cs = n;
    const auto batch =
        Describe(1, &n, &kd, nullptr, &off, nullptr, &nrhs, &off, &rs, &cs);
    Solve(batch, ab_dev, nullptr, b_dev, false);
  });
}

int sl3_band_lu_factor(double* ab_dev, int n, int kl, int ku, int* piv_dev,
                       int* info) {
  return Guard([&] {
    Check(ab_dev != nullptr && piv_dev != nullptr && info != nullptr,
          "null argument");
    const int64_t off = 0;
    const auto batch = Describe(1, &n, &kl, &ku, &off, &off, nullptr, nullptr,
                                nullptr, nullptr);
    DeviceArray<int> infoDev(1);
    FactorLu(batch, ab_dev, piv_dev, infoDev.data());
    SL3_HIP(hipMemcpy(info, infoDev.data(), sizeof(int), hipMemcpyDeviceToHost));
  });
}

int sl3_band_lu_solve(const double* ab_dev, int n, int kl, int ku,
                      const int* piv_dev, double* b_dev, int nrhs) {
  return Guard([&] {
    Check(ab_dev != nullptr && piv_dev != nullptr && b_dev != nullptr,
          "null argument");
    const int64_t off = 0, rs = 1, cs = n;
    const auto batch =
        Describe(1, &n, &kl, &ku, &off, &off, &nrhs, &off, &rs, &cs);
    Solve(batch, ab_dev, piv_dev, b_dev, true);
  });
}

int sl3_band_cholesky_factor_batched(int count, const int* n, const int* kd,
                                     const int64_t* ab_off, double* ab_dev,
                                     int* info_dev) {
  return Guard([&] {
    Check(count == 0 || (n && kd && ab_off && ab_dev && info_dev),
          "null argument");
    const auto batch = Describe(count, n, kd, nullptr, ab_off, nullptr,
                                nullptr, nullptr, nullptr, nullptr);
    FactorCholesky(batch, ab_dev, info_dev);
  });
}

int sl3_band_cholesky_solve_batched(int count, const int* n, const int* kd,
                                    const int64_t* ab_off,
                                    const double* ab_dev, const int* nrhs,
                                    const int64_t* b_off,
                                    const int64_t* row_stride,
                                    const int64_t* col_stride, double* b_dev) {
  return Guard([&] {
    Check(count == 0 || (n && kd && ab_off && ab_dev && nrhs && b_off && b_dev),
          "null argument");
    const auto batch = Describe(count, n, kd, nullptr, ab_off, nullptr, nrhs,
                                b_off, row_stride, col_stride);
    Solve(batch, ab_dev, nullptr, b_dev, false);
  });
}

int sl3_band_lu_factor_batched(int count, const int* n, const int* kl,
                               const int* ku, const int64_t* ab_off,
                               double* ab_dev, const int64_t* piv_off,
                               int* piv_dev, int* info_dev) {
  return Guard([&] {
    Check(count == 0 || (n && kl && ku && ab_off && ab_dev && piv_off &&
                         piv_dev && info_dev),
          "null argument");
    const auto batch = Describe(count, n, kl, ku, ab_off, piv_off, nullptr,
                                nullptr, nullptr, nullptr);
    FactorLu(batch, ab_dev, piv_dev, info_dev);
  });
}

int sl3_band_lu_solve_batched(int count, const int* n, const int* kl,
                              const int* ku, const int64_t* ab_off,
                              const double* ab_dev, const int64_t* piv_off,
                              const int* piv_dev, const int* nrhs,
                              const int64_t* b_off, const int64_t* row_stride,
                              const int64_t* col_stride, double* b_dev) {
  return Guard([&] {
    Check(count == 0 || (n && kl && ku && ab_off && ab_dev && piv_off &&
                         piv_dev && nrhs && b_off && b_dev),
          "null argument");
    const auto batch = Describe(count, n, kl, ku, ab_off, piv_off, nrhs, b_off,
                                row_stride, col_stride);
    Solve(batch, ab_dev, piv_dev, b_dev, true);
  });
}

}  // extern "C"
