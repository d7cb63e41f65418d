// Launch wrappers of the HIP kernels (the kernels themselves are file-local
// in the .hip sources).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace sl3amd {

// One band system of a batch. Cholesky uses kl as kd.
struct BandSystem {
  int n = 0;
  int kl = 0;
  int ku = 0;
  int nrhs = 0;
  long long ab_off = 0;
  long long piv_off = 0;
  long long b_off = 0;
  long long row_stride = 1;
  long long col_stride = 0;
};

hipStream_t Stream();

size_t CholeskyFactorLds(int kdMax);
size_t LuFactorLds(int klMax, int kuMax);
void LaunchCholeskyFactor(const BandSystem* systems_dev, int count, int kdMax,
                          double* ab_dev, int* info_dev, hipStream_t stream);
void LaunchLuFactor(const BandSystem* systems_dev, int count, int klMax,
                    int kuMax, double* ab_dev, int* piv_dev, int* info_dev,
                    hipStream_t stream);
void LaunchCholeskySolve(const BandSystem* systems_dev, const int2* tasks_dev,
                         int tasks, int kdMax, const double* ab_dev,
                         double* b_dev, hipStream_t stream);
void LaunchLuSolve(const BandSystem* systems_dev, const int2* tasks_dev,
                   int tasks, int kvMax, const double* ab_dev,
                   const int* piv_dev, double* b_dev, hipStream_t stream);
// Blocks of 64 right-hand sides: (system, first column) per workgroup.
std::vector<int2> SolveTasks(const std::vector<BandSystem>& systems);

}  // namespace sl3amd
