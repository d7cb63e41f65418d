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
// panelRows: the rows of a staged band panel in the fixed-width sweeps
// (4, 8 or 16; 0: the default 8); every panel gives the same bits.
void LaunchCholeskySolve(const BandSystem* systems_dev, const int2* tasks_dev,
                         int tasks, int kdMax, const double* ab_dev,
                         double* b_dev, hipStream_t stream, int panelRows = 0);
constexpr int kPanelChoices = 3;
constexpr int kPanelRowsChoices[kPanelChoices] = {4, 8, 16};
// The tuner of a batch's band panel: times each panel's solve in place on
// b_dev (scratch of the batch's layout) and returns the fastest; ms_out
// (kPanelChoices, may be null) takes the ms per solve.
int TuneCholeskySolvePanel(const BandSystem* systems_dev, const int2* tasks_dev, int tasks,
                           int kdMax, const double* ab_dev, double* b_dev, hipStream_t stream,
                           double* ms_out);
void LaunchLuSolve(const BandSystem* systems_dev, const int2* tasks_dev,
                   int tasks, int kvMax, const double* ab_dev,
                   const int* piv_dev, double* b_dev, hipStream_t stream);
// Tile shapes fixed for every coupling and band batch tuned from now on
// (0: tuned); for tests and timing studies (defined with the Legendre
// tuner, step.hip).
void SetTuningOverride(int analysisRows, int synthesisTheta, int bandPanelRows);
int BandPanelOverride();
// Blocks of 64 right-hand sides: (system, first column) per workgroup.
std::vector<int2> SolveTasks(const std::vector<BandSystem>& systems);

}  // namespace sl3amd
