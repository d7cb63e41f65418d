// The memory response of the degree systems: per degree task, the dense
// matrix G = D A_s^{-1} D restricted to the dofs of the memory elements
// (A_s the diagonally scaled band matrix, D its scaling). The fixed point's
// force lives only on those dofs and strain recovery reads only those, so
// one pass's band solves (DegreeOperator::Solve with BeginSolve/FinishSolve,
// src/DegreeOperator.cpp:861-901) become x = G f on the memory dofs: a
// grouped FP64 GEMM on the matrix cores instead of a latency-bound
// substitution per right-hand side. G is built once per (scheme, dt) by the
// band solves themselves, one unit column per memory dof.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "coupling.hpp"

namespace sl3amd {

// One workgroup's share of the response GEMM: rows 16 rb .. 16 rb + 63 and
// columns 16 cb .. 16 cb + 127 of one task's output.
struct ResponseWork {
  int task = 0;
  int rb = 0;
  int cb = 0;
  int pad = 0;
};

// Every workgroup's work over the tasks, the largest contraction first:
// the narrow tasks' (at most kNarrowColTiles column tiles; pad = 1, 128 x 64
// blocks) first, then the wide tasks' (64 x 128 blocks).
constexpr int kNarrowColTiles = 4;
std::vector<ResponseWork> ResponseWorkList(const std::vector<DegreeTask>& tasks);
// The narrow items at the list's front.
int ResponseNarrowCount(const std::vector<ResponseWork>& work);

// z (per task at zOff[task], ndof x nm row-major, zero on entry): column j
// gets scale[dm[j]] in row dm[j] (BeginSolve of the unit vector).
void LaunchUnitColumns(const DegreeTask* tasks_dev, int taskCount, int nmMax,
                       const int* dm, const double* scale, const long long* zOff,
                       double* z, hipStream_t stream);

// G[i][j] = scale[dm[i]] z[dm[i]][j] (FinishSolve), zero-padded to nmPad and
// stored in the matrix-core operand order the response GEMM streams.
void LaunchPackResponse(const DegreeTask* tasks_dev, int taskCount, int nmPadMax,
                        const int* dm, const double* scale, const long long* zOff,
                        const double* z, double* g, hipStream_t stream);

// x = G f for every task (f: nmPad x colsPad row-major at fOff; x the
// same size at fOff, column-major: x[col][row]).
// live (per realisation, or null): the column tiles of realisations marked
// 0 (converged; realisation r's columns at r cols) are skipped.
void LaunchResponse(const DegreeTask* tasks_dev, const ResponseWork* work_dev,
                    int workCount, int narrowCount, const double* g, const double* f, double* x,
                    hipStream_t stream, const int* live = nullptr, int realisations = 1);

// The full scaled right-hand sides from the memory-dof force (for a band
// solve of the whole system): rhs[dm[i]][c] = scale[dm[i]] f[i][c]; rows
// outside the memory dofs are left as they are.
void LaunchExpandForce(const DegreeTask* tasks_dev, int taskCount, int nmMax,
                       int colsMax, const int* dm, const double* scale,
                       const double* f, double* rhs, hipStream_t stream);

// The force rows (row-major) -> and the response layout (column-major) <-
// the reference's solution layout (every task's ndof x (l + 1) complex
// columns, column-major, task t at solOff[t]): the boundary's force and
// strain maps. Dofs outside the memory dofs (and the
// rows' padding) are left as they are.
void LaunchRowsToSolutions(const DegreeTask* tasks_dev, int taskCount, int nmMax, int lmax,
                           const int* dm, const long long* solOff, const double* rows,
                           double* solutions, hipStream_t stream);
void LaunchSolutionsToRows(const DegreeTask* tasks_dev, int taskCount, int nmMax, int lmax,
                           const int* dm, const long long* solOff, const double* solutions,
                           double* rows, hipStream_t stream);

}  // namespace sl3amd
