// Launch interface of the coupling kernels (coupling.hip): the memory
// force on the degree systems, strain recovery from their solutions, the
// residual and the Anderson sweeps, and layout permutations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace sl3amd {

// One degree system of the step's batch (spheroidal or toroidal).
struct DegreeTask {
  int l = 0;
  int kind = 0;  // 0 spheroidal, 1 toroidal
  int ndof = 0;
  int cols = 0;         // 2 (l + 1) real columns: Re of m = 0..l, then Im
  int firstNode = 0;    // the degree's first memory node
  int firstElement = 0;
  int pinned = -1;      // a pinned dof (toroidal l = 1), or -1
  int pad = 0;
  long long rhsOff = 0;    // [dof][cols] row-major
  long long scaleOff = 0;  // ndof doubles
  long long dofOff = 0;    // memory elements x ngll x (U, V) or (W)
  long long unitOff = -1;  // the unit load, then tide, responses (2 ndof), spheroidal
  // the memory dofs: the force's rows and the strain's reads
  int nm = 0;              // count (the pinned dof excluded)
  int nmPad = 0;           // nm rounded up to 16
  int colsPad = 0;         // cols rounded up to 16
  int pad2 = 0;
  long long dmOff = 0;     // their dof numbers (nm ints), compact unit responses
  long long gOff = 0;      // the memory response, nmPad^2 doubles
  long long fOff = 0;      // force and response rows: nmPad x colsPad doubles
};

// The radial geometry the force and strain maps read (device pointers).
struct RadialGeometry {
  int nodes = 0;       // memory nodes = memory elements x np
  int np = 0;          // points per element
  int ngll = 0;
  int elem0 = 0;       // first memory element
  int elements = 0;    // memory elements (contiguous)
  // independent realisations side by side: realisation r's node n is the
  // spectral column r nodes + n, its degree-task columns r cols + c
  int realisations = 1;
  const double* radius = nullptr;  // per node
  const double* jac = nullptr;     // per memory element
  const double* weight = nullptr;  // per point (reference weights)
  const double* value = nullptr;   // (np, ngll)
  const double* deriv = nullptr;   // (np, ngll)
};

struct SpectralView {
  double* block[3] = {};           // device layout blocks (see grid.hpp)
  const int* rowOffset[3] = {};    // per m - mlo
  int lmax = 0;
};

void LaunchFill(double* dst, size_t count, double value, hipStream_t stream);
void LaunchCopy(double* dst, const double* src, size_t count, hipStream_t stream);
// dst = a dst + b src (a stage's combined state, StageState).
void LaunchCombine(double* dst, double a, const double* src, double b, size_t count,
                   hipStream_t stream);

// Forces of the analysed tensor on every task's memory dofs (rows of
// nmPad x colsPad at fOff, unscaled: the memory response carries the
// scaling; realisation r's columns at r cols). tasks: spheroidal l = 0..L,
// then toroidal l = 1..L. dofs: per task, memory element x GLL node x
// (U, V) or (W, -), the memory-dof row or -1 (none, or the pinned dof).
// work: the (degree, first order) chunks of ForceWork(lmax).
std::vector<int2> ForceWork(int lmax);
// live (per realisation, or null): a realisation marked 0 has converged
// and the maps below skip it.
void LaunchForce(const DegreeTask* tasks_dev, const RadialGeometry& g,
                 const SpectralView& weighted, const int* dofs, const int2* work_dev,
                 int workCount, double* force, hipStream_t stream,
                 const int* live = nullptr);

// Strain of the responses on the memory dofs into the spectral iterate
// (every value written; realisation r from its response columns r cols,
// the channels shared), the channels' unit responses (per memory dof at
// dmOff) added with their amplitudes (complex per (l, m), offset
// l(l+1)/2 + m): the load's from units and amp, the tide's from tideUnits
// and tideAmp (tideAmp null: no tide).
void LaunchStrain(const DegreeTask* sphTasks_dev, const DegreeTask* torTasks_dev,
                  int lmax, const RadialGeometry& g, const int* dofs,
                  const double* response, const double* units, const double* amp,
                  const double* tideUnits, const double* tideAmp, SpectralView out,
                  hipStream_t stream, const int* live = nullptr);

// Where a reduction's results also go for the host: a one-workgroup kernel
// after the reduction writes them to `values` (host-mapped memory), then
// `ticket` to `seq` (host-mapped) after a system-scope fence; the host polls
// `seq` (no copy, no event).
struct HostPublish {
  double* values = nullptr;
  unsigned long long* seq = nullptr;
  unsigned long long ticket = 0;
};
void LaunchPublish(const double* src, int n, const HostPublish& pub, hipStream_t stream);

// max |n|^2, max |n - o|^2 and a non-finite flag over n values, into
// result[0..2] (device; with pub, also published to the host).
void LaunchRelativeChange(const double* next, const double* now, size_t n,
                          double* partials, int partialBlocks, double* result,
                          hipStream_t stream, const HostPublish* pub = nullptr);

// The active entries of a spectral tensor: per row (all blocks, in memory
// order) the run of nodes at or above the row degree's first memory node;
// the others are exact zeros of every iterate. full[r]: the run's first
// double in the device layout; compact[r]: its offset in the compact
// vector (runs + 1 entries, the last the total).
struct ActiveRuns {
  const long long* full = nullptr;
  const long long* compact = nullptr;
  int runs = 0;
  long long total = 0;
};

// The deepest Anderson history the kernels hold.
constexpr int kAndersonMax = 16;
// The sweep's sums for a history of m columns: the Gram column at [0, w),
// the right-hand side at [w, 2 w), w = 8 up to m = 8, else 16.
inline int AndersonWidth(int m) { return m <= 8 ? 8 : kAndersonMax; }

// Anderson (AndersonMixer.hpp:123-218) over the active entries: x and gx
// in the device layout, the history (xPrev, fPrev, dF, dG) compact.
struct AndersonSweep {
  double* x;            // in: iterate; (second sweep) out: mixed
  const double* gx;     // the pass's image
  double* xPrev;
  double* fPrev;
  double* dF[kAndersonMax];
  double* dG[kAndersonMax];
  int m;                // active history columns
  int slot;             // the slot being written
};
// RelativeChange over the active entries only (exact when the inactive
// entries of both iterates are zero).
void LaunchRelativeChangeActive(const double* next, const double* now, const ActiveRuns& runs,
                                double* partials, int partialBlocks, double* result,
                                hipStream_t stream, const HostPublish* pub = nullptr);

// The first mixed pass: xPrev = x, fPrev = gx - x (compact), then x = gx
// (every entry, n doubles).
void LaunchAndersonFirst(double* x, const double* gx, double* xPrev,
                         double* fPrev, const ActiveRuns& runs, size_t n,
                         hipStream_t stream);
// The same with x = gx over the entries of `every` (one realisation's
// columns of every row) in place of the whole tensor.
void LaunchAndersonFirst(double* x, const double* gx, double* xPrev, double* fPrev,
                         const ActiveRuns& runs, const ActiveRuns& every, hipStream_t stream);
// dst = src on the entries of the runs (device layout).
void LaunchCopyRuns(double* dst, const double* src, const ActiveRuns& runs, hipStream_t stream);
// The Anderson sweep (differences, the Gram column and right-hand side into
// sums[0, 2 w), w = AndersonWidth(m)); with residual non-null also the
// pass's RelativeChange (max |gx|^2, max |gx - x|^2, a non-finite flag)
// into residual[0..2]. partials holds partialBlocks x (2 kAndersonMax + 3)
// doubles. With pub, the sums and then the residual's three values (if
// any; residual must then be sums + 2 w) are also published to the host.
void LaunchAndersonSweep(const AndersonSweep& s, const ActiveRuns& runs, double* partials,
                         int partialBlocks, double* sums, double* residual,
                         hipStream_t stream, const HostPublish* pub = nullptr);
void LaunchAndersonMix(const AndersonSweep& s, const double* gamma, const ActiveRuns& runs,
                       hipStream_t stream);

// Several realisations' vector sweeps in one launch (realisations side by
// side, every realisation's compact vector laid out alike): entry e works
// on realisation r[e], whose runs are the given runs' full offsets moved by
// r[e] runs, whose Anderson history is the base sweep's moved by r[e]
// strides, with its own m and slot (sweeps and mixes) and, for its
// RelativeChange, every[e]: over every entry of its columns. Each entry's
// arithmetic is its single-realisation launch's, bit for bit.
constexpr int kBatchMax = 64;
struct RealisationBatch {
  int count = 0;
  long long prevStride = 0;     // doubles between realisations' xPrev, and fPrev
  long long historyStride = 0;  // doubles between realisations' dF[i], and dG[i]
  int r[kBatchMax] = {};
  int m[kBatchMax] = {};
  int slot[kBatchMax] = {};
  int every[kBatchMax] = {};
};
// partials: per entry partialBlocks x (2 kAndersonMax + 3) doubles. Results
// go to result + r stride: RelativeChange's three maxima at 0; a sweep's
// sums at [0, 2 W) and, with residual, its maxima after them; W =
// AndersonWidth of the entries' largest m, returned.
void LaunchRelativeChangeBatch(const double* next, const double* now, const ActiveRuns& runs,
                               const ActiveRuns& every, const RealisationBatch& b,
                               double* partials, int partialBlocks, double* result, int stride,
                               hipStream_t stream);
int LaunchAndersonSweepBatch(const AndersonSweep& base, const ActiveRuns& runs,
                             const RealisationBatch& b, bool residual, double* partials,
                             int partialBlocks, double* result, int stride, hipStream_t stream);
// gammas: per entry kAndersonMax doubles (device).
void LaunchAndersonMixBatch(const AndersonSweep& base, const ActiveRuns& runs,
                            const RealisationBatch& b, const double* gammas,
                            hipStream_t stream);
// The first mixed pass of every entry, x = gx over every entry of its columns.
void LaunchAndersonFirstBatch(const AndersonSweep& base, const ActiveRuns& runs,
                              const ActiveRuns& every, const RealisationBatch& b,
                              hipStream_t stream);
// dst = src over every entry of each entry's columns.
void LaunchCopyRunsBatch(double* dst, const double* src, const ActiveRuns& every,
                         const RealisationBatch& b, hipStream_t stream);
// n values of every entry's result (result + r stride) to pub.values +
// r valuesStride (host-mapped), then pub.ticket to pub.seq[r].
void LaunchPublishBatch(const double* result, int stride, int n, const RealisationBatch& b,
                        const HostPublish& pub, int valuesStride, hipStream_t stream);

// GridCoupling::ScaleByNode (src/GridCoupling.cpp:524-540) on a device-
// layout spectral tensor ([row][node] complex, count complex values):
// every coefficient of a node times mu[node].
void LaunchScaleByNode(const double* mu, double* spectral, size_t count, int nodes,
                       hipStream_t stream);
// GridCoupling::Multiply (src/GridCoupling.cpp:497-509) on a grid tensor:
// c0, c1 (complex) and r times the weight, point by point (points values
// per block, the weight per point).
void LaunchMultiply(double* c0, double* c1, double* r, const double* weight, size_t points,
                    hipStream_t stream);

// Reference layout (per node, l-major, SurfaceGrid::SpinIndex / Index)
// <-> device layout, for one block over `nodes` nodes.
void LaunchToDeviceLayout(const double* ref, double* dev, const int* rowOffset,
                          int spin, int lmax, int nodes, hipStream_t stream);
void LaunchToReferenceLayout(const double* dev, double* ref, const int* rowOffset,
                             int spin, int lmax, int nodes, hipStream_t stream);

}  // namespace sl3amd
