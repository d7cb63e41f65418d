// Launch interface of the transform kernels (transform.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace sl3amd {

// The Legendre stage's tables for the three stored blocks (spin -2, -1
// complex over all orders; spin 0 real over m >= 0), device pointers.
struct LegendrePlan {
  const double* table[3] = {};
  const int* rowOffset[3] = {};
  int mlo[3] = {};
  int spin[3] = {};
  const int* orders = nullptr;  // (block, m) per launch z, device
  int orderCount = 0;
  // a node block (its nodes nodeOffset .. nodeOffset + nodes - 1 of the
  // spectral tensor's nodesTotal, transformed to degree lmax on its own
  // ntheta colatitudes): the tables and the spectral rows keep the full
  // degree's row layout (rowOffset, mlo), the block's Fourier coefficients
  // their own orders (from fourierMlo)
  int lmax = 0, nodes = 0, ntheta = 0;
  int nodeOffset = 0, nodesTotal = 0;
  int fourierMlo[3] = {};
  // the Fourier layout's colatitude stride, ntheta rounded up to 8 (whole
  // 128-byte lines per (order, node) row of complex values)
  int fstride = 0;
  // the matrix-core kernels' tile shapes (0: the defaults below): degrees
  // per analysis workgroup (16, 32, 64, 128), colatitudes per synthesis
  // workgroup (32, 64, 112, 208); picked per shape by the tuner
  int analysisRows = 0;
  int synthesisTheta = 0;
  // realisations side by side: a column block whose nodes (nodeOffset +
  // node) all have live[node / liveNodes] == 0 is skipped (null: none)
  const int* live = nullptr;
  int liveNodes = 1;
  // a truncated analysis whose output is zero below each degree's
  // truncation already (and read only above it) leaves those entries as
  // they are: neither the zeros written nor the addend read there
  bool keepInactive = false;
};

constexpr int kDefaultAnalysisRows = 64;
constexpr int kDefaultSynthesisTheta = 208;
constexpr int kTileChoices = 4;
constexpr int kAnalysisRowsChoices[kTileChoices] = {16, 32, 64, 128};
constexpr int kSynthesisThetaChoices[kTileChoices] = {32, 64, 112, 208};

// ntheta rounded up to 8 complex values (128 bytes).
inline int FourierStride(int ntheta) { return (ntheta + 7) / 8 * 8; }

// The longitude DFT's tables: nphi = R P; cos/sin(2 pi j / P) and
// cos/sin(2 pi j / nphi), device pointers.
struct DftTables {
  int R = 1, P = 1;
  const double* cosP = nullptr;
  const double* sinP = nullptr;
  const double* cosN = nullptr;
  const double* sinN = nullptr;
  // The matrix-core form (odd P): jt tiles of 16 frequencies; tabC / tabS
  // hold cos / sin(2 pi jk / P) for j, k = 1..(P-1)/2 (zero beyond) in the
  // A-operand order of v_mfma_f64_16x16x4f64: tile (t, kb), lane i ->
  // j = 1 + 16 t + i % 16, k = 1 + 4 kb + i / 16. jt = 0: no such form.
  int jt = 0;
  const double* tabC = nullptr;
  const double* tabS = nullptr;
};

// The largest jt the matrix-core ring kernel is instantiated for.
constexpr int kRingMfmaMaxTiles = 5;

enum RingOp {
  kRingSynthesise = 0,        // F -> grid
  kRingAnalyse = 1,           // grid -> F
  kRingWeighted = 2,          // grid x weight -> F
  kRingPass = 3,              // F -> ring x weight -> F (no grid round trip)
  kRingStart = 4,             // step start from a stored strain grid
  kRingStartFromSpectral = 5, // step start, d0 synthesised from F
  kRingFinish = 6,            // F -> d grid, m += beta d
  kRingNodeScaled = 7,        // grid x mu(node) -> F
};

struct RingLaunch {
  DftTables tables;
  int lmax = 0, nodes = 0, ntheta = 0, nphi = 0;
  int fstride = 0;  // the Fourier layout's colatitude stride (FourierStride)
  int block = 0;
  int op = kRingAnalyse;
  const double* wtheta = nullptr;
  const double* fourierIn = nullptr;
  double* fourierOut = nullptr;
  double* fourierOut2 = nullptr;
  const double* gridIn = nullptr;
  double* gridOut = nullptr;
  const double* weight = nullptr;
  const double* memIn = nullptr;
  double* memOut = nullptr;
  const double* itau = nullptr;
  double* betaOut = nullptr;
  double* weightOut = nullptr;
  const double* beta = nullptr;
  const double* muNode = nullptr;
  const double* itauRefNode = nullptr;
  double dt = 0.0;
  int scheme = 0;
  // realisations side by side: live[node / liveNodes] == 0 marks a
  // converged realisation, whose nodes the pass skips (null: none)
  const int* live = nullptr;
  int liveNodes = 1;
};

// The three blocks' Fourier buffers of one fused pass launch.
struct RingBlocks {
  const double* in[3] = {};
  double* out[3] = {};
};

size_t RingLds(int nphi, int P);
// The matrix-core ring kernel (ring_mfma.hip): its tile count for P (0 if
// it does not apply), its LDS bytes, its launch.
int RingMfmaTiles(int P);
size_t RingMfmaLds(const RingLaunch& r);
void LaunchRingMfma(const RingLaunch& r, hipStream_t stream);
// The pass (op kRingPass) over all three blocks in one launch, for R = 4
// grids the matrix-core kernel takes; r carries the common fields.
bool RingPassFusable(const RingLaunch& r);
void LaunchRingPassMfma(const RingLaunch& r, const RingBlocks& blocks, hipStream_t stream);
void LaunchLegendreSynthesis(const LegendrePlan& p, const double* const spectral[3],
                             double* const fourier[3], const int* reach,
                             hipStream_t stream);
// reach (per node, the largest degree it carries) and firstNode (per
// degree, its first node) give the degree truncation, or null for none.
void LaunchLegendreAnalysis(const LegendrePlan& p, const double* const fourier[3],
                            double* const spectral[3], const double* const addend[3],
                            const int* reach, const int* firstNode, hipStream_t stream);
// The matrix-core Legendre stage (legendre_mfma.hip).
void LaunchLegendreSynthesisMfma(const LegendrePlan& p, const double* const spectral[3],
                                 double* const fourier[3], const int* reach,
                                 hipStream_t stream);
void LaunchLegendreAnalysisMfma(const LegendrePlan& p, const double* const fourier[3],
                                double* const spectral[3], const double* const addend[3],
                                const int* reach, const int* firstNode, hipStream_t stream);
void LaunchRing(const RingLaunch& r, hipStream_t stream);

}  // namespace sl3amd
