// The 3D lateral-viscosity step on the GPU: the device-resident grid
// coupling, the per-(scheme, dt) flavours of factorised degree systems,
// and Viscoelastic3D's Step (StartStep, the Converge fixed point with
// adaptive Anderson mixing, Finish), mirroring the reference's
// include/sl3dpp/Viscoelastic3D.hpp:89-675 for one Maxwell branch, Gauss
// memory points, one node block and one process.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <list>
#include <memory>
#include <vector>

#include "common.hpp"
#include "coupling.hpp"
#include "grid.hpp"
#include "kernels.hpp"
#include "lateral_cg.hpp"
#include "model.hpp"
#include "degree_systems.hpp"
#include "response.hpp"
#include "transform.hpp"

namespace sl3amd {

// How Step closes a step's linear system (Options::StepSolver,
// include/sl3dpp/Viscoelastic3D.hpp:122-125).
enum { kFixedPoint = 0, kConjugateGradients = 1 };

// Viscoelastic3D::Options (include/sl3dpp/Viscoelastic3D.hpp:89-187): the
// fixed point's controls and the step solver's.
struct StepOptions {
  double eps = 1e-8;
  double oversampling = 1.5;
  int maxIterations = 40;
  double rtol = 1e-11;
  double damping = 1.0;
  int andersonMemory = 5;
  double andersonRate = 0.5;
  int solver = kFixedPoint;
  double solveRtol = 1e-11;
  int solveMaxIterations = 200;
  // each pass (synthesis, ring pass, analysis, force, response, strain)
  // replays a hipGraph captured on its first use
  bool graphPasses = false;
  // Options::transform_degrees (include/sl3dpp/Viscoelastic3D.hpp:
  // 172-175): the degrees the transforms may be cut to with depth
  std::vector<int> transformDegrees;
  // independent realisations (each its own lateral field and state, the
  // Earth, degree, load and tide shared) stepped together: every kernel
  // takes their memory nodes as extra columns, each converges on its own
  int realisations = 1;
};

// The most realisations one operator steps together.
constexpr int kRealisationsMax = 64;

// Viscoelastic3D::StepInfo (include/sl3dpp/Viscoelastic3D.hpp:357-375).
struct StepInfo {
  int iterations = 0;
  double residual = 0.0;
  bool converged = false;
  int mixedPasses = 0;
  int solveIterations = 0;  // conjugate-gradient iterations
  bool warmFallback = false;  // a guarded warm start was dropped for the step's own
};

// LoadCallable3D: the m >= 0 columns of degree l at time t, (re, im) pairs,
// or false for none.
using LoadFn = std::function<bool(int l, double t, double* columns)>;

// The step's channels (Viscoelastic3D::Step's load and tide, either may be
// empty). The tide's degree-1 columns must vanish (src/Viscoelastic3D.cpp:
// 138-154).
struct Channels {
  LoadFn load;
  LoadFn tide;
};

// A node block (GridCoupling's Options::transform_degrees,
// include/sl3dpp/GridCoupling.hpp:23-40, src/GridCoupling.cpp:163-190):
// contiguous memory nodes transformed at the block's degree on a grid of
// that degree, with its own Wigner and DFT tables; every grid field lays the
// blocks end to end, every spin block's Fourier buffer too.
struct NodeBlock {
  NodeBlock(int lmax, double oversampling) : lmax(lmax), grid(lmax, oversampling) {}
  int firstNode = 0, nodes = 0, lmax = 0;
  SurfaceGrid grid;
  size_t fieldOffset = 0;        // grid values before the block
  size_t fourierOffset[3] = {};  // Fourier doubles before the block, per spin block
  DeviceArray<double> table[3];  // rows in the full degree's layout
  // the rows of degree above the block's (zero at the block's nodes after
  // an analysis), per spin block
  DeviceArray<int> cutRows[3];
  int cutRowCount[3] = {};
  DeviceArray<int> orders;
  DeviceArray<double> wtheta, cosP, sinP, cosN, sinN, tabC, tabS;
  LegendrePlan plan;
  DftTables dft;
  std::vector<double> tuning;
};

// The device grid coupling (GridCoupling + SurfaceGrid): memory nodes,
// their node blocks, Wigner tables, DFT tables, layouts.
class Coupling {
 public:
  // transformDegrees: the degrees the transforms may be cut to with depth
  // (Options::transform_degrees; empty: one block at lmax). realisations:
  // independent copies of the memory nodes side by side (realisation r's
  // node n is column r N + n of every spectral row and node r N + n of
  // every grid field, N = RadialNodeCount()); not with transformDegrees.
  Coupling(const Earth& earth, int lmax, double oversampling, double eps,
           const std::vector<int>& transformDegrees = {}, int realisations = 1);
  int LMax() const { return lmax_; }
  // the memory nodes of every realisation (the spectral rows' columns)
  int NodeCount() const { return geometry_.nodes * geometry_.realisations; }
  // the memory nodes of one realisation
  int RadialNodeCount() const { return geometry_.nodes; }
  int Realisations() const { return geometry_.realisations; }
  // the full degree's grid (the layouts, and every node's grid without cuts)
  const SurfaceGrid& Grid() const { return grid_; }
  int BlockCount() const { return static_cast<int>(blocks_.size()); }
  const NodeBlock& Block(int i) const { return *blocks_[i]; }
  // the values of a grid field over every node (the blocks end to end)
  size_t GridValueCount() const { return gridValues_; }
  // doubles of each grid component (complex: 2 per value) and of a
  // spectral tensor (three blocks end to end, device layout)
  size_t GridDoubles(int block) const { return (block == 2 ? 1 : 2) * GridValueCount(); }
  size_t SpectralDoubles() const { return spectralOffset_[3]; }
  size_t SpectralOffset(int block) const { return spectralOffset_[block]; }
  size_t FourierDoubles(int block) const;
  const RadialGeometry& Geometry() const { return geometry_; }
  const std::vector<int>& FirstNodes() const { return firstNodeHost_; }
  // per node of one realisation
  const std::vector<double>& NodeRadii() const { return radiusHost_; }
  const std::vector<int>& NodeElement() const { return nodeElement_; }
  int FirstElementOfDegree(int l) const { return firstElement_[l]; }
  SpectralView View(double* spectral) const;
  const int* RowOffsetDevice(int block) const { return rowOffset_[block].data(); }
  // The active (row, node) runs of a spectral tensor (degree truncation)
  // of realisation r, their compact vector laid out alike for every
  // realisation; with one realisation, of the tensor.
  ActiveRuns Runs(int r = 0) const {
    ActiveRuns a = runs_;
    a.full += static_cast<size_t>(r) * runs_.runs;
    return a;
  }
  // Every entry of realisation r's columns, as runs.
  ActiveRuns EveryRuns(int r) const {
    ActiveRuns a = everyRuns_;
    a.full += static_cast<size_t>(r) * everyRuns_.runs;
    return a;
  }

  // Batched transforms of all three blocks over every node. live (per
  // realisation, or null): realisations marked 0 are skipped (their
  // outputs left as they were).
  void SynthesiseToFourier(const double* spectral, double* const fourier[3],
                           bool truncated, hipStream_t s, const int* live = nullptr) const;
  // keepInactive (truncated only): the output is zero below each degree's
  // truncation already and read only above it; those entries are left.
  void AnalyseFromFourier(const double* const fourier[3], double* spectral,
                          const double* addend, bool truncated, hipStream_t s,
                          const int* live = nullptr, bool keepInactive = false) const;
  void Ring(RingLaunch r, hipStream_t s) const;  // fills tables and sizes
  // The pass's longitude stage (Fourier -> ring x weight -> Fourier) over
  // all three blocks: one launch where the grid allows it.
  void RingPass(double* const in[3], double* const out[3], const double* weight,
                hipStream_t s, const int* live = nullptr) const;
  // The tuner: times every tile shape of the Legendre kernels on this
  // coupling's own shape (the degree truncation included) and keeps the
  // fastest; shapes met before take the choice made then. spectral: a
  // zeroed tensor, fourier and spectralOut: the pass's buffers (written).
  void TuneLegendre(const double* spectral, double* const fourier[3], double* spectralOut,
                    hipStream_t s);
  // Per node block: [analysis rows, synthesis colatitudes] chosen, then the
  // ms per launch of each analysis choice and each synthesis choice.
  std::vector<double> Tuning() const;
  // One block's Legendre stage alone (the C-ABI transforms).
  void SynthesiseBlockToFourier(int block, const double* spectral, double* fourier,
                                hipStream_t s) const;
  void AnalyseBlockFromFourier(int block, const double* fourier, double* spectral,
                               hipStream_t s) const;
  void Synthesise(const double* spectral, double* const grid[3], double* const fourier[3],
                  hipStream_t s) const;
  void Analyse(const double* const grid[3], double* spectral, double* const fourier[3],
               hipStream_t s) const;

 private:
  void BuildBlock(NodeBlock& k, hipStream_t s);
  // After an analysis: the rows above each block's degree at its nodes
  // hold the addend (or zero), as if the block's grid carried them.
  void FillCutRows(double* const spectral[3], const double* const addend[3],
                   hipStream_t s) const;
  // A ring launch's fields for one node block: sizes, tables, and every
  // pointer moved to the block's part.
  RingLaunch BlockLaunch(RingLaunch r, const NodeBlock& k) const;
  int lmax_;
  SurfaceGrid grid_;
  RadialGeometry geometry_;
  std::vector<int> nodeElement_;
  std::vector<double> radiusHost_;
  std::vector<int> firstElement_;  // per degree
  std::vector<int> firstNodeHost_, reachHost_;
  DeviceArray<double> radius_, jac_, weight_, value_, deriv_;
  DeviceArray<int> rowOffset_[3];
  DeviceArray<int> firstNode_, reach_;
  std::vector<std::unique_ptr<NodeBlock>> blocks_;
  size_t gridValues_ = 0;
  size_t fourierDoubles_[3] = {};
  size_t spectralOffset_[4] = {};
  DeviceArray<long long> runFull_, runCompact_, everyFull_, everyCompact_;
  ActiveRuns runs_, everyRuns_;  // realisation 0's; full offsets per realisation follow
};


// The factorised systems of one (scheme, dt): Viscoelastic3D::Flavour
// (src/Viscoelastic3D.cpp:158-167, 535-558).
struct Flavour {
  unsigned long long id = 0;  // unique per operator: keys the captured passes
  int scheme = 0;
  double dt = 0.0;
  std::vector<double> pointShear;
  DeviceArray<double> scale, factors;
  DeviceArray<double> units;  // per spheroidal task at unitOff: load, tide (ndof each)
  DeviceArray<int> pivots;
  std::vector<BandSystem> cholesky, lu;  // the solve batches
  DeviceArray<BandSystem> choleskyDev, luDev;
  DeviceArray<int2> choleskyTasks, luTasks;
  int choleskyTaskCount = 0, luTaskCount = 0;
  int kdMax = 0;
  std::vector<char> usedCholesky;  // per task
  std::vector<int> choleskyTask, luTask;  // the task of each solve system
  // the memory response of every task (G = D A_s^{-1} D on its memory dofs)
  // and the unit load and tide responses on the memory dofs (load for every
  // memory dof, then tide for every memory dof)
  DeviceArray<double> response, unitsMem;
  // per degree, the unit channel forces against their responses over every
  // dof: load.zLoad, load.zTide, tide.zLoad, tide.zTide
  std::vector<double> channelEnergy;
};

// Viscoelastic3D::State on the device: one branch's memory tensor, the
// strain on the grid and its coefficients (device layout).
struct DeviceState {
  double t = 0.0;
  DeviceArray<double> m[3];
  bool hasStrain = false;
  DeviceArray<double> strain[3];
  bool hasCoefficients = false;
  DeviceArray<double> coefficients;
};

class Viscoelastic3D {
 public:
  // delta_dev: delta(r, theta, phi) sampled on every memory node's grid
  // ([node][theta][phi], device), or null for none.
  // The operator shares the Earth's ownership: the caller's handle may go
  // first (the reference copies the rheology, src/Viscoelastic3D.cpp:229).
  Viscoelastic3D(std::shared_ptr<const Earth> earth, int lmax, const double* delta_dev,
                 StepOptions options);
  ~Viscoelastic3D();
  Viscoelastic3D(const Viscoelastic3D&) = delete;
  Viscoelastic3D& operator=(const Viscoelastic3D&) = delete;
  const Coupling& GetCoupling() const { return coupling_; }
  const StepOptions& Options() const { return options_; }
  StepOptions& MutableOptions() { return options_; }
  std::unique_ptr<DeviceState> ZeroState() const;
  // A copy of a state (the reference's Step returns a new state and leaves
  // its argument as it was).
  std::unique_ptr<DeviceState> CloneState(const DeviceState& from) const;
  void SetDelta(const double* delta_dev);
  // Step (src/Viscoelastic3D.cpp:1982-2030) for ExponentialTrapezoid,
  // BackwardEuler or Sdirk2 (two backward Euler stages by the recipe of
  // src/StepRecipe.cpp:25-32), in place on the state. `start` (device
  // spectral, or null) is the warm start of Converge (:1916-1921); a
  // multi-stage step takes none. warmGuard > 0: when the first pass from
  // `start` changes it by a relative residual above warmGuard (or a
  // non-finite one), the step drops it and converges from its own start.
  StepInfo Step(DeviceState& state, double dt, int scheme, const Channels& channels,
                const double* start = nullptr, double warmGuard = 0.0);
  // The strain coefficients a step from `state` starts from (StartStep's
  // s0, src/Viscoelastic3D.cpp:1252-1263): the state's coefficients, the
  // analysis of its strain grid, or the elastic strain of its memory and the
  // channels at its time; into out (device layout). No step may be open.
  void StartStrain(DeviceState& state, const Channels& channels, double* out);
  // The last pass's solutions on every dof, unscaled with channels, per
  // task (band solves of the last pass's force): host export for the
  // surface outputs and tests.
  std::vector<double> SolutionOfTask(int task);
  const std::vector<DegreeTask>& Tasks() const { return tasks_; }
  Flavour& FlavourOf(int scheme, double dt);
  // Timings of the last Step's phases (device events), milliseconds.
  struct Timing {
    float startStep = 0, passes = 0, finish = 0;
  } lastTiming;
  // Per-pass residuals of the last step (conjugate gradients: the relative
  // residual before each iteration and after the last; realisations side
  // by side: the largest of the pass's unconverged realisations).
  std::vector<double> residualHistory;
  // Per realisation, the last step's (or converge's) own StepInfo; the
  // StepInfo returned for several realisations takes the most passes, the
  // largest residual, and converged when all converged.
  const std::vector<StepInfo>& RealisationInfos() const { return realisationInfos_; }
  // The last conjugate-gradient solve's scale <x0, f> (x0 = K^-1 f).
  std::vector<double> solveEnergy;
  // The device buffers, for export.
  const double* Weight() const { return weight_.data(); }
  const double* Beta() const { return beta_.data(); }
  const double* Itau() const { return itau_.data(); }
  const double* FixedForce() const { return fixed_.data(); }
  const double* Analysed() const { return analysed_.data(); }
  const double* Rhs() const { return rhs_.data(); }
  const double* AmplitudesDevice() const { return amp_.data(); }
  const int* Dofs() const { return dofs_.data(); }
  const std::vector<int>& DofsHost() const { return dofsHost_; }
  const std::vector<double>& MuNode() const { return muNodeHost_; }
  const std::vector<double>& ItauRefNode() const { return itauRefHost_; }
  int CholeskyFallbacks(int scheme, double dt);

  // One pass: Respond(s) + AddChannels + StrainOf into sNext; with
  // graphPasses a replay of the pass captured for (s, sNext, f, tide).
  void Pass(const double* s, double* sNext, bool useStartForce, Flavour& f);
  // The captured passes held (tests).
  int PassGraphCount() const { return static_cast<int>(passGraphs_.size()); }

  // OpenStep (include/sl3dpp/Viscoelastic3D.hpp:461-647) on this
  // operator's buffers, one open step at a time; Step is OpenStart +
  // OpenClose + OpenFinish.
  // StartStep (src/Viscoelastic3D.cpp:1223-1378), one stage (exponential
  // trapezoid or backward Euler): the state's memory becomes the decayed
  // memory m~ in place until OpenFinish; the weight, beta, the fixed force,
  // the channels at the end time and the start strain (the iterate) are
  // formed.
  void OpenStart(DeviceState& state, double dt, int scheme, const Channels& channels);
  bool IsOpen() const { return open_.state != nullptr; }
  // One pass of the open step (Respond, AddChannels, StrainOf) of s into
  // next (device spectral tensors).
  void OpenPass(const double* s, double* next);
  using PassFn = std::function<void(const double* s, double* next)>;
  // Converge (:1916-1974) with a caller's pass (empty: OpenPass) from the
  // start strain or `start`; the result is the iterate.
  StepInfo OpenConverge(const PassFn& pass, const double* start, double warmGuard);
  // Close (:2008-2015): the fixed point, or the conjugate gradients.
  StepInfo OpenClose(const double* start, double warmGuard);
  // Finish (:1885-1914) with `strain` (device coefficients; null: the
  // iterate): m+ = m~ + beta d, the strain and coefficients, the end time.
  void OpenFinish(const double* strain);
  // AddMemory (increment per grid block, device) and AddForce's tensor
  // (an analysed tensor, device layout) (:576-598).
  void OpenAddMemory(const double* const increment[3]);
  void OpenAddForce(const double* analysed);
  // The discrete adjoint's transposed pass of the open step
  // (src/AdjointDiscrete3D.cpp:998-1037, without the sea-level and
  // measurement duals): out = Analyse(W Synthesise(StrainOf(SolveForces(
  // Entry(z)))) + src), Entry the forces of z, StrainOf without channels,
  // src a grid tensor (per block, device; null for none). The pass's own
  // pieces in the other order: the adjoint of OpenPass's linear part in
  // the quadrature-weighted pairing of spectral tensors.
  void OpenTransposedPass(const double* z, const double* const src[3], double* out);
  // Drops the open step without finishing it (the state keeps the decayed
  // memory m~ the start wrote).
  void OpenCancel() { open_.state = nullptr; }
  DeviceState* OpenState() const { return open_.state; }
  // RelativeChange (src/Viscoelastic3D.cpp:1636-1669) over every entry.
  double RelativeChangeOf(const double* next, const double* now) {
    return RelativeChange(next, now, true);
  }
  // The iterate (the start strain after OpenStart, the converged strain
  // after OpenConverge), device layout.
  double* Iterate() { return s_.data(); }

  // GridCoupling's maps at the boundary, in the reference's solution
  // layout (every task's ndof x (l + 1) complex columns, column-major, end
  // to end): SpheroidalForce / ToroidalForce (src/GridCoupling.cpp:
  // 691-809) of a device-layout spectral tensor, and StrainOf's
  // AddStrainOfSolution (:595-689) of solutions (no channels). Both use
  // their own buffers, not the step's.
  size_t SolutionDoubles() const { return solutionDoubles_; }
  void ForceOf(const double* weighted, double* forces);
  void StrainOfSolutions(const double* solutions, double* spectral);

 private:
  StepInfo StepOnce(DeviceState& state, double dt, int scheme, const Channels& channels,
                    const double* start, double warmGuard);
  StepInfo StepSdirk2(DeviceState& state, double dt, const Channels& channels);
  // Converge for several realisations side by side: every pass takes all
  // of them (one launch per kernel), each realisation keeps its own
  // residual, adaptive switch and Anderson history (launched on its runs),
  // and stops changing once it converges, so that each follows the passes
  // it takes alone.
  StepInfo ConvergeBatch(const PassFn& pass, const double* start, double warmGuard);
  void RequireOneRealisation(const char* what) const;
  std::vector<StepInfo> realisationInfos_;
  // realisations side by side: per realisation nonzero while it converges
  // (a pass skips the others); every one nonzero outside ConvergeBatch
  DeviceArray<int> live_;
  const int* Live() const { return coupling_.Realisations() > 1 ? live_.data() : nullptr; }
  // the mixes' gammas of one pass (per realisation kAndersonMax): pinned on
  // the host, uploaded in one copy
  double* gammaHost_ = nullptr;
  DeviceArray<double> gammaDev_;
  void SetAllLive();
  // OpenStep::Forces + Solve + StrainOf (src/Viscoelastic3D.cpp:2008-2015):
  // the step's system by conjugate gradients on the memory dofs, its strain
  // coefficients into s_.
  StepInfo CloseByConjugateGradients(Flavour& f);
  // The lateral force of a response on the memory dofs (with the channels'
  // unit responses added when `channels`): strain, synthesis, the weight,
  // analysis and the force map, into `force` (the force layout).
  void LateralForce(const double* response, bool channels, double* force, Flavour& f);
  // LateralPair of a response and a force (minus g when given), on the host.
  double Pair(const double* x, const double* force, const double* g, const ChannelTerms& ch);
  void CopyState(const DeviceState& from, DeviceState& to) const;
  void BuildTasks();
  void BuildResponse(Flavour& f);
  void SolveAll(const double* weighted, Flavour& f);
  // The channels' columns at t into amp_ (load, then tide); hasTide_ says
  // whether the tide half is in use.
  void Amplitudes(const Channels& channels, double t);
  // Throws when the tide's degree-1 columns at t are not zero.
  void CheckTide(const Channels& channels, double t) const;
  // RelativeChange (src/Viscoelastic3D.cpp:1636-1669): over every entry,
  // or over the active ones (where the inactive ones are zero in both).
  double RelativeChange(const double* next, const double* now, bool everyEntry);
  void ElasticStrainCoefficients(DeviceState& state, const Channels& channels, double* out);
  const double* TideAmplitudes() const {
    return hasTide_ ? amp_.data() + ampDoubles_ : nullptr;
  }

  std::shared_ptr<const Earth> earth_;
  StepOptions options_;
  Coupling coupling_;
  std::vector<double> muNodeHost_, itauRefHost_;
  DeviceArray<double> muNode_, itauRefNode_, itau_;
  std::vector<DegreeTask> tasks_;  // spheroidal l = 0..L, then toroidal 1..L
  DeviceArray<DegreeTask> tasksDev_;
  std::vector<int> dofsHost_;
  DeviceArray<int> dofs_;
  size_t rhsDoubles_ = 0, scaleDoubles_ = 0, unitDoubles_ = 0;
  // the memory dofs: per task their dof numbers (dm), and the dof table
  // with memory-dof rows in place of dof numbers
  std::vector<int> dmHost_, dofsMemHost_;
  DeviceArray<int> dm_, dofsMem_;
  size_t gDoubles_ = 0, fDoubles_ = 0;
  int nmMax_ = 0, nmPadMax_ = 0, colsMax_ = 0;
  DeviceArray<ResponseWork> work_;
  int workCount_ = 0, narrowCount_ = 0;
  DeviceArray<int2> forceWork_;  // the force kernel's (degree, order chunk) list
  int forceWorkCount_ = 0;
  DeviceArray<PairWork> pairWork_;  // the pairing's tiles
  int pairWorkCount_ = 0;
  // the conjugate gradients' memory-dof vectors (allocated on first use):
  // the force so far and its response, the residual and its response, the
  // direction and its response, the direction's lateral force
  DeviceArray<double> cgPhi_, cgY_, cgRho_, cgGrho_, cgPi_, cgGpi_, cgQ_, cgPartials_;
  std::list<std::unique_ptr<Flavour>> flavours_;
  // the open step's buffers
  DeviceArray<double> weight_, beta_, fixed_, analysed_, amp_, s_, sNext_;
  // the transposed pass's buffers: zero channel amplitudes, its strain, the
  // source's analysis
  DeviceArray<double> zeroAmp_, strainT_, sourceT_;
  DeviceArray<double> coldStart_;  // the step's own start under a guarded warm start
  size_t ampDoubles_ = 0;  // one channel's amplitudes, (L + 1)(L + 2)
  std::vector<double> ampHost_;  // the last amplitudes gathered (load, then tide)
  struct {
    int scheme = -1;  // none yet
    double dt = 0.0;
  } lastStep_;  // the flavour of the last pass, for the solution export
  bool hasTide_ = false;
  // the open step: its state, flavour, end time, and whether analysed_
  // holds the start strain's force for the first pass
  struct {
    DeviceState* state = nullptr;
    Flavour* flavour = nullptr;
    int scheme = 0;
    double dt = 0.0, tEnd = 0.0;
    bool startForce = false;
  } open_;
  void RequireOpen() const;
  DeviceArray<double> force_, response_;  // per pass, on the memory dofs
  DeviceArray<double> boundaryRows_;      // ForceOf / StrainOfSolutions
  DeviceArray<long long> solutionOffset_;  // per task, in the solution layout
  size_t solutionDoubles_ = 0;
  DeviceArray<double> rhs_;               // full solves for export only
  bool rhsCurrent_ = false;               // rhs_ holds the solves of force_
  DeviceArray<double> fourierA_[3], fourierB_[3];
  DeviceArray<double> partials_, result_;
  DeviceArray<double> history_, xPrev_, fPrev_;
  std::unique_ptr<DeviceState> stage_;  // Sdirk2's first stage
  hipEvent_t events_[4] = {};
  // a pass's few result doubles come back through pinned memory, the host
  // polling an event behind the copy (ReadBack)
  static constexpr int kReadBackMax = 40;  // the sweep's 2 x 16 sums and 3 maxima
  // (per realisation)
  double* pinned_ = nullptr;
  hipEvent_t readEvent_ = nullptr;
  void ReadBack(double* dst, const double* src, int count);
  // The residual's and the sweep's results reach the host without a copy:
  // a kernel writes them to host-mapped memory (per realisation
  // kReadBackMax doubles and a ticket word), the host polls the ticket
  // (Publish, then Published).
  double* mapped_ = nullptr;  // host-mapped: values, then the ticket words
  double* mappedValuesDev_ = nullptr;
  unsigned long long* mappedSeqDev_ = nullptr;
  unsigned long long ticket_ = 0;
  HostPublish Publish(int r);
  // Waits for realisation r's ticket `ticket`; its values then.
  const double* Published(int r, unsigned long long ticket);
  // The pass's kernels on the stream (what Pass captures).
  void LaunchPass(const double* s, double* sNext, bool useStartForce, Flavour& f);
  // A pass captured as a graph, keyed by the tensors it reads and writes,
  // its flavour and whether the tide channel is in use (the pointers its
  // kernels take; every other buffer is fixed for the operator's life).
  struct PassGraph {
    const double* s = nullptr;
    double* next = nullptr;
    unsigned long long flavour = 0;
    bool tide = false;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<PassGraph> passGraphs_;
  unsigned long long flavourCount_ = 0;
  // Destroys the captured passes of one flavour (0: all).
  void DropPassGraphs(unsigned long long flavour);
};

}  // namespace sl3amd
