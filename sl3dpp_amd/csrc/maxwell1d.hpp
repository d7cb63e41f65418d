// The radially symmetric (1D) Maxwell evolution in the degree basis on the
// GPU: ViscoelasticOperator::Step (include/sl3dpp/Viscoelastic.hpp:283-301,
// src/Viscoelastic.cpp:914-1076) for every spheroidal degree at once, one
// branch, the memory held as the three deviatoric scalars (M, R, S) per
// (degree, Gauss memory point, real column). A step is four batched
// launches over all degrees (the decayed memory's force, the Cholesky and
// LU band solves with the step's factors, strain and memory update) and a
// counter bump, and takes no host decision, so a trajectory replays one
// captured hipGraph per step. The load is one separable channel, a fixed
// pattern of surface-mass coefficients times a ramp sampled at the step
// times.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "common.hpp"
#include "degree_systems.hpp"
#include "kernels.hpp"
#include "model.hpp"

namespace sl3amd {

// One degree's block of the 1D state and systems.
struct Degree1D {
  int l = 0;
  int ndof = 0;
  int cols = 0;        // 2 (l + 1) real columns: Re of m = 0..l, then Im
  int firstNode = 0;   // the degree's first memory node
  int pad = 0;
  long long memOff = 0;   // memory / strain block: [node][scalar][col]
  long long rhsOff = 0;   // ndof x cols, row-major [dof][col]
  long long dofOff = 0;   // per mesh node: (U, V, phi) dofs
  long long unitOff = 0;  // the unit load response (ndof)
};

// The factors of every degree's system for one (scheme, dt), with the
// per-node step weights and the unit load responses.
struct Flavour1D {
  int scheme = 0;
  double dt = 0.0;
  DeviceArray<double> scale, factors, units;
  DeviceArray<int> pivots;
  std::vector<BandSystem> cholesky, lu;
  DeviceArray<BandSystem> choleskyDev, luDev;
  DeviceArray<int2> choleskyTasks, luTasks;
  int choleskyTaskCount = 0, luTaskCount = 0, kdMax = 0, kvMax = 0;
  // per memory node: e = the memory's decay, alpha, beta (m+ = e m +
  // alpha d0 + beta d)
  DeviceArray<double> e, alpha, beta;
  // the Cholesky solve's band panel (rows per staged panel) and the ms per
  // solve the tuner timed for each choice
  int panelRows = 0;
  double panelMs[kPanelChoices] = {};
};

class Maxwell1D {
 public:
  Maxwell1D(std::shared_ptr<const Earth> earth, int lmax, double eps);
  ~Maxwell1D();
  Maxwell1D(const Maxwell1D&) = delete;
  Maxwell1D& operator=(const Maxwell1D&) = delete;
  int LMax() const { return lmax_; }
  int Nodes() const { return nodes_; }
  const std::vector<Degree1D>& Degrees() const { return degrees_; }
  double Time() const { return t_; }
  size_t MemoryDoubles() const { return memDoubles_; }
  size_t SolutionDoubles() const { return rhsDoubles_; }

  // ZeroState.
  void Reset();
  // The load channel: a fixed pattern (complex per (l, m) at l(l+1)/2 + m,
  // non-dimensional surface mass) times a ramp r_k at the step times
  // t_k = t0 + k dt, k = 0..steps.
  void SetLoad(const std::vector<double>& pattern, const std::vector<double>& ramp);
  // steps steps of length dt with the scheme (exponential trapezoid or
  // backward Euler); the first from rest takes its elastic start under
  // the load at t0. graph: each step one replay of a captured hipGraph.
  void Evolve(double dt, int scheme, int steps, bool graph);
  // Host copies: the memory, the strain (the layout of Degree1D::memOff),
  // the solutions of the last step (unscaled, with the channel; layout of
  // rhsOff, row-major [dof][col]), the step index.
  std::vector<double> Memory() const;
  std::vector<double> Strain() const;
  std::vector<double> Solutions() const;
  // The step flavour's band panel and the ms per solve timed for each
  // choice (4, 8, 16 rows), of the last Evolve.
  std::vector<double> Tuning() const { return tuning_; }
  int StepIndex() const { return stepIndex_; }
  // One component's (0 U, 1 V, 2 phi) surface value of the last solutions
  // per degree and column, at out[l (l + 1) + col] ((lmax+1)(lmax+2)
  // doubles on the device, zero where the component has no dof;
  // ViscoelasticOperator::Surface, include/sl3dpp/Viscoelastic.hpp:262-267).
  void Surface(int component, double* out_dev, hipStream_t s) const;

 private:
  const Flavour1D& FlavourOf(int scheme, double dt);
  void EnqueueStep(const Flavour1D& f, hipStream_t s);
  void ElasticStart(const Flavour1D& f, hipStream_t s);
  void DropGraph();

  std::shared_ptr<const Earth> earth_;
  int lmax_;
  double eps_;
  int nodes_ = 0, np_ = 0, ngll_ = 0, elem0_ = 0, elements_ = 0, meshNodes_ = 0;
  std::vector<Degree1D> degrees_;
  DeviceArray<Degree1D> degreesDev_;
  DeviceArray<int> dofs_;  // per degree: (U, V, phi) dof of every mesh node
  std::vector<int> surfaceDof_[3];
  DeviceArray<int> surfaceDofDev_[3];
  DeviceArray<double> radius_, jac_, pweight_, pvalue_, pderiv_, mu_;
  std::vector<double> itauRef_;  // per memory node
  size_t memDoubles_ = 0, rhsDoubles_ = 0, unitDoubles_ = 0;
  int colsMax_ = 0;
  DeviceArray<double> m_, d_, rhs_, pattern_, ramp_;
  DeviceArray<int> counter_;  // the device's step index
  int stepIndex_ = 0;
  double t_ = 0.0;
  std::vector<double> tuning_;
  bool hasStrain_ = false;
  std::vector<std::unique_ptr<Flavour1D>> flavours_;
  const Flavour1D* last_ = nullptr;  // the last step's flavour
  // the captured step of one flavour
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graphExec_ = nullptr;
  const Flavour1D* graphFlavour_ = nullptr;
};

}  // namespace sl3amd
