/* The C++ face of the C ABI (sl3dpp_amd.h), header only: the reference's
 * Viscoelastic3D interface (include/sl3dpp/Viscoelastic3D.hpp:89-440) with
 * its names, argument meanings and exception types, for C++ callers of the
 * reference (Evolution3D's step loop, src/Evolution3D.cpp:759-800;
 * benchmarks/step_timing.cpp).
 *
 *  - Handles are RAII (Earth, Viscoelastic3D, State); moves transfer them.
 *  - Status codes become the reference's exceptions: SL3_EINVAL ->
 *    std::invalid_argument, SL3_ENUMERIC (a factorisation failed) ->
 *    std::runtime_error, a device failure -> std::runtime_error.
 *  - LoadCallable3D is the reference's: degree l's complex m >= 0 columns
 *    at time t, or nothing for no load at that degree. An exception thrown
 *    by it ends the step (before the state changes, as the channels are
 *    gathered first; in Sdirk2's second stage after the stages combine)
 *    and is rethrown from Step.
 *  - Step(state, ...) returns a new State and leaves its argument as it
 *    was, as the reference's does (a device copy of the state);
 *    StepInPlace is the boundary's own in-place form.
 *  - The lateral field delta(r, theta, phi) is sampled by the caller on
 *    every memory node's grid and passed as a device array ([node][theta]
 *    [phi]), where the reference takes a functor. */
#ifndef SL3DPP_AMD_HPP
#define SL3DPP_AMD_HPP

#include <complex>
#include <cstdint>
#include <exception>
#include <functional>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sl3dpp_amd.h"

namespace sl3dpp_amd {

using Complex = std::complex<double>;

inline void Check(int status) {
  if (status == SL3_OK) return;
  const std::string message = sl3_last_error();
  if (status == SL3_EINVAL) throw std::invalid_argument(message);
  throw std::runtime_error(message);
}

/* The layered Earth and its radial set-up (host). */
class Earth {
 public:
  explicit Earth(const sl3_earth_spec& spec) { Check(sl3_earth_create(&spec, &handle_)); }
  ~Earth() {
    if (handle_) sl3_earth_destroy(handle_);
  }
  Earth(const Earth&) = delete;
  Earth& operator=(const Earth&) = delete;
  Earth(Earth&& other) noexcept : handle_(std::exchange(other.handle_, nullptr)) {}
  const sl3_earth* Handle() const { return handle_; }

 private:
  sl3_earth* handle_ = nullptr;
};

/* Scheme (include/sl3dpp/StepWeights.hpp, StepRecipe.hpp). */
enum class Scheme {
  ExponentialTrapezoid = SL3_EXPONENTIAL_TRAPEZOID,
  BackwardEuler = SL3_BACKWARD_EULER,
  Sdirk2 = SL3_SDIRK2,
};

class Viscoelastic3D {
 public:
  /* Viscoelastic3D::Options (include/sl3dpp/Viscoelastic3D.hpp:89-187),
   * the members this step reads; the defaults are the reference's. */
  struct Options {
    enum class StepSolver { FixedPoint, ConjugateGradients };
    double eps = 1e-8;
    double oversampling = 1.5;
    int max_iterations = 40;
    double rtol = 1e-11;
    double damping = 1.0;
    int anderson_memory = 5;
    double anderson_rate = 0.5;
    StepSolver solver = StepSolver::FixedPoint;
    double solve_rtol = 1e-11;
    int solve_max_iterations = 200;
    /* not in the reference: each pass replays a captured hipGraph */
    bool graph_passes = false;
    /* the degrees the transforms may be cut to with depth (at most
     * SL3_MAX_TRANSFORM_DEGREES) */
    std::vector<int> transform_degrees{};
    /* not in the reference: independent realisations stepped side by side */
    int realisations = 1;
  };

  /* Viscoelastic3D::StepInfo (:357-375). */
  struct StepInfo {
    int iterations = 0;
    double residual = 0.0;
    bool converged = true;
    int mixedPasses = 0;
    int solveIterations = 0;
    bool warmFallback = false;  // a guarded warm start was dropped
  };

  /* LoadCallable3D: degree l's complex m >= 0 columns at time t. */
  using LoadCallable3D = std::function<std::optional<std::vector<Complex>>(int, double)>;

  /* The step's state on the device (Viscoelastic3D::State, :295-309). */
  class State {
   public:
    State() = default;
    ~State() {
      if (handle_) sl3_state_destroy(handle_);
    }
    State(const State&) = delete;
    State& operator=(const State&) = delete;
    State(State&& other) noexcept : handle_(std::exchange(other.handle_, nullptr)) {}
    State& operator=(State&& other) noexcept {
      if (this != &other) {
        if (handle_) sl3_state_destroy(handle_);
        handle_ = std::exchange(other.handle_, nullptr);
      }
      return *this;
    }
    double t() const {
      double v = 0.0;
      Check(sl3_state_time(handle_, &v));
      return v;
    }
    sl3_state* Handle() const { return handle_; }

   private:
    friend class Viscoelastic3D;
    explicit State(sl3_state* h) : handle_(h) {}
    sl3_state* handle_ = nullptr;
  };

  /* Per-degree solutions (Viscoelastic3D::Solutions, :326-329):
   * ndof x (l + 1) complex, column-major, spheroidal from degree 0 and
   * toroidal from degree 1. */
  struct Solutions {
    std::map<int, std::vector<Complex>> spheroidal;
    std::map<int, std::vector<Complex>> toroidal;
  };

  /* Viscoelastic3D::StepResult (:401-405). */
  struct StepResult {
    State state;
    Solutions solutions;
    StepInfo info;
  };

  Viscoelastic3D(const Earth& earth, int lmax, const double* delta_dev, const Options& options) {
    const sl3_step_options o = ToC(options);
    Check(sl3_operator_create(earth.Handle(), lmax, delta_dev, &o, &handle_));
  }
  Viscoelastic3D(const Earth& earth, int lmax, const double* delta_dev)
      : Viscoelastic3D(earth, lmax, delta_dev, Options()) {}
  ~Viscoelastic3D() {
    if (handle_) sl3_operator_destroy(handle_);
  }
  Viscoelastic3D(const Viscoelastic3D&) = delete;
  Viscoelastic3D& operator=(const Viscoelastic3D&) = delete;
  Viscoelastic3D(Viscoelastic3D&& other) noexcept
      : handle_(std::exchange(other.handle_, nullptr)) {}

  void SetOptions(const Options& options) {
    const sl3_step_options o = ToC(options);
    Check(sl3_operator_set_options(handle_, &o));
  }

  /* ZeroState (:310). */
  State ZeroState() const {
    sl3_state* s = nullptr;
    Check(sl3_state_zero(handle_, &s));
    return State(s);
  }

  /* Step (:407-410): the new state, the end-time solutions and the
   * iteration diagnostics; the argument is left as it was. */
  StepResult Step(const State& state, double dt, Scheme scheme = Scheme::Sdirk2,
                  const LoadCallable3D& load = {}, const LoadCallable3D& tide = {}) const {
    sl3_state* copy = nullptr;
    Check(sl3_state_clone(handle_, state.Handle(), &copy));
    State next(copy);
    StepResult result{std::move(next), {}, {}};
    result.info = StepInPlace(result.state, dt, scheme, load, tide);
    result.solutions = LastSolutions();
    return result;
  }

  /* The step in place on the state (the boundary's form, no copy). */
  StepInfo StepInPlace(State& state, double dt, Scheme scheme = Scheme::Sdirk2,
                       const LoadCallable3D& load = {}, const LoadCallable3D& tide = {}) const {
    return StepInPlace(state, dt, scheme, load, tide, nullptr, 0.0);
  }

  /* The step from a warm start: Converge's `start` (:1916-1921) as a
   * device-layout spectral tensor; guard > 0 drops a start whose first pass
   * changes it by more than guard (StepInfo::warmFallback). */
  StepInfo StepInPlace(State& state, double dt, Scheme scheme, const LoadCallable3D& load,
                       const LoadCallable3D& tide, const double* start_dev, double guard) const {
    return WithChannels(load, tide, [&](sl3_load_fn l, sl3_load_fn t, void* user,
                                        sl3_step_info* info) {
      return sl3_step_guarded(handle_, state.Handle(), dt, static_cast<int>(scheme), l, t, user,
                              start_dev, guard, info);
    });
  }

  /* Warm starts along a trajectory: the Lagrange extrapolation in time of
   * the last order + 1 converged strains of the states stepped through it
   * (sl3_extrapolator). */
  class Extrapolator {
   public:
    Extrapolator(const Viscoelastic3D& op, int order = 3) {
      Check(sl3_extrapolator_create(op.Handle(), order, &handle_));
    }
    ~Extrapolator() {
      if (handle_) sl3_extrapolator_destroy(handle_);
    }
    Extrapolator(const Extrapolator&) = delete;
    Extrapolator& operator=(const Extrapolator&) = delete;
    void Reset() { Check(sl3_extrapolator_reset(handle_)); }
    sl3_extrapolator* Handle() const { return handle_; }

   private:
    sl3_extrapolator* handle_ = nullptr;
  };

  /* A trajectory's step in place, warm-started by the extrapolator (the
   * step's own start until it holds two steps of this state's history). */
  StepInfo StepInPlace(State& state, Extrapolator& ex, double dt, Scheme scheme,
                       const LoadCallable3D& load = {}, const LoadCallable3D& tide = {},
                       double guard = 0.0) const {
    return WithChannels(load, tide, [&](sl3_load_fn l, sl3_load_fn t, void* user,
                                        sl3_step_info* info) {
      return sl3_extrapolated_step(ex.Handle(), state.Handle(), dt, static_cast<int>(scheme), l,
                                   t, user, guard, info);
    });
  }

  /* The last step's solutions on every dof (its last pass's, with the
   * channels), per degree. */
  Solutions LastSolutions() const {
    const std::vector<double> tasks = Export("tasks");
    Solutions out;
    for (size_t task = 0; 5 * task + 4 < tasks.size(); ++task) {
      const int l = static_cast<int>(tasks[5 * task]);
      const int kind = static_cast<int>(tasks[5 * task + 1]);
      const std::vector<double> x = Export("solution:" + std::to_string(task));
      std::vector<Complex> columns(x.size() / 2);
      for (size_t i = 0; i < columns.size(); ++i) columns[i] = Complex(x[2 * i], x[2 * i + 1]);
      (kind == 0 ? out.spheroidal : out.toroidal)[l] = std::move(columns);
    }
    return out;
  }

  /* sl3_operator_export by name (host copy). */
  std::vector<double> Export(const std::string& name) const {
    int64_t count = 0;
    Check(sl3_operator_export(handle_, name.c_str(), nullptr, &count));
    std::vector<double> out(static_cast<size_t>(count));
    Check(sl3_operator_export(handle_, name.c_str(), out.data(), &count));
    return out;
  }

  /* sl3_state_export by name (host copy). */
  std::vector<double> Export(const State& state, const std::string& name) const {
    int64_t count = 0;
    Check(sl3_state_export(handle_, state.Handle(), name.c_str(), nullptr, &count));
    std::vector<double> out(static_cast<size_t>(count));
    Check(sl3_state_export(handle_, state.Handle(), name.c_str(), out.data(), &count));
    return out;
  }

  /* Not in the reference: realisations side by side
   * (Options::realisations). Realisation r's memory nodes are nodes
   * r N .. r N + N - 1 of every node-indexed tensor, N = RadialNodes(). */
  int Realisations() const {
    int r = 0;
    Check(sl3_operator_realisations(handle_, &r, nullptr));
    return r;
  }
  int RadialNodes() const {
    int n = 0;
    Check(sl3_operator_realisations(handle_, nullptr, &n));
    return n;
  }
  /* Per realisation, the last step's own StepInfo. */
  std::vector<StepInfo> RealisationInfos() const {
    int count = 0;
    Check(sl3_step_realisation_infos(handle_, 0, &count, nullptr));
    std::vector<sl3_step_info> c(static_cast<size_t>(count));
    Check(sl3_step_realisation_infos(handle_, count, &count, c.data()));
    std::vector<StepInfo> out;
    for (const auto& i : c) {
      out.push_back(StepInfo{i.iterations, i.residual, i.converged != 0, i.mixed_passes,
                             i.solve_iterations, i.warm_fallback != 0});
    }
    return out;
  }

  sl3_operator* Handle() const { return handle_; }

 private:
  struct Channel {
    const LoadCallable3D* fn;
    std::exception_ptr error;
  };

  // Runs a step call with the load and tide channels bound to the C
  // callbacks; a channel's own exception is rethrown, then the status.
  template <class F>
  static StepInfo WithChannels(const LoadCallable3D& load, const LoadCallable3D& tide, F&& call) {
    Channel channels[2] = {{&load, nullptr}, {&tide, nullptr}};
    sl3_step_info info{};
    const int status = call(load ? &Trampoline : nullptr, tide ? &Trampoline2 : nullptr,
                            static_cast<void*>(channels), &info);
    for (const auto& c : channels) {
      if (c.error) std::rethrow_exception(c.error);
    }
    Check(status);
    return StepInfo{info.iterations, info.residual, info.converged != 0, info.mixed_passes,
                    info.solve_iterations, info.warm_fallback != 0};
  }

  static int Call(Channel& c, int l, double t, double* columns) {
    if (c.error) return -1;
    try {
      const auto cols = (*c.fn)(l, t);
      if (!cols) return 0;
      if (cols->size() != static_cast<size_t>(l + 1)) {
        throw std::invalid_argument("a load channel returned " + std::to_string(cols->size()) +
                                    " columns for degree " + std::to_string(l));
      }
      for (int m = 0; m <= l; ++m) {
        columns[2 * m] = (*cols)[m].real();
        columns[2 * m + 1] = (*cols)[m].imag();
      }
      return 1;
    } catch (...) {
      c.error = std::current_exception();
      return -1;  // the library ends the step before the state changes
    }
  }
  static int Trampoline(int l, double t, double* columns, void* user) {
    return Call(static_cast<Channel*>(user)[0], l, t, columns);
  }
  static int Trampoline2(int l, double t, double* columns, void* user) {
    return Call(static_cast<Channel*>(user)[1], l, t, columns);
  }

  static sl3_step_options ToC(const Options& o) {
    sl3_step_options c{};
    sl3_step_options_default(&c);
    c.eps = o.eps;
    c.oversampling = o.oversampling;
    c.max_iterations = o.max_iterations;
    c.rtol = o.rtol;
    c.damping = o.damping;
    c.anderson_memory = o.anderson_memory;
    c.anderson_rate = o.anderson_rate;
    c.solver = o.solver == Options::StepSolver::ConjugateGradients ? SL3_CONJUGATE_GRADIENTS
                                                                     : SL3_FIXED_POINT;
    c.solve_rtol = o.solve_rtol;
    c.solve_max_iterations = o.solve_max_iterations;
    c.graph_passes = o.graph_passes ? 1 : 0;
    if (o.transform_degrees.size() > SL3_MAX_TRANSFORM_DEGREES) {
      throw std::invalid_argument("at most 32 transform degrees");
    }
    c.transform_degree_count = static_cast<int>(o.transform_degrees.size());
    for (size_t i = 0; i < o.transform_degrees.size(); ++i) {
      c.transform_degrees[i] = o.transform_degrees[i];
    }
    c.realisations = o.realisations;
    return c;
  }

  sl3_operator* handle_ = nullptr;
};

}  // namespace sl3dpp_amd

#endif  // SL3DPP_AMD_HPP
