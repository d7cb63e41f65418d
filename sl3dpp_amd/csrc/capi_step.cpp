// C ABI of the Earth set-up, the 3D step and the batched transforms
// (include/sl3dpp_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "common.hpp"
#include "coupling.hpp"
#include "kernels.hpp"
#include "learned.hpp"
#include "step.hpp"

#include "capi_handles.hpp"

struct sl3_operator {
  std::unique_ptr<sl3amd::Viscoelastic3D> op;
  // per degree cut: the device-layout offsets of the rows of degree <= cut
  // (the learned warm start's gather and scatter)
  std::mutex rowsMutex;
  std::map<int, std::pair<int, sl3amd::DeviceArray<long long>>> cutRows;
};

struct sl3_state {
  std::unique_ptr<sl3amd::DeviceState> state;
  // the operator whose grid and spectral sizes the state's buffers have:
  // every entry point taking both checks it before any launch
  const sl3_operator* owner = nullptr;
  uint64_t id = 0;  // unique per state made (zero or clone)
};

// The converged strain coefficients of a trajectory's last steps (device
// copies, oldest first) and the buffer their extrapolation is formed in.
struct sl3_extrapolator {
  sl3_operator* op = nullptr;
  int order = 3;
  std::vector<sl3amd::DeviceArray<double>> buffers;
  sl3amd::DeviceArray<double> start;
  std::vector<std::pair<double, int>> recorded;  // (time, buffer)
  uint64_t last = 0;  // the id of the state they are of (0: none)
};

using namespace sl3amd;

namespace {

void Check(bool ok, const std::string& what) {
  if (!ok) throw ArgumentError(what);
}

std::atomic<uint64_t> g_nextStateId{1};

// A state's buffers are sized by the operator that made it.
void CheckOwner(const sl3_operator* op, const sl3_state* state) {
  Check(op != nullptr && state != nullptr, "null argument");
  Check(state->owner == op, "the state belongs to another operator");
}

StepOptions ToOptions(const sl3_step_options* o) {
  StepOptions s;
  if (o == nullptr) return s;
  s.eps = o->eps;
  s.oversampling = o->oversampling;
  s.maxIterations = o->max_iterations;
  s.rtol = o->rtol;
  s.damping = o->damping;
  s.andersonMemory = o->anderson_memory;
  s.andersonRate = o->anderson_rate;
  s.solver = o->solver;
  s.solveRtol = o->solve_rtol;
  s.solveMaxIterations = o->solve_max_iterations;
  s.graphPasses = o->graph_passes != 0;
  Check(o->transform_degree_count >= 0 &&
            o->transform_degree_count <= SL3_MAX_TRANSFORM_DEGREES,
        "transform_degree_count in [0, 32]");
  s.transformDegrees.assign(o->transform_degrees,
                            o->transform_degrees + o->transform_degree_count);
  s.realisations = o->realisations;
  Check(s.realisations >= 1 && s.realisations <= kRealisationsMax, "realisations in [1, 64]");
  Check(s.realisations == 1 || (s.transformDegrees.empty() && s.solver == kFixedPoint),
        "realisations side by side take the fixed point and no transform degrees");
  Check(s.solver == kFixedPoint || s.solver == kConjugateGradients,
        "solver: SL3_FIXED_POINT or SL3_CONJUGATE_GRADIENTS");
  Check(s.solveMaxIterations >= 0, "solve_max_iterations >= 0");
  Check(s.andersonMemory >= 0 && s.andersonMemory <= kAndersonMax, "anderson memory in [0, 16]");
  Check(s.maxIterations >= 1, "max_iterations >= 1");
  Check(s.eps > 0.0 && s.eps < 1.0, "eps in (0, 1)");
  Check(s.oversampling >= 1.0, "oversampling >= 1");
  return s;
}

// Copies n doubles from the device to the host (dst), or reports n.
void Export(const double* src_dev, size_t n, double* dst, int64_t* count) {
  if (count) *count = static_cast<int64_t>(n);
  if (dst == nullptr) return;
  hipStream_t s = Stream();
  SL3_HIP(hipMemcpyAsync(dst, src_dev, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SL3_HIP(hipStreamSynchronize(s));
}

void ExportHost(const std::vector<double>& v, double* dst, int64_t* count) {
  if (count) *count = static_cast<int64_t>(v.size());
  if (dst) std::memcpy(dst, v.data(), sizeof(double) * v.size());
}

size_t ReferenceBlockSize(int block, int lmax) {
  const size_t L1 = static_cast<size_t>(lmax + 1);
  return block == 2 ? L1 * (lmax + 2) / 2 : L1 * L1 - (block == 0 ? 4 : 1);
}

void ToReference(const Coupling& c, const double* dev, double* ref, hipStream_t s) {
  size_t refOff = 0;
  for (int b = 0; b < 3; ++b) {
    LaunchToReferenceLayout(dev + c.SpectralOffset(b), ref + refOff, c.RowOffsetDevice(b),
                            SurfaceGrid::SpinOf(b), c.LMax(), c.NodeCount(), s);
    refOff += 2 * ReferenceBlockSize(b, c.LMax()) * c.NodeCount();
  }
}

void FromReference(const Coupling& c, const double* ref, double* dev, hipStream_t s) {
  size_t refOff = 0;
  for (int b = 0; b < 3; ++b) {
    LaunchToDeviceLayout(ref + refOff, dev + c.SpectralOffset(b), c.RowOffsetDevice(b),
                         SurfaceGrid::SpinOf(b), c.LMax(), c.NodeCount(), s);
    refOff += 2 * ReferenceBlockSize(b, c.LMax()) * c.NodeCount();
  }
}

int BlockOfSpin(int spin) {
  if (spin == -2) return 0;
  if (spin == -1) return 1;
  if (spin == 0) return 2;
  throw ArgumentError("the stored spins are -2, -1 and 0");
}

}  // namespace

extern "C" {

int sl3_earth_create(const sl3_earth_spec* spec, sl3_earth** out) {
  return Guard([&] {
    Check(spec != nullptr && out != nullptr, "null argument");
    Check(spec->layers >= 1 && spec->layers <= SL3_MAX_LAYERS, "1 to 8 layers");
    EarthSpec es;
    for (int i = 0; i < spec->layers; ++i) {
      es.layers.push_back({spec->r0[i], spec->r1[i], spec->rho0[i], spec->rho1[i],
                           spec->kappa0[i], spec->kappa1[i], spec->mu0[i], spec->mu1[i],
                           spec->elements[i]});
    }
    es.ngll = spec->ngll;
    es.viscousLayer = spec->viscous_layer;
    es.viscosityBoundary = spec->viscosity_boundary;
    es.viscosityUpper = spec->viscosity_upper;
    es.viscosityLower = spec->viscosity_lower;
    auto e = std::make_unique<sl3_earth>();
    e->earth = BuildEarth(es);
    *out = e.release();
  });
}

int sl3_earth_destroy(sl3_earth* earth) {
  delete earth;
  return SL3_OK;
}

int sl3_earth_scales(const sl3_earth* earth, double* length, double* time,
                     double* mass) {
  return Guard([&] {
    Check(earth != nullptr, "null earth");
    const auto& s = earth->earth->scales;
    if (length) *length = s.length;
    if (time) *time = s.Time();
    if (mass) *mass = s.Mass();
  });
}

int sl3_earth_mesh(const sl3_earth* earth, int* elements, int* ngll) {
  return Guard([&] {
    Check(earth != nullptr, "null earth");
    if (elements) *elements = earth->earth->mesh->ElementCount();
    if (ngll) *ngll = earth->earth->mesh->NGll();
  });
}

int sl3_earth_export(const sl3_earth* earth, const char* name, double* dst,
                     int64_t* count) {
  return Guard([&] {
    Check(earth != nullptr && name != nullptr, "null argument");
    const auto& e = *earth->earth;
    const auto& mesh = *e.mesh;
    const auto& pts = e.reference.points;
    const std::string key(name);
    std::vector<double> v;
    if (key == "radius") {
      v = mesh.Radii();
    } else if (key == "jac") {
      for (int k = 0; k < mesh.ElementCount(); ++k) v.push_back(mesh.Jacobian(k));
    } else if (key == "gll_weight") {
      v = mesh.Weights();
    } else if (key == "gll_deriv") {
      for (int k = 0; k < mesh.NGll(); ++k)
        for (int i = 0; i < mesh.NGll(); ++i) v.push_back(mesh.Derivative(k, i));
    } else if (key == "point_value") {
      v = pts.value;
    } else if (key == "point_deriv") {
      v = pts.derivative;
    } else if (key == "point_weight") {
      v = pts.weight;
    } else if (key == "point_radius") {
      for (int k = 0; k < mesh.ElementCount(); ++k)
        for (int p = 0; p < pts.count; ++p) v.push_back(pts.RadiusAt(mesh, k, p));
    } else if (key == "point_mu") {
      v = e.reference.pointMu;
    } else if (key == "point_itau") {
      v = e.reference.pointItau;
    } else if (key == "fluid") {
      for (int k = 0; k < mesh.ElementCount(); ++k) v.push_back(e.material.IsFluid(k));
    } else if (key == "surface_gravity") {
      v.push_back(e.material.surfaceGravity);
    } else if (key == "memory_elements") {
      int first = -1, n = 0;
      for (int k = 0; k < mesh.ElementCount(); ++k) {
        if (!e.material.IsFluid(k) && e.rheology.ElementActive(mesh, k)) {
          if (first < 0) first = k;
          ++n;
        }
      }
      v = {static_cast<double>(first), static_cast<double>(n)};
    } else {
      throw ArgumentError("unknown earth export: " + key);
    }
    if (count) *count = static_cast<int64_t>(v.size());
    if (dst) std::memcpy(dst, v.data(), sizeof(double) * v.size());
  });
}

int sl3_earth_degree_system(const sl3_earth* earth, int l, int kind, int scheme,
                            double dt, double eps, int* ndof, int* kd,
                            int* first_element, int* pinned, double* band,
                            int* dofs, double* load, double* tide) {
  return Guard([&] {
    Check(earth != nullptr, "null earth");
    Check(kind == 0 || kind == 1, "kind 0 (spheroidal) or 1 (toroidal)");
    const auto& e = *earth->earth;
    std::vector<double> shear;
    PointShear ps{&e.reference.points, &shear};
    const PointShear* psp = nullptr;
    if (dt >= 0.0) {
      shear = e.reference.PointShearModulus(scheme, dt);
      psp = &ps;
    }
    const auto& mesh = *e.mesh;
    if (kind == 0) {
      DegreeOperator op(e.reference.nodal, l, eps, psp);
      if (ndof) *ndof = op.DofCount();
      if (kd) *kd = op.Kd();
      if (first_element) *first_element = op.FirstElement();
      if (pinned) *pinned = -1;
      if (band) std::memcpy(band, op.Band().data(), sizeof(double) * op.Band().size());
      if (dofs) {
        for (int gn = 0; gn < mesh.NodeCount(); ++gn)
          for (int c = 0; c < 3; ++c) dofs[gn * 3 + c] = op.Dof(gn, Component(c));
      }
      if (load) {
        const auto b = op.Load();
        std::memcpy(load, b.data(), sizeof(double) * b.size());
      }
      if (tide) {
        const auto b = op.Tide(l == 0 ? 2 : -1);
        std::memcpy(tide, b.data(), sizeof(double) * b.size());
      }
    } else {
      ToroidalOperator op(e.reference.nodal, l, eps, psp);
      if (ndof) *ndof = op.DofCount();
      if (kd) *kd = op.Kd();
      if (first_element) *first_element = op.FirstElement();
      // one pinned dof per solid region at degree 1: the first (the
      // innermost region's) here; an Earth with a solid inner core has a
      // second, at the mantle's top, which the device step pins as well
      if (pinned) *pinned = op.Pinned().empty() ? -1 : op.Pinned()[0];
      if (band) std::memcpy(band, op.Band().data(), sizeof(double) * op.Band().size());
      if (dofs) {
        for (int gn = 0; gn < mesh.NodeCount(); ++gn) dofs[gn] = op.Dof(gn);
      }
    }
  });
}

void sl3_step_options_default(sl3_step_options* o) {
  if (o == nullptr) return;
  const StepOptions s;
  o->eps = s.eps;
  o->oversampling = s.oversampling;
  o->max_iterations = s.maxIterations;
  o->rtol = s.rtol;
  o->damping = s.damping;
  o->anderson_memory = s.andersonMemory;
  o->anderson_rate = s.andersonRate;
  o->solver = s.solver;
  o->solve_rtol = s.solveRtol;
  o->solve_max_iterations = s.solveMaxIterations;
  o->graph_passes = s.graphPasses ? 1 : 0;
  o->transform_degree_count = 0;
  for (int i = 0; i < SL3_MAX_TRANSFORM_DEGREES; ++i) o->transform_degrees[i] = 0;
  o->realisations = s.realisations;
}

int sl3_operator_create(const sl3_earth* earth, int lmax, const double* delta_dev,
                        const sl3_step_options* options, sl3_operator** out) {
  return Guard([&] {
    Check(earth != nullptr && out != nullptr, "null argument");
    Check(lmax >= 2, "lmax >= 2");
    auto op = std::make_unique<sl3_operator>();
    op->op = std::make_unique<Viscoelastic3D>(earth->earth, lmax, delta_dev,
                                              ToOptions(options));
    *out = op.release();
  });
}

int sl3_operator_destroy(sl3_operator* op) {
  return Guard([&] {
    if (op) SL3_HIP(hipStreamSynchronize(Stream()));
    delete op;
  });
}

int sl3_operator_set_options(sl3_operator* op, const sl3_step_options* options) {
  return Guard([&] {
    Check(op != nullptr && options != nullptr, "null argument");
    const auto o = ToOptions(options);
    const auto& cur = op->op->Options();
    Check(o.eps == cur.eps && o.oversampling == cur.oversampling,
          "eps and oversampling are fixed at construction");
    auto sorted = [](std::vector<int> d) {
      std::sort(d.begin(), d.end());
      d.erase(std::unique(d.begin(), d.end()), d.end());
      return d;
    };
    Check(sorted(o.transformDegrees) == sorted(cur.transformDegrees),
          "the transform degrees are fixed at construction");
    Check(o.realisations == cur.realisations, "the realisations are fixed at construction");
    op->op->MutableOptions() = o;
  });
}

int sl3_operator_set_delta(sl3_operator* op, const double* delta_dev) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    op->op->SetDelta(delta_dev);
  });
}

int sl3_operator_shape(const sl3_operator* op, int* nodes, int* ntheta, int* nphi,
                       int64_t* spectral_doubles) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    const auto& c = op->op->GetCoupling();
    if (nodes) *nodes = c.NodeCount();
    if (ntheta) *ntheta = c.Grid().NTheta();
    if (nphi) *nphi = c.Grid().NPhi();
    if (spectral_doubles) *spectral_doubles = static_cast<int64_t>(c.SpectralDoubles());
  });
}

int sl3_operator_realisations(const sl3_operator* op, int* realisations, int* radial_nodes) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    const auto& c = op->op->GetCoupling();
    if (realisations) *realisations = c.Realisations();
    if (radial_nodes) *radial_nodes = c.RadialNodeCount();
  });
}

int sl3_step_realisation_infos(const sl3_operator* op, int capacity, int* count,
                               sl3_step_info* infos) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    const auto& all = op->op->RealisationInfos();
    if (count) *count = static_cast<int>(all.size());
    for (int r = 0; r < static_cast<int>(all.size()) && r < capacity && infos; ++r) {
      sl3_step_info& i = infos[r];
      i = sl3_step_info{};
      i.iterations = all[r].iterations;
      i.residual = all[r].residual;
      i.converged = all[r].converged ? 1 : 0;
      i.mixed_passes = all[r].mixedPasses;
      i.solve_iterations = all[r].solveIterations;
      i.warm_fallback = all[r].warmFallback ? 1 : 0;
    }
  });
}

int sl3_operator_blocks(const sl3_operator* op, int capacity, int* count, int* first_node,
                        int* nodes, int* lmax, int* ntheta, int* nphi, int64_t* grid_values) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    const auto& c = op->op->GetCoupling();
    if (count) *count = c.BlockCount();
    if (grid_values) *grid_values = static_cast<int64_t>(c.GridValueCount());
    for (int i = 0; i < c.BlockCount() && i < capacity; ++i) {
      const NodeBlock& k = c.Block(i);
      if (first_node) first_node[i] = k.firstNode;
      if (nodes) nodes[i] = k.nodes;
      if (lmax) lmax[i] = k.lmax;
      if (ntheta) ntheta[i] = k.grid.NTheta();
      if (nphi) nphi[i] = k.grid.NPhi();
    }
  });
}

int sl3_operator_block_geometry(const sl3_operator* op, int block, double* theta,
                                double* wtheta) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    const auto& c = op->op->GetCoupling();
    Check(block >= 0 && block < c.BlockCount(), "no such block");
    const SurfaceGrid& g = c.Block(block).grid;
    if (theta) std::memcpy(theta, g.CoLatitudes().data(), sizeof(double) * g.NTheta());
    if (wtheta) std::memcpy(wtheta, g.CoLatitudeWeights().data(), sizeof(double) * g.NTheta());
  });
}

int sl3_operator_geometry(const sl3_operator* op, double* radii, double* theta,
                          double* wtheta) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    const auto& c = op->op->GetCoupling();
    if (radii) {
      const auto& r = c.NodeRadii();  // one realisation's nodes, repeated
      for (int node = 0; node < c.NodeCount(); ++node) radii[node] = r[node % c.RadialNodeCount()];
    }
    if (theta) {
      std::memcpy(theta, c.Grid().CoLatitudes().data(), sizeof(double) * c.Grid().NTheta());
    }
    if (wtheta) {
      std::memcpy(wtheta, c.Grid().CoLatitudeWeights().data(),
                  sizeof(double) * c.Grid().NTheta());
    }
  });
}

int sl3_state_zero(sl3_operator* op, sl3_state** out) {
  return Guard([&] {
    Check(op != nullptr && out != nullptr, "null argument");
    auto s = std::make_unique<sl3_state>();
    s->state = op->op->ZeroState();
    s->owner = op;
    s->id = g_nextStateId++;
    SL3_HIP(hipStreamSynchronize(Stream()));
    *out = s.release();
  });
}

int sl3_state_clone(const sl3_operator* op, const sl3_state* state, sl3_state** out) {
  return Guard([&] {
    Check(out != nullptr, "null argument");
    CheckOwner(op, state);
    auto s = std::make_unique<sl3_state>();
    s->state = op->op->CloneState(*state->state);
    s->owner = op;
    s->id = g_nextStateId++;
    *out = s.release();
  });
}

int sl3_state_destroy(sl3_state* state) {
  return Guard([&] {
    if (state) SL3_HIP(hipStreamSynchronize(Stream()));
    delete state;
  });
}

int sl3_state_time(const sl3_state* state, double* t) {
  return Guard([&] {
    Check(state != nullptr && t != nullptr, "null argument");
    *t = state->state->t;
  });
}

int sl3_step(sl3_operator* op, sl3_state* state, double dt, int scheme, sl3_load_fn load,
             sl3_load_fn tide, void* user, const double* start_dev, sl3_step_info* info) {
  return sl3_step_guarded(op, state, dt, scheme, load, tide, user, start_dev, 0.0, info);
}

int sl3_step_guarded(sl3_operator* op, sl3_state* state, double dt, int scheme,
                     sl3_load_fn load, sl3_load_fn tide, void* user, const double* start_dev,
                     double guard, sl3_step_info* info) {
  return Guard([&] {
    CheckOwner(op, state);
    auto wrap = [user](sl3_load_fn fn) {
      LoadFn out;
      if (fn) {
        out = [fn, user](int l, double t, double* columns) {
          const int r = fn(l, t, columns, user);
          if (r < 0) {
            throw ArgumentError("a load channel reported a failure at degree " +
                                std::to_string(l));
          }
          return r != 0;
        };
      }
      return out;
    };
    const Channels channels{wrap(load), wrap(tide)};
    const auto r = op->op->Step(*state->state, dt, scheme, channels, start_dev, guard);
    if (info) {
      info->iterations = r.iterations;
      info->residual = r.residual;
      info->converged = r.converged ? 1 : 0;
      info->mixed_passes = r.mixedPasses;
      info->start_ms = op->op->lastTiming.startStep;
      info->passes_ms = op->op->lastTiming.passes;
      info->finish_ms = op->op->lastTiming.finish;
      info->solve_iterations = r.solveIterations;
      info->warm_fallback = r.warmFallback ? 1 : 0;
    }
  });
}

namespace {

Channels WrapChannels(sl3_load_fn load, sl3_load_fn tide, void* user) {
  auto wrap = [user](sl3_load_fn fn) {
    LoadFn out;
    if (fn) {
      out = [fn, user](int l, double t, double* columns) {
        const int r = fn(l, t, columns, user);
        if (r < 0) {
          throw ArgumentError("a load channel reported a failure at degree " +
                              std::to_string(l));
        }
        return r != 0;
      };
    }
    return out;
  };
  return Channels{wrap(load), wrap(tide)};
}

void FillInfo(sl3_step_info* info, const StepInfo& r) {
  if (info == nullptr) return;
  info->iterations = r.iterations;
  info->residual = r.residual;
  info->converged = r.converged ? 1 : 0;
  info->mixed_passes = r.mixedPasses;
  info->start_ms = info->passes_ms = info->finish_ms = 0.0f;
  info->solve_iterations = r.solveIterations;
  info->warm_fallback = r.warmFallback ? 1 : 0;
}

}  // namespace

int sl3_step_start(sl3_operator* op, sl3_state* state, double dt, int scheme, sl3_load_fn load,
                   sl3_load_fn tide, void* user) {
  return Guard([&] {
    CheckOwner(op, state);
    op->op->OpenStart(*state->state, dt, scheme, WrapChannels(load, tide, user));
  });
}

int sl3_step_pass(sl3_operator* op, const double* s_dev, double* next_dev) {
  return Guard([&] {
    Check(op != nullptr && s_dev != nullptr && next_dev != nullptr, "null argument");
    Check(s_dev != next_dev, "a pass writes a tensor other than its input");
    op->op->OpenPass(s_dev, next_dev);
  });
}

int sl3_step_converge(sl3_operator* op, sl3_pass_fn pass, void* user, const double* start_dev,
                      sl3_step_info* info) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    Viscoelastic3D::PassFn fn;
    if (pass != nullptr) {
      fn = [pass, user](const double* s, double* next) {
        if (pass(s, next, user) != 0) throw ArgumentError("the caller's pass reported a failure");
      };
    }
    FillInfo(info, op->op->OpenConverge(fn, start_dev, 0.0));
  });
}

int sl3_step_close(sl3_operator* op, const double* start_dev, sl3_step_info* info) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    FillInfo(info, op->op->OpenClose(start_dev, 0.0));
  });
}

int sl3_step_finish(sl3_operator* op, const double* strain_dev) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    op->op->OpenFinish(strain_dev);
    SL3_HIP(hipStreamSynchronize(Stream()));
  });
}

int sl3_step_cancel(sl3_operator* op) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    op->op->OpenCancel();
  });
}

int sl3_step_add_memory(sl3_operator* op, const double* const* increment_dev) {
  return Guard([&] {
    Check(op != nullptr && increment_dev != nullptr, "null argument");
    for (int b = 0; b < 3; ++b) Check(increment_dev[b] != nullptr, "null increment block");
    op->op->OpenAddMemory(increment_dev);
  });
}

int sl3_step_add_force(sl3_operator* op, const double* analysed_dev) {
  return Guard([&] {
    Check(op != nullptr && analysed_dev != nullptr, "null argument");
    op->op->OpenAddForce(analysed_dev);
  });
}

int sl3_step_transposed_pass(sl3_operator* op, const double* z_dev,
                             const double* const* src_dev, double* out_dev) {
  return Guard([&] {
    Check(op != nullptr && z_dev != nullptr && out_dev != nullptr, "null argument");
    Check(z_dev != out_dev, "z and out must not overlap");
    op->op->OpenTransposedPass(z_dev, src_dev, out_dev);
    SL3_HIP(hipStreamSynchronize(Stream()));
  });
}

int sl3_step_view(sl3_operator* op, const char* name, const double** dev, int64_t* count) {
  return Guard([&] {
    Check(op != nullptr && name != nullptr && dev != nullptr, "null argument");
    auto& v = *op->op;
    const auto& c = v.GetCoupling();
    const std::string key(name);
    DeviceState* st = v.OpenState();
    if (key == "weight" || key == "beta") {
      *dev = key == "weight" ? v.Weight() : v.Beta();
      if (count) *count = static_cast<int64_t>(c.GridValueCount());
    } else if (key == "fixed" || key == "strain") {
      *dev = key == "fixed" ? v.FixedForce() : v.Iterate();
      if (count) *count = static_cast<int64_t>(c.SpectralDoubles());
    } else if (key.size() == 7 && key.rfind("memory", 0) == 0 && key[6] >= '0' && key[6] <= '2') {
      Check(st != nullptr, "no step is open on this operator");
      const int b = key[6] - '0';
      *dev = st->m[b].data();
      if (count) *count = static_cast<int64_t>(st->m[b].size());
    } else {
      throw ArgumentError("unknown step view: " + key);
    }
  });
}

int sl3_step_relative_change(sl3_operator* op, const double* next_dev, const double* now_dev,
                             double* out) {
  return Guard([&] {
    Check(op != nullptr && next_dev != nullptr && now_dev != nullptr && out != nullptr,
          "null argument");
    *out = op->op->RelativeChangeOf(next_dev, now_dev);
  });
}

int sl3_coupling_solution_doubles(const sl3_operator* op, int64_t* count) {
  return Guard([&] {
    Check(op != nullptr && count != nullptr, "null argument");
    *count = static_cast<int64_t>(op->op->SolutionDoubles());
  });
}

int sl3_coupling_force(sl3_operator* op, const double* weighted_dev, double* forces_dev) {
  return Guard([&] {
    Check(op != nullptr && weighted_dev != nullptr && forces_dev != nullptr, "null argument");
    op->op->ForceOf(weighted_dev, forces_dev);
  });
}

int sl3_coupling_strain(sl3_operator* op, const double* solutions_dev, double* spectral_dev) {
  return Guard([&] {
    Check(op != nullptr && solutions_dev != nullptr && spectral_dev != nullptr, "null argument");
    op->op->StrainOfSolutions(solutions_dev, spectral_dev);
  });
}

int sl3_coupling_multiply(sl3_operator* op, double* c0_dev, double* c1_dev, double* r_dev,
                          const double* weight_dev) {
  return Guard([&] {
    Check(op != nullptr && c0_dev && c1_dev && r_dev && weight_dev, "null argument");
    LaunchMultiply(c0_dev, c1_dev, r_dev, weight_dev, op->op->GetCoupling().GridValueCount(),
                   Stream());
  });
}

int sl3_coupling_scale_by_node(sl3_operator* op, const double* factor, double* spectral_dev) {
  return Guard([&] {
    Check(op != nullptr && factor != nullptr && spectral_dev != nullptr, "null argument");
    const auto& c = op->op->GetCoupling();
    DeviceArray<double> f;
    f.Upload(factor, c.NodeCount(), Stream());
    LaunchScaleByNode(f.data(), spectral_dev, c.SpectralDoubles() / 2, c.NodeCount(), Stream());
    SL3_HIP(hipStreamSynchronize(Stream()));
  });
}

int sl3_state_export(const sl3_operator* op, const sl3_state* state, const char* name,
                     double* dst, int64_t* count) {
  return Guard([&] {
    Check(name != nullptr, "null argument");
    CheckOwner(op, state);
    const auto& c = op->op->GetCoupling();
    const auto& st = *state->state;
    const std::string key(name);
    hipStream_t s = Stream();
    if (key == "coefficients" || key == "coefficients_device") {
      Check(st.hasCoefficients, "the state carries no strain coefficients");
      const size_t n = c.SpectralDoubles();
      if (key == "coefficients_device") return Export(st.coefficients.data(), n, dst, count);
      DeviceArray<double> ref(n);
      ToReference(c, st.coefficients.data(), ref.data(), s);
      return Export(ref.data(), n, dst, count);
    }
    for (int b = 0; b < 3; ++b) {
      if (key == "m" + std::to_string(b)) return Export(st.m[b].data(), st.m[b].size(), dst, count);
      if (key == "strain" + std::to_string(b)) {
        Check(st.hasStrain, "the state carries no strain grid");
        return Export(st.strain[b].data(), st.strain[b].size(), dst, count);
      }
    }
    throw ArgumentError("unknown state export: " + key);
  });
}

int sl3_tuning_set(int analysis_rows, int synthesis_theta, int band_panel_rows) {
  return Guard([&] { SetTuningOverride(analysis_rows, synthesis_theta, band_panel_rows); });
}

int sl3_operator_export(sl3_operator* op, const char* name, double* dst, int64_t* count) {
  return Guard([&] {
    Check(op != nullptr && name != nullptr, "null argument");
    auto& v = *op->op;
    const auto& c = v.GetCoupling();
    const std::string key(name);
    const size_t G = c.GridValueCount();
    if (key == "weight") return Export(v.Weight(), G, dst, count);
    if (key == "beta") return Export(v.Beta(), G, dst, count);
    if (key == "itau") return Export(v.Itau(), G, dst, count);
    if (key == "fixed") return Export(v.FixedForce(), c.SpectralDoubles(), dst, count);
    if (key == "analysed") return Export(v.Analysed(), c.SpectralDoubles(), dst, count);
    if (key == "residuals") return ExportHost(v.residualHistory, dst, count);
    if (key == "solve_energy") return ExportHost(v.solveEnergy, dst, count);
    if (key == "mu_node") return ExportHost(v.MuNode(), dst, count);
    if (key == "itau_ref_node") return ExportHost(v.ItauRefNode(), dst, count);
    if (key == "tuning") return ExportHost(c.Tuning(), dst, count);
    if (key == "tasks") {
      std::vector<double> t;
      for (const auto& task : v.Tasks()) {
        t.push_back(task.l);
        t.push_back(task.kind);
        t.push_back(task.ndof);
        t.push_back(task.firstNode);
        t.push_back(task.nm);
      }
      return ExportHost(t, dst, count);
    }
    if (key.rfind("solution:", 0) == 0) {
      const int task = std::stoi(key.substr(9));
      Check(task >= 0 && task < static_cast<int>(v.Tasks().size()), "no such task");
      return ExportHost(v.SolutionOfTask(task), dst, count);
    }
    throw ArgumentError("unknown operator export: " + key);
  });
}

int sl3_spectral_to_reference(const sl3_operator* op, const double* device_dev,
                              double* reference_dev) {
  return Guard([&] {
    Check(op && device_dev && reference_dev, "null argument");
    ToReference(op->op->GetCoupling(), device_dev, reference_dev, Stream());
    SL3_HIP(hipStreamSynchronize(Stream()));
  });
}

int sl3_state_spectral(const sl3_state* state, const double** spectral_dev) {
  return Guard([&] {
    Check(state != nullptr && spectral_dev != nullptr, "null argument");
    const auto& st = *state->state;
    *spectral_dev = st.hasCoefficients ? st.coefficients.data() : nullptr;
  });
}

namespace {

// The rows of degree <= cut of the device layout, blocks spin -2, -1, 0 in
// memory order (each row: the nodes' complex values), as double offsets.
const std::pair<int, DeviceArray<long long>>& CutRows(sl3_operator* op, int cut) {
  const auto& c = op->op->GetCoupling();
  Check(cut >= 0 && cut <= c.LMax(), "degree cut must be in 0..lmax");
  std::lock_guard<std::mutex> lock(op->rowsMutex);
  auto it = op->cutRows.find(cut);
  if (it != op->cutRows.end()) return it->second;
  std::vector<long long> rows;
  const long long nodes = c.NodeCount();
  for (int b = 0; b < 3; ++b) {
    const auto& lay = c.Grid().Layout(b);
    for (int m = lay.mlo; m <= lay.lmax; ++m)
      for (int l = lay.LMin(m); l <= std::min(cut, lay.lmax); ++l)
        rows.push_back(static_cast<long long>(c.SpectralOffset(b)) +
                       2ll * nodes * lay.Row(l, m));
  }
  auto& entry = op->cutRows[cut];
  entry.first = static_cast<int>(rows.size());
  entry.second.Upload(rows.data(), rows.size(), Stream());
  SL3_HIP(hipStreamSynchronize(Stream()));
  return entry;
}

}  // namespace

int sl3_spectral_rows(sl3_operator* op, int cut, int64_t* count) {
  return Guard([&] {
    Check(op != nullptr && count != nullptr, "null argument");
    *count = CutRows(op, cut).first;
  });
}

int sl3_spectral_gather(sl3_operator* op, const double* spectral_dev, const double* base_dev,
                        int cut, const float* mean_dev, const float* scale_dev, float* out_dev,
                        int ld, int col0) {
  return Guard([&] {
    Check(op && spectral_dev && out_dev, "null argument");
    Check((mean_dev == nullptr) == (scale_dev == nullptr), "mean and scale go together");
    const auto& rows = CutRows(op, cut);
    Check(col0 >= 0 && ld >= col0 + 2 * rows.first, "ld must hold col0 + 2 rows");
    LaunchSpectralGather(spectral_dev, base_dev, rows.second.data(), rows.first,
                         op->op->GetCoupling().NodeCount(), mean_dev, scale_dev, out_dev, ld,
                         col0, Stream());
    SL3_HIP(hipStreamSynchronize(Stream()));
  });
}

int sl3_spectral_scatter(sl3_operator* op, const float* in_dev, int ld, int cut,
                         const float* mean_dev, const float* scale_dev, const double* base_dev,
                         double* out_dev) {
  return Guard([&] {
    Check(op && in_dev && out_dev, "null argument");
    Check((mean_dev == nullptr) == (scale_dev == nullptr), "mean and scale go together");
    const auto& rows = CutRows(op, cut);
    Check(ld >= 2 * rows.first, "ld must hold 2 rows");
    const auto& c = op->op->GetCoupling();
    hipStream_t s = Stream();
    if (base_dev != nullptr) {
      if (base_dev != out_dev) LaunchCopy(out_dev, base_dev, c.SpectralDoubles(), s);
    } else {
      LaunchFill(out_dev, c.SpectralDoubles(), 0.0, s);
    }
    LaunchSpectralScatterAdd(in_dev, ld, rows.second.data(), rows.first, c.NodeCount(),
                             mean_dev, scale_dev, out_dev, s);
    SL3_HIP(hipStreamSynchronize(s));
  });
}

int sl3_extrapolator_create(sl3_operator* op, int order, sl3_extrapolator** out) {
  return Guard([&] {
    Check(op != nullptr && out != nullptr, "null argument");
    Check(order >= 1 && order <= 5, "extrapolation order must be 1 to 5");
    *out = nullptr;
    auto ex = std::make_unique<sl3_extrapolator>();
    ex->op = op;
    ex->order = order;
    const size_t S = op->op->GetCoupling().SpectralDoubles();
    ex->buffers.resize(order + 1);
    for (auto& b : ex->buffers) b.Resize(S);
    ex->start.Resize(S);
    *out = ex.release();
  });
}

int sl3_extrapolator_destroy(sl3_extrapolator* ex) {
  return Guard([&] { delete ex; });
}

int sl3_extrapolator_reset(sl3_extrapolator* ex) {
  return Guard([&] {
    Check(ex != nullptr, "null argument");
    ex->recorded.clear();
    ex->last = 0;
  });
}

int sl3_extrapolated_step(sl3_extrapolator* ex, sl3_state* state, double dt, int scheme,
                          sl3_load_fn load, sl3_load_fn tide, void* user, double guard,
                          sl3_step_info* info) {
  const double* start = nullptr;
  int status = Guard([&] {
    Check(ex != nullptr && state != nullptr, "null argument");
    CheckOwner(ex->op, state);
    Check(guard >= 0.0, "the warm-start guard must be non-negative");
    const double t0 = state->state->t;
    // another state than the last one stepped, or the same one moved to
    // another time since, starts a new history
    if (!ex->recorded.empty() && (ex->last != state->id || ex->recorded.back().first != t0))
      ex->recorded.clear();
    if (ex->recorded.empty() && scheme != SL3_SDIRK2 && dt > 0.0) {
      // a new history starts at the strain the state's step starts from, at
      // the state's time: the second step is then extrapolated already
      ex->op->op->StartStrain(*state->state, WrapChannels(load, tide, user),
                              ex->buffers[0].data());
      ex->recorded.emplace_back(t0, 0);
      ex->last = state->id;
    }
    if (scheme == SL3_SDIRK2 || !(dt > 0.0) || ex->recorded.size() < 2) return;
    // Lagrange extrapolation through the recorded times to t0 + dt
    const double target = t0 + dt;
    SpectralTerms t;
    t.count = static_cast<int>(ex->recorded.size());
    for (int j = 0; j < t.count; ++j) {
      double w = 1.0;
      for (int k = 0; k < t.count; ++k)
        if (k != j)
          w *= (target - ex->recorded[k].first) / (ex->recorded[j].first - ex->recorded[k].first);
      t.c[j] = w;
      t.src[j] = ex->buffers[ex->recorded[j].second].data();
    }
    LaunchCombineTerms(t, ex->start.data(), ex->start.size(), Stream());
    start = ex->start.data();
  });
  if (status != SL3_OK) return status;
  status = sl3_step_guarded(ex->op, state, dt, scheme, load, tide, user, start,
                            start != nullptr ? guard : 0.0, info);
  if (status != SL3_OK) return status;
  return Guard([&] {
    const auto& st = *state->state;
    if (!st.hasCoefficients) return;
    int slot = -1;
    if (!ex->recorded.empty() && ex->recorded.back().first == st.t) {
      // a zero-length step ends where the last one did: its strain replaces
      // that entry (two entries at one time would make the weights divide
      // by zero)
      slot = ex->recorded.back().second;
      ex->recorded.pop_back();
    } else if (ex->recorded.size() == ex->buffers.size()) {
      slot = ex->recorded.front().second;
      ex->recorded.erase(ex->recorded.begin());
    } else {
      for (int b = 0; b < static_cast<int>(ex->buffers.size()) && slot < 0; ++b) {
        bool used = false;
        for (const auto& r : ex->recorded) used = used || r.second == b;
        if (!used) slot = b;
      }
    }
    LaunchCopy(ex->buffers[slot].data(), st.coefficients.data(), ex->buffers[slot].size(),
               Stream());
    SL3_HIP(hipStreamSynchronize(Stream()));
    ex->recorded.emplace_back(st.t, slot);
    ex->last = state->id;
  });
}

int sl3_spectral_combine(const sl3_operator* op, int count, const double* coefficients,
                         const double* const* sources_dev, double* dst_dev) {
  return Guard([&] {
    Check(op && coefficients && sources_dev && dst_dev, "null argument");
    Check(count >= 1 && count <= kSpectralTermsMax, "combine: 1 to 8 terms");
    SpectralTerms t;
    t.count = count;
    for (int j = 0; j < count; ++j) {
      Check(sources_dev[j] != nullptr, "null source");
      t.c[j] = coefficients[j];
      t.src[j] = sources_dev[j];
    }
    LaunchCombineTerms(t, dst_dev, op->op->GetCoupling().SpectralDoubles(), Stream());
    SL3_HIP(hipStreamSynchronize(Stream()));
  });
}

int sl3_spectral_from_reference(const sl3_operator* op, const double* reference_dev,
                                double* device_dev) {
  return Guard([&] {
    Check(op && device_dev && reference_dev, "null argument");
    FromReference(op->op->GetCoupling(), reference_dev, device_dev, Stream());
    SL3_HIP(hipStreamSynchronize(Stream()));
  });
}

int sl3_gsh_synthesise(const sl3_operator* op, int spin, const double* coef_dev,
                       double* field_dev) {
  return Guard([&] {
    Check(op && coef_dev && field_dev, "null argument");
    const int b = BlockOfSpin(spin);
    const auto& c = op->op->GetCoupling();
    hipStream_t s = Stream();
    const size_t blockDoubles = c.SpectralOffset(b + 1) - c.SpectralOffset(b);
    DeviceArray<double> dev(blockDoubles), fourier(c.FourierDoubles(b));
    LaunchToDeviceLayout(coef_dev, dev.data(), c.RowOffsetDevice(b), spin, c.LMax(),
                         c.NodeCount(), s);
    c.SynthesiseBlockToFourier(b, dev.data(), fourier.data(), s);
    RingLaunch r;
    r.block = b;
    r.op = kRingSynthesise;
    r.fourierIn = fourier.data();
    r.gridOut = field_dev;
    c.Ring(r, s);
    SL3_HIP(hipStreamSynchronize(s));
  });
}

int sl3_gsh_analyse(const sl3_operator* op, int spin, const double* field_dev,
                    double* coef_dev) {
  return Guard([&] {
    Check(op && coef_dev && field_dev, "null argument");
    const int b = BlockOfSpin(spin);
    const auto& c = op->op->GetCoupling();
    hipStream_t s = Stream();
    const size_t blockDoubles = c.SpectralOffset(b + 1) - c.SpectralOffset(b);
    DeviceArray<double> dev(blockDoubles), fourier(c.FourierDoubles(b));
    RingLaunch r;
    r.block = b;
    r.op = kRingAnalyse;
    r.gridIn = field_dev;
    r.fourierOut = fourier.data();
    c.Ring(r, s);
    c.AnalyseBlockFromFourier(b, fourier.data(), dev.data(), s);
    LaunchToReferenceLayout(dev.data(), coef_dev, c.RowOffsetDevice(b), spin, c.LMax(),
                            c.NodeCount(), s);
    SL3_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
