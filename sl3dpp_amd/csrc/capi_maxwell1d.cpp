// C ABI of the batched 1D Maxwell evolution (include/sl3dpp_amd.h):
// ViscoelasticOperator's steps for every spheroidal degree on the GPU.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "capi_handles.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "maxwell1d.hpp"

using namespace sl3amd;

struct sl3_maxwell1d {
  std::unique_ptr<Maxwell1D> impl;
};

namespace {

void Check(bool ok, const std::string& what) {
  if (!ok) throw ArgumentError(what);
}

}  // namespace

extern "C" {

int sl3_maxwell1d_create(const sl3_earth* earth, int lmax, double eps, sl3_maxwell1d** out) {
  return Guard([&] {
    Check(earth != nullptr && out != nullptr, "null argument");
    auto h = std::make_unique<sl3_maxwell1d>();
    h->impl = std::make_unique<Maxwell1D>(earth->earth, lmax, eps);
    *out = h.release();
  });
}

int sl3_maxwell1d_destroy(sl3_maxwell1d* h) {
  return Guard([&] { delete h; });
}

int sl3_maxwell1d_reset(sl3_maxwell1d* h) {
  return Guard([&] {
    Check(h != nullptr, "null handle");
    h->impl->Reset();
  });
}

int sl3_maxwell1d_set_load(sl3_maxwell1d* h, const double* pattern, int64_t pattern_count,
                           const double* ramp, int64_t ramp_count) {
  return Guard([&] {
    Check(h != nullptr && pattern != nullptr && ramp != nullptr, "null argument");
    Check(pattern_count >= 0 && ramp_count > 0, "counts");
    h->impl->SetLoad(std::vector<double>(pattern, pattern + pattern_count),
                     std::vector<double>(ramp, ramp + ramp_count));
  });
}

int sl3_maxwell1d_evolve(sl3_maxwell1d* h, double dt, int scheme, int steps, int use_graph) {
  return Guard([&] {
    Check(h != nullptr, "null handle");
    Check(scheme == SL3_EXPONENTIAL_TRAPEZOID || scheme == SL3_BACKWARD_EULER,
          "scheme: the exponential trapezoid or backward Euler");
    h->impl->Evolve(dt, scheme == SL3_BACKWARD_EULER ? kBackwardEuler : kExponentialTrapezoid,
                    steps, use_graph != 0);
  });
}

int sl3_maxwell1d_shape(const sl3_maxwell1d* h, int64_t* memory_doubles,
                        int64_t* solution_doubles, int* nodes, int* steps_taken) {
  return Guard([&] {
    Check(h != nullptr, "null handle");
    if (memory_doubles != nullptr) *memory_doubles = static_cast<int64_t>(h->impl->MemoryDoubles());
    if (solution_doubles != nullptr) {
      *solution_doubles = static_cast<int64_t>(h->impl->SolutionDoubles());
    }
    if (nodes != nullptr) *nodes = h->impl->Nodes();
    if (steps_taken != nullptr) *steps_taken = h->impl->StepIndex();
  });
}

int sl3_maxwell1d_degree(const sl3_maxwell1d* h, int l, int* ndof, int* first_node,
                         int64_t* memory_offset, int64_t* solution_offset) {
  return Guard([&] {
    Check(h != nullptr, "null handle");
    Check(l >= 0 && l <= h->impl->LMax(), "degree out of range");
    const Degree1D& d = h->impl->Degrees()[l];
    if (ndof != nullptr) *ndof = d.ndof;
    if (first_node != nullptr) *first_node = d.firstNode;
    if (memory_offset != nullptr) *memory_offset = d.memOff;
    if (solution_offset != nullptr) *solution_offset = d.rhsOff;
  });
}

int sl3_maxwell1d_export(const sl3_maxwell1d* h, const char* name, double* dst, int64_t count) {
  return Guard([&] {
    Check(h != nullptr && name != nullptr && dst != nullptr, "null argument");
    std::vector<double> v;
    if (std::strcmp(name, "memory") == 0) {
      v = h->impl->Memory();
    } else if (std::strcmp(name, "strain") == 0) {
      v = h->impl->Strain();
    } else if (std::strcmp(name, "solutions") == 0) {
      v = h->impl->Solutions();
    } else if (std::strcmp(name, "tuning") == 0) {
      v = h->impl->Tuning();
    } else {
      throw ArgumentError(std::string("unknown array '") + name + "'");
    }
    Check(count == static_cast<int64_t>(v.size()),
          "count must be " + std::to_string(v.size()) + " doubles");
    std::memcpy(dst, v.data(), sizeof(double) * v.size());
  });
}

int sl3_maxwell1d_surface(const sl3_maxwell1d* h, int component, double* out_dev) {
  return Guard([&] {
    Check(h != nullptr && out_dev != nullptr, "null argument");
    hipStream_t s = Stream();
    h->impl->Surface(component, out_dev, s);
    SL3_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
