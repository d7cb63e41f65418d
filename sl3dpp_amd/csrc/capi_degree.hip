// C ABI of the Earth from nodal arrays and of one degree's radial system
// (include/sl3dpp_amd.h): DegreeOperator / ToroidalOperator assembled on
// the host, factorised and solved on the GPU by the band kernels.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "capi_handles.hpp"
#include "common.hpp"
#include "degree_systems.hpp"
#include "kernels.hpp"

using namespace sl3amd;

namespace {

void Check(bool ok, const std::string& what) {
  if (!ok) throw ArgumentError(what);
}

// x(i, k) *= scale(i) over ndof x nrhs column-major values; the pinned rows
// (the toroidal degree-1 system's, one per solid region: at most two, -1
// for none) set to zero (DegreeOperator::BeginSolve / FinishSolve,
// src/DegreeOperator.cpp:861-901).
__global__ void ScaleRowsKernel(double* __restrict__ x, const double* __restrict__ scale, int n,
                                long long count, int pinned0, int pinned1) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= count) return;
  const int row = static_cast<int>(i % n);
  x[i] = row == pinned0 || row == pinned1 ? 0.0 : x[i] * scale[row];
}

void ScaleRows(double* x, const double* scale, int n, int nrhs, const int pinned[2],
               hipStream_t s) {
  const long long count = static_cast<long long>(n) * nrhs;
  if (count == 0) return;
  hipLaunchKernelGGL(ScaleRowsKernel, dim3(CeilDiv(count, 256)), dim3(256), 0, s, x, scale, n,
                     count, pinned[0], pinned[1]);
  SL3_HIP(hipGetLastError());
}

}  // namespace

struct sl3_degree_operator {
  std::shared_ptr<const Earth> earth;
  int kind = 0;
  std::vector<double> shear;  // non-dimensional, (element, point)
  MemoryPoints points;
  PointShear pointShear;
  std::unique_ptr<DegreeOperator> sph;
  std::unique_ptr<ToroidalOperator> tor;
  // the device factorisation, made on the first solve
  bool factored = false;
  bool cholesky = false;
  int kd = 0, ndof = 0, pinned = -1;
  int pins[2] = {-1, -1};  // every pinned dof
  DeviceArray<double> factor, scale;
  DeviceArray<int> piv;
};

extern "C" {

int sl3_earth_create_nodal(const sl3_earth_nodal* spec, sl3_earth** out) {
  return Guard([&] {
    Check(spec != nullptr && out != nullptr, "null argument");
    Check(spec->elements >= 1 && spec->ngll >= 2, "elements >= 1, ngll >= 2");
    Check(spec->branches == 0 || spec->branches == 1, "0 or 1 Maxwell branch");
    const size_t E = static_cast<size_t>(spec->elements);
    const size_t size = E * spec->ngll;
    auto take = [&](const double* p, size_t n, const char* what) {
      Check(p != nullptr, std::string("missing ") + what);
      return std::vector<double>(p, p + n);
    };
    NodalEarthSpec ns;
    ns.ngll = spec->ngll;
    ns.edges = take(spec->edges, E + 1, "edges");
    ns.rho = take(spec->rho, size, "rho");
    ns.drho = take(spec->drho, size, "drho");
    ns.g = take(spec->gravity, size, "gravity");
    ns.A = take(spec->A, size, "A");
    ns.C = take(spec->C, size, "C");
    ns.F = take(spec->F, size, "F");
    ns.L = take(spec->L, size, "L");
    ns.N = take(spec->N, size, "N");
    Check(spec->fluid != nullptr, "missing fluid");
    ns.fluid.assign(spec->fluid, spec->fluid + E);
    if (spec->branches == 1) {
      ns.branchMu = take(spec->branch_mu, size, "branch_mu");
      ns.branchItau = take(spec->branch_itau, size, "branch_itau");
    }
    if (spec->length > 0.0) ns.length = spec->length;
    if (spec->density > 0.0) ns.density = spec->density;
    auto e = std::make_unique<sl3_earth>();
    e->earth = BuildEarth(ns);
    *out = e.release();
  });
}

void sl3_degree_options_default(sl3_degree_options* o) {
  if (o == nullptr) return;
  o->eps = 1e-8;
  o->gravitating = 1;
  o->material = SL3_MATERIAL_ELASTIC;
  o->point_shear = nullptr;
  o->point_set = SL3_POINTS_GAUSS;
}

int sl3_degree_operator_create(const sl3_earth* earth, int l, int kind,
                               const sl3_degree_options* options, sl3_degree_operator** out) {
  return Guard([&] {
    Check(earth != nullptr && out != nullptr, "null argument");
    Check(kind == SL3_SPHEROIDAL || kind == SL3_TOROIDAL, "kind: spheroidal or toroidal");
    sl3_degree_options o;
    sl3_degree_options_default(&o);
    if (options != nullptr) o = *options;
    Check(o.eps > 0.0 && o.eps < 1.0, "eps in (0, 1)");
    Check(o.material == SL3_MATERIAL_ELASTIC || o.material == SL3_MATERIAL_LONG_TERM,
          "material: elastic or long-term");
    Check(o.point_set == SL3_POINTS_GAUSS || o.point_set == SL3_POINTS_GLL,
          "point set: Gauss or GLL");
    Check(kind == SL3_SPHEROIDAL || o.gravitating, "the toroidal system does not gravitate");
    auto d = std::make_unique<sl3_degree_operator>();
    d->earth = earth->earth;
    d->kind = kind;
    const Earth& e = *d->earth;
    const RadialMesh& mesh = *e.mesh;
    const Material& material = o.material == SL3_MATERIAL_ELASTIC ? e.material : e.reference.nodal;
    const PointShear* shear = nullptr;
    if (o.point_shear != nullptr) {
      d->points = o.point_set == SL3_POINTS_GLL ? MemoryPoints::Gll(mesh) : MemoryPoints::Gauss(mesh);
      const size_t n = static_cast<size_t>(mesh.ElementCount()) * d->points.count;
      d->shear.assign(o.point_shear, o.point_shear + n);
      for (auto& v : d->shear) v /= e.scales.Modulus();
      d->pointShear = PointShear{&d->points, &d->shear};
      shear = &d->pointShear;
    }
    if (kind == SL3_SPHEROIDAL) {
      d->sph = std::make_unique<DegreeOperator>(material, l,
                                                DegreeOptions{o.eps, o.gravitating != 0, shear});
      d->ndof = d->sph->DofCount();
      d->kd = d->sph->Kd();
    } else {
      d->tor = std::make_unique<ToroidalOperator>(material, l, o.eps, shear);
      d->ndof = d->tor->DofCount();
      d->kd = d->tor->Kd();
      const auto& pins = d->tor->Pinned();
      Check(pins.size() <= 2, "at most two solid regions (inner core, mantle) are supported");
      d->pinned = pins.empty() ? -1 : pins[0];
      for (size_t k = 0; k < pins.size(); ++k) d->pins[k] = pins[k];
    }
    *out = d.release();
  });
}

int sl3_degree_operator_destroy(sl3_degree_operator* op) {
  return Guard([&] {
    if (op) SL3_HIP(hipStreamSynchronize(Stream()));
    delete op;
  });
}

int sl3_degree_operator_shape(const sl3_degree_operator* op, int* ndof, int* kd,
                              int* first_element, int* pinned) {
  return Guard([&] {
    Check(op != nullptr, "null operator");
    if (ndof) *ndof = op->ndof;
    if (kd) *kd = op->kd;
    if (first_element) *first_element = op->sph ? op->sph->FirstElement() : op->tor->FirstElement();
    if (pinned) *pinned = op->pinned;
  });
}

int sl3_degree_operator_export(const sl3_degree_operator* op, const char* name, double* dst,
                               int64_t* count) {
  return Guard([&] {
    Check(op != nullptr && name != nullptr, "null argument");
    const std::string key(name);
    const RadialMesh& mesh = *op->earth->mesh;
    std::vector<double> v;
    if (key == "band") {
      v = op->sph ? op->sph->Band() : op->tor->Band();
    } else if (key == "scale") {
      v = op->sph ? op->sph->Scale() : op->tor->Scale();
    } else if (key == "dofs") {
      for (int node = 0; node < mesh.NodeCount(); ++node) {
        if (op->sph) {
          for (int c = 0; c < 3; ++c) v.push_back(op->sph->Dof(node, Component(c)));
        } else {
          v.push_back(op->tor->Dof(node));
        }
      }
    } else if (op->sph && key == "load") {
      v = op->sph->Load();
    } else if (op->sph && key == "load_force") {
      v = op->sph->LoadForce();
    } else if (op->sph && key == "load_potential") {
      v = op->sph->LoadPotential();
    } else if (op->sph && key == "tangential") {
      v = op->sph->Tangential();
    } else if (op->sph && key == "traction") {
      v = op->sph->Traction();
    } else if (op->sph && key == "traction_tangential") {
      v = op->sph->TractionTangential();
    } else if (op->sph && key.rfind("tide:", 0) == 0) {
      v = op->sph->Tide(std::stoi(key.substr(5)));
    } else {
      throw ArgumentError("unknown degree operator export: " + key);
    }
    if (count) *count = static_cast<int64_t>(v.size());
    if (dst) std::memcpy(dst, v.data(), sizeof(double) * v.size());
  });
}

int sl3_degree_operator_solve(sl3_degree_operator* op, double* b_dev, int nrhs,
                              int* used_cholesky) {
  return Guard([&] {
    Check(op != nullptr && (b_dev != nullptr || nrhs == 0), "null argument");
    Check(nrhs >= 0, "nrhs >= 0");
    hipStream_t s = Stream();
    if (!op->factored) {
      // DegreeOperator::Factorise (src/DegreeOperator.cpp:775-852): the
      // scaled band by Cholesky; a spheroidal system that is not positive
      // definite by LU with partial pivoting on the full band.
      const int n = op->ndof, kd = op->kd;
      const auto& scale = op->sph ? op->sph->Scale() : op->tor->Scale();
      op->scale.Upload(scale.data(), scale.size(), s);
      bool tryCholesky = op->tor != nullptr || op->sph->PositiveDiagonal();
      int info = 0;
      if (tryCholesky) {
        const auto band = op->sph ? op->sph->ScaledUpperBand() : op->tor->ScaledUpperBand();
        op->factor.Upload(band.data(), band.size(), s);
        const int status = sl3_band_cholesky_factor(op->factor.data(), n, kd, &info);
        if (status != SL3_OK) throw HipError(sl3_last_error());
        op->cholesky = info == 0;
        if (!op->cholesky && op->tor) {
          throw NumericalError("toroidal degree " + std::to_string(op->tor->Degree()) +
                               ": Cholesky failed (info = " + std::to_string(info) + ")");
        }
      }
      if (!op->cholesky) {
        const auto band = op->sph->ScaledLuBand();
        op->factor.Upload(band.data(), band.size(), s);
        op->piv.Resize(n);
        const int status = sl3_band_lu_factor(op->factor.data(), n, kd, kd, op->piv.data(), &info);
        if (status != SL3_OK) throw HipError(sl3_last_error());
        if (info != 0) {
          throw NumericalError("degree " + std::to_string(op->sph->Degree()) +
                               ": the factorisation failed (zero pivot " + std::to_string(info) +
                               ")");
        }
      }
      op->factored = true;
    }
    if (used_cholesky) *used_cholesky = op->cholesky ? 1 : 0;
    if (nrhs == 0) return;
    ScaleRows(b_dev, op->scale.data(), op->ndof, nrhs, op->pins, s);
    const int status = op->cholesky
                           ? sl3_band_cholesky_solve(op->factor.data(), op->ndof, op->kd, b_dev, nrhs)
                           : sl3_band_lu_solve(op->factor.data(), op->ndof, op->kd, op->kd,
                                               op->piv.data(), b_dev, nrhs);
    if (status != SL3_OK) throw HipError(sl3_last_error());
    const int none[2] = {-1, -1};
    ScaleRows(b_dev, op->scale.data(), op->ndof, nrhs, none, s);
    SL3_HIP(hipStreamSynchronize(s));
  });
}

int sl3_degree_operator_apply(const sl3_degree_operator* op, const double* x, int nrhs,
                              double* y) {
  return Guard([&] {
    Check(op != nullptr && x != nullptr && y != nullptr, "null argument");
    Check(nrhs >= 0, "nrhs >= 0");
    const std::vector<double> in(x, x + static_cast<size_t>(op->ndof) * nrhs);
    const auto out = op->sph ? op->sph->Apply(in, nrhs) : op->tor->Apply(in, nrhs);
    std::memcpy(y, out.data(), sizeof(double) * out.size());
  });
}

}  // extern "C"
