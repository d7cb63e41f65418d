// The C++ face of the library (include/sl3dpp_amd.hpp) used as a caller of
// the reference's Viscoelastic3D would use it, after the reference's own
// tests (tests/TestViscoelastic3D.cpp):
//   cpu: an Earth from a layered spec; a bad spec throws
//        std::invalid_argument with the library's message;
//   gpu: Step returns a new state and leaves its argument unchanged, and two
//        steps by Step equal two by StepInPlace bit for bit; the solutions
//        hold every degree (spheroidal 0..L, toroidal 1..L, ndof x (l + 1));
//        a negative dt and a tide with a degree-1 column throw
//        std::invalid_argument; an exception thrown by a load channel comes
//        out of Step as itself; two realisations side by side report their
//        own infos, the step's the larger pass count.
// Prints one line per check and "ok" at the end; exits non-zero on failure.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <memory>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sl3dpp_amd.hpp"

namespace {

using sl3dpp_amd::Complex;
using sl3dpp_amd::Earth;
using sl3dpp_amd::Scheme;
using sl3dpp_amd::Viscoelastic3D;

int failures = 0;

void Expect(bool ok, const char* what) {
  std::printf("[cpp] %s: %s\n", what, ok ? "pass" : "FAIL");
  std::fflush(stdout);
  if (!ok) ++failures;
}

// Core 0-3480 km (fluid), mantle to 6271 km, lithosphere to 6371 km; the
// viscous mantle 1e21 Pa s above 5701 km and 1e22 below.
sl3_earth_spec SmallEarth() {
  sl3_earth_spec s{};
  s.layers = 3;
  const double r[4] = {0.0, 3480e3, 6271e3, 6371e3};
  const double rho[3][2] = {{12000.0, 9900.0}, {5500.0, 3400.0}, {3300.0, 3300.0}};
  const double kappa[3][2] = {{1300e9, 650e9}, {650e9, 130e9}, {130e9, 130e9}};
  const double mu[3][2] = {{0.0, 0.0}, {290e9, 65e9}, {65e9, 65e9}};
  const int elements[3] = {6, 5, 1};
  for (int i = 0; i < 3; ++i) {
    s.r0[i] = r[i];
    s.r1[i] = r[i + 1];
    s.rho0[i] = rho[i][0];
    s.rho1[i] = rho[i][1];
    s.kappa0[i] = kappa[i][0];
    s.kappa1[i] = kappa[i][1];
    s.mu0[i] = mu[i][0];
    s.mu1[i] = mu[i][1];
    s.elements[i] = elements[i];
  }
  s.ngll = 6;
  s.viscous_layer = 1;
  s.viscosity_boundary = 5701e3;
  s.viscosity_upper = 1e21;
  s.viscosity_lower = 1e22;
  return s;
}

template <class E, class F>
bool Throws(F&& f, std::string* message = nullptr) {
  try {
    f();
  } catch (const E& e) {
    if (message) *message = e.what();
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

int Cpu() {
  Earth earth(SmallEarth());
  Expect(earth.Handle() != nullptr, "earth from a layered spec");
  sl3_earth_spec bad = SmallEarth();
  bad.layers = 0;
  std::string message;
  Expect(Throws<std::invalid_argument>([&] { Earth e(bad); }, &message) && !message.empty(),
         "a spec without layers throws std::invalid_argument");
  return failures;
}

struct LoadFailure : std::exception {
  const char* what() const noexcept override { return "the load channel failed"; }
};

int Gpu() {
  Earth earth(SmallEarth());
  const int lmax = 8;
  Viscoelastic3D::Options options;
  options.max_iterations = 60;
  options.rtol = 1e-12;
  Viscoelastic3D op(earth, lmax, nullptr, options);
  // delta = 0.6 sin(theta) cos(phi) on every memory node's grid
  int nodes = 0, ntheta = 0, nphi = 0;
  int64_t spectral = 0;
  sl3dpp_amd::Check(sl3_operator_shape(op.Handle(), &nodes, &ntheta, &nphi, &spectral));
  std::vector<double> radii(nodes), theta(ntheta), wtheta(ntheta);
  sl3dpp_amd::Check(sl3_operator_geometry(op.Handle(), radii.data(), theta.data(),
                                          wtheta.data()));
  std::vector<double> delta(static_cast<size_t>(nodes) * ntheta * nphi);
  for (int n = 0; n < nodes; ++n)
    for (int i = 0; i < ntheta; ++i)
      for (int j = 0; j < nphi; ++j)
        delta[(static_cast<size_t>(n) * ntheta + i) * nphi + j] =
            0.6 * std::sin(theta[i]) * std::cos(2.0 * M_PI * j / nphi);
  double* delta_dev = nullptr;
  if (hipMalloc(&delta_dev, sizeof(double) * delta.size()) != hipSuccess) return 1;
  if (hipMemcpy(delta_dev, delta.data(), sizeof(double) * delta.size(),
                hipMemcpyHostToDevice) != hipSuccess)
    return 1;
  sl3dpp_amd::Check(sl3_operator_set_delta(op.Handle(), delta_dev));
  const std::vector<double> itau = op.Export("itau");
  double itauMax = 0.0;
  for (double v : itau) itauMax = std::max(itauMax, v);
  const double dt = 0.5 / itauMax;
  const Viscoelastic3D::LoadCallable3D load = [](int l, double) {
    std::vector<Complex> c(l + 1);
    for (int m = 0; m <= l; ++m) c[m] = Complex(0.1 * (l + 1) - 0.05 * m, m == 0 ? 0.0 : 0.03 * m);
    return std::optional<std::vector<Complex>>(c);
  };

  const auto s0 = op.ZeroState();
  const std::vector<double> m0 = op.Export(s0, "m0");
  auto r1 = op.Step(s0, dt, Scheme::ExponentialTrapezoid, load);
  auto r2 = op.Step(r1.state, dt, Scheme::ExponentialTrapezoid, load);
  Expect(r1.info.converged && r2.info.converged, "both steps converge");
  Expect(op.Export(s0, "m0") == m0 && std::abs(s0.t()) == 0.0,
         "Step leaves its argument unchanged");
  auto p = op.ZeroState();
  op.StepInPlace(p, dt, Scheme::ExponentialTrapezoid, load);
  op.StepInPlace(p, dt, Scheme::ExponentialTrapezoid, load);
  Expect(op.Export(r2.state, "coefficients") == op.Export(p, "coefficients") &&
             op.Export(r2.state, "m1") == op.Export(p, "m1"),
         "two steps by Step equal two by StepInPlace bit for bit");
  Expect(std::abs(r2.state.t() - 2.0 * dt) < 1e-15 * dt, "the new state's time");
  bool shapes = r2.solutions.spheroidal.size() == static_cast<size_t>(lmax + 1) &&
                r2.solutions.toroidal.size() == static_cast<size_t>(lmax);
  for (const auto& [l, x] : r2.solutions.spheroidal) shapes = shapes && x.size() % (l + 1) == 0;
  for (const auto& [l, x] : r2.solutions.toroidal) shapes = shapes && l >= 1 && x.size() % (l + 1) == 0;
  Expect(shapes, "solutions: spheroidal 0..L and toroidal 1..L, ndof x (l + 1)");
  double top = 0.0;
  for (const auto& c : r2.solutions.spheroidal.at(2)) top = std::max(top, std::abs(c));
  Expect(top > 0.0 && std::isfinite(top), "degree 2 responds to the load");

  Expect(Throws<std::invalid_argument>([&] { op.Step(r2.state, -1.0, Scheme::BackwardEuler, load); }),
         "a negative dt throws std::invalid_argument");
  const Viscoelastic3D::LoadCallable3D tide = [](int l, double) {
    std::vector<Complex> c(l + 1, Complex(0.0, 0.0));
    if (l == 1) c[0] = Complex(1e-3, 0.0);
    return std::optional<std::vector<Complex>>(c);
  };
  const std::vector<double> before = op.Export(p, "m2");
  std::string message;
  Expect(Throws<std::invalid_argument>(
             [&] { op.StepInPlace(p, dt, Scheme::ExponentialTrapezoid, load, tide); }, &message) &&
             message.find("degree-1") != std::string::npos && op.Export(p, "m2") == before,
         "a degree-1 tide throws std::invalid_argument, the state unchanged");
  const Viscoelastic3D::LoadCallable3D failing =
      [](int, double) -> std::optional<std::vector<Complex>> { throw LoadFailure{}; };
  Expect(Throws<LoadFailure>([&] { op.Step(r2.state, dt, Scheme::ExponentialTrapezoid, failing); }),
         "a load channel's exception comes out of Step");
  // a trajectory warm-started by extrapolation reaches the plain
  // trajectory's states, in fewer passes once it holds a history
  {
    auto plain = op.ZeroState();
    auto warm = op.ZeroState();
    Viscoelastic3D::Extrapolator ex(op, 2);
    int plainPasses = 0, warmPasses = 0;
    double worst = 0.0;
    for (int k = 0; k < 6; ++k) {
      const auto a = op.StepInPlace(plain, dt, Scheme::ExponentialTrapezoid, load);
      const auto b = op.StepInPlace(warm, ex, dt, Scheme::ExponentialTrapezoid, load);
      if (k >= 3) {
        plainPasses += a.iterations;
        warmPasses += b.iterations;
      }
      const auto x = op.Export(plain, "coefficients");
      const auto y = op.Export(warm, "coefficients");
      double top = 0.0, diff = 0.0;
      for (size_t i = 0; i < x.size(); ++i) {
        top = std::max(top, std::abs(x[i]));
        diff = std::max(diff, std::abs(x[i] - y[i]));
      }
      worst = std::max(worst, diff / top);
    }
    std::printf("[cpp] extrapolated trajectory: passes %d vs %d, state difference %.2e\n",
                warmPasses, plainPasses, worst);
    Expect(worst < 1e-8, "the extrapolated trajectory reaches the plain one's states");
    Expect(warmPasses < plainPasses, "the extrapolated trajectory takes fewer passes");
    // a warm start that fails its guard falls back to the step's own start
    auto q = op.ZeroState();
    std::vector<double> zeros(static_cast<size_t>(spectral), 0.0);
    double* bad = nullptr;
    if (hipMalloc(&bad, sizeof(double) * zeros.size()) != hipSuccess) return 1;
    for (auto& v : zeros) v = 1.0;
    if (hipMemcpy(bad, zeros.data(), sizeof(double) * zeros.size(), hipMemcpyHostToDevice) !=
        hipSuccess)
      return 1;
    const auto g = op.StepInPlace(q, dt, Scheme::ExponentialTrapezoid, load, {}, bad, 1e-3);
    Expect(g.warmFallback && g.converged, "a guarded warm start is dropped");
    (void)hipFree(bad);
  }
  // the operator shares the Earth: destroyed first, stepping with two new
  // dt values (each a new flavour assembled from the Earth) still works
  {
    auto owned = std::make_unique<Earth>(SmallEarth());
    Viscoelastic3D::Options o2;
    o2.max_iterations = 40;
    Viscoelastic3D other(*owned, 4, nullptr, o2);
    owned.reset();
    auto s = other.ZeroState();
    const auto a = other.StepInPlace(s, 0.5 * dt, Scheme::ExponentialTrapezoid, load);
    const auto b = other.StepInPlace(s, 0.25 * dt, Scheme::BackwardEuler, load);
    Expect(a.converged && b.converged, "an operator outlives its Earth handle");
    // a state of one operator is refused by another before any launch
    Expect(Throws<std::invalid_argument>(
               [&] { op.StepInPlace(s, dt, Scheme::ExponentialTrapezoid, load); }),
           "a state of another operator is refused");
    Expect(Throws<std::invalid_argument>([&] { (void)op.Export(s, "m0"); }),
           "exporting another operator's state is refused");
  }
  // two realisations side by side (the second without a lateral field: the
  // first's field and zeros), each with its own info
  {
    Viscoelastic3D::Options o3 = options;
    o3.realisations = 2;
    Viscoelastic3D pair(earth, lmax, nullptr, o3);
    Expect(pair.Realisations() == 2 && pair.RadialNodes() == nodes,
           "an operator of two realisations");
    double* both = nullptr;
    const size_t field = static_cast<size_t>(nodes) * ntheta * nphi;
    if (hipMalloc(&both, sizeof(double) * 2 * field) != hipSuccess) return 1;
    if (hipMemcpy(both, delta_dev, sizeof(double) * field, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemset(both + field, 0, sizeof(double) * field) != hipSuccess)
      return 1;
    sl3dpp_amd::Check(sl3_operator_set_delta(pair.Handle(), both));
    auto s2 = pair.ZeroState();
    const auto info = pair.StepInPlace(s2, dt, Scheme::ExponentialTrapezoid, load);
    const auto per = pair.RealisationInfos();
    Expect(per.size() == 2 && per[0].converged && per[1].converged && info.converged,
           "both realisations converge");
    Expect(per.size() == 2 &&
               info.iterations == std::max(per[0].iterations, per[1].iterations),
           "the step's passes are the most of its realisations'");
    (void)hipFree(both);
  }
  (void)hipFree(delta_dev);
  return failures;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  try {
    const int f = mode == "gpu" ? Gpu() : Cpu();
    if (f == 0) std::printf("[cpp] ok\n");
    return f == 0 ? 0 : 1;
  } catch (const std::exception& e) {
    std::printf("[cpp] unexpected exception: %s\n", e.what());
    return 2;
  }
}
