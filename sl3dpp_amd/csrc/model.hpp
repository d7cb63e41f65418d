// Host-side set-up of the 3D step (FP64, non-dimensional like the
// reference): the layered Earth, the radial GLL mesh, the nodal material,
// the Maxwell rheology, the memory points and the 1D reference's point
// moduli. These build what the GPU path consumes; none of it is per-pass
// work.
#pragma once

#include <cmath>
#include <memory>
#include <vector>

namespace sl3amd {

// The non-dimensional units of EarthModel::Dimensionless
// (include/sl3dpp/Scales.hpp:56-92, Geophysical): length the Earth's
// radius, density the mean density, time such that G density T^2 = 1.
struct Scales {
  static constexpr double kG = 6.6743e-11;
  static constexpr double kRadius = 6.371e6;
  static constexpr double kDensity = 5515.0;
  double length = kRadius;
  double density = kDensity;
  double Time() const { return 1.0 / std::sqrt(kG * density); }
  double Mass() const { return density * length * length * length; }
  double Modulus() const { return Mass() / (length * Time() * Time()); }
  double Viscosity() const { return Modulus() * Time(); }
  double SurfaceDensity() const { return Mass() / (length * length); }
};

// One layer: properties linear in radius between its bounds (SI), meshed
// into `elements` equal elements. mu = 0 everywhere makes it fluid.
struct LayerSpec {
  double r0, r1;
  double rho0, rho1, kappa0, kappa1, mu0, mu1;
  int elements;
};

struct EarthSpec {
  std::vector<LayerSpec> layers;  // bottom up, contiguous, from r = 0
  int ngll = 6;
  // Maxwell viscosity (Pa s) of one layer: `upper` above `boundary` (m),
  // `lower` at or below; the other layers elastic.
  int viscousLayer = 1;
  double viscosityBoundary = 5.701e6;
  double viscosityUpper = 5e20;
  double viscosityLower = 1e22;
};

// The radial GLL mesh (include/sl3dpp/RadialMesh.hpp; src/RadialMesh.cpp:
// 131-222): elements per layer, GLL nodes, the Lagrange derivative matrix.
class RadialMesh {
 public:
  RadialMesh(const std::vector<double>& breakpoints,
             const std::vector<int>& elementsPerSpan, int ngll);
  // One element between each pair of consecutive edges.
  RadialMesh(const std::vector<double>& edges, int ngll)
      : RadialMesh(edges, std::vector<int>(edges.empty() ? 0 : edges.size() - 1, 1), ngll) {}
  int ElementCount() const { return static_cast<int>(left_.size()); }
  int NGll() const { return ngll_; }
  int NodeCount() const { return ElementCount() * (ngll_ - 1) + 1; }
  double Left(int e) const { return left_[e]; }
  double Right(int e) const { return right_[e]; }
  double Jacobian(int e) const { return jac_[e]; }
  double Radius(int e, int i) const { return radius_[e * ngll_ + i]; }
  const std::vector<double>& Radii() const { return radius_; }
  int GlobalIndex(int e, int i) const { return e * (ngll_ - 1) + i; }
  int LayerOf(int e) const { return layer_[e]; }
  const std::vector<double>& Points() const { return xi_; }
  const std::vector<double>& Weights() const { return w_; }
  double Derivative(int k, int i) const { return deriv_[k * ngll_ + i]; }
  int ElementOf(double r) const;
  // Graded truncation r_l = rmax eps^(1/(l+1)) (RadialMesh.hpp:48-50).
  double TruncationRadius(int l, double eps) const;
  int StartElement(int l, double eps) const;

 private:
  int ngll_;
  std::vector<double> left_, right_, jac_, radius_, xi_, w_, deriv_;
  std::vector<int> layer_;
};

// Nodal material (include/sl3dpp/Material.hpp): density, its slope,
// gravity, fluid flags and the isotropic moduli A = C = kappa + 4mu/3,
// F = kappa - 2mu/3, L = N = mu, per (element, node).
struct Material {
  std::shared_ptr<const RadialMesh> mesh;
  std::vector<double> rho, drho, g, kappa, mu;
  std::vector<double> A, C, F, L, N;
  std::vector<char> fluid;  // per element
  double G = 1.0;
  double radius = 1.0;
  double surfaceGravity = 1.0;
  bool IsFluid(int e) const { return fluid[e] != 0; }
  // The same material with the shear modulus replaced (long-term solid).
  Material WithShear(const std::vector<double>& newMu) const;
};

// Maxwell rheology (include/sl3dpp/Rheology.hpp; src/Rheology.cpp:129-163):
// one branch, mu_1 = the unrelaxed shear, itau = mu / eta at the nodes.
struct Rheology {
  std::vector<double> branchMu, branchItau;  // (element, node)
  std::vector<double> muLongTerm;
  std::vector<char> active;  // per (element, node)
  bool ElementActive(const RadialMesh& mesh, int e) const;
};

// The memory points: the ngll - 1 interior Gauss-Legendre points and the
// element's GLL Lagrange basis there (include/sl3dpp/MemoryPoints.hpp:
// 60-78; src/MemoryPoints.cpp).
struct MemoryPoints {
  int count = 0, ngll = 0;
  std::vector<double> xi, weight, value, derivative;  // (count, ngll)
  double Value(int p, int i) const { return value[p * ngll + i]; }
  double Derivative(int p, int i) const { return derivative[p * ngll + i]; }
  double RadiusAt(const RadialMesh& mesh, int e, int p) const;
  // The element's GLL Lagrange basis at the points xi (ascending, in
  // [-1, 1]) with quadrature weights w.
  static MemoryPoints At(const RadialMesh& mesh, const std::vector<double>& xi,
                         const std::vector<double>& w);
  static MemoryPoints Gauss(const RadialMesh& mesh);
  // The GLL nodes themselves (MemoryPointSet::Gll).
  static MemoryPoints Gll(const RadialMesh& mesh);
};

// The 1D reference's quantities at the memory points
// (src/Viscoelastic.cpp:114-176, 324-345): each branch's modulus and
// inverse time interpolated and clamped at zero; the long-term material
// at the nodes; the elastic point shear sum_k mu_k.
struct Reference1D {
  MemoryPoints points;
  std::vector<double> pointMu, pointItau;  // (element, point), one branch
  Material nodal;                          // the long-term material
  std::vector<double> elasticShear;        // (element, point)
  // sum_k mu_k (1 - beta_k) at the points, the trapezoid's unrelaxed
  // fraction (ExponentialTrapezoid); 1/(1 + h) for BackwardEuler.
  std::vector<double> PointShearModulus(int scheme, double dt) const;
};

// Everything the step is built from.
struct Earth {
  Scales scales;
  EarthSpec spec;
  std::shared_ptr<const RadialMesh> mesh;
  Material material;
  Rheology rheology;
  Reference1D reference;
};

std::shared_ptr<const Earth> BuildEarth(const EarthSpec& spec);

// An Earth given by its nodal material on a given mesh, the reference's
// Material and Rheology::Maxwell as a caller holds them
// (include/sl3dpp/Material.hpp, include/sl3dpp/Rheology.hpp:100-190). SI
// throughout; arrays per (element, GLL node) unless noted.
struct NodalEarthSpec {
  std::vector<double> edges;  // element boundaries, elements + 1, increasing
  int ngll = 5;
  std::vector<double> rho, drho, g;  // kg m^-3, kg m^-4, m s^-2
  std::vector<double> A, C, F, L, N;  // Pa
  std::vector<char> fluid;            // per element
  // At most one Maxwell branch: its shear (Pa) and inverse relaxation time
  // mu / eta (1/s); empty for an elastic Earth.
  std::vector<double> branchMu, branchItau;
  double length = Scales::kRadius, density = Scales::kDensity;  // the scales
};
std::shared_ptr<const Earth> BuildEarth(const NodalEarthSpec& spec);

// Gauss-Legendre nodes (descending x) and weights; GLL nodes (ascending)
// and weights.
void GaussLegendre(int n, std::vector<double>& x, std::vector<double>& w);
void GaussLobattoLegendre(int n, std::vector<double>& x, std::vector<double>& w);
double LegendreP(int n, double x);

// Step weights (include/sl3dpp/StepWeights.hpp:98-144).
inline double Phi1(double h, double e) {
  if (std::abs(h) < 1e-3) {
    return 1.0 - h / 2.0 + h * h / 6.0 - h * h * h / 24.0 +
           h * h * h * h / 120.0;
  }
  return (1.0 - e) / h;
}

enum SchemeId { kExponentialTrapezoid = 0, kBackwardEuler = 1, kSdirk2 = 2 };

// Sdirk2's stage fraction gamma = 1 - 1/sqrt(2) (include/sl3dpp/StepRecipe.hpp:100).
constexpr double kSdirk2Gamma = 0.29289321881345247560;

}  // namespace sl3amd
