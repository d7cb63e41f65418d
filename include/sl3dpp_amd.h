/*
 * sl3dpp_amd.h -- the C ABI of the MI355X (gfx950) implementation of the
 * data-parallel core of sl3dpp's 3D lateral-viscosity step.
 *
 * Plain C: pointers, sizes and status codes; no HIP or torch types. Device
 * pointers are marked `_dev`; everything else is host memory borrowed for
 * the duration of the call. Calls run on the library's stream of the calling
 * thread (see sl3_synchronize) and are asynchronous unless they return host
 * data.
 *
 * Status codes: 0 success; < 0 argument or runtime failure (message from
 * sl3_last_error); > 0 a numerical failure the reference reports by
 * exception (a factorisation that failed outright).
 */
#ifndef SL3DPP_AMD_H
#define SL3DPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SL3_OK = 0,
  SL3_EINVAL = -1,   /* std::invalid_argument in the reference */
  SL3_EHIP = -2,     /* a HIP runtime failure */
  SL3_ENOMEM = -3,
  SL3_EINTERNAL = -4,
  SL3_ENUMERIC = 1   /* std::runtime_error: a factorisation failed */
};

/* The message of the last failed call on the calling thread. */
const char* sl3_last_error(void);
/* Library version string. */
const char* sl3_version(void);
/* Waits for every call issued so far on the calling thread's library
 * stream (one per host thread). */
int sl3_synchronize(void);
/* Copies bytes between device addresses on the library's stream (for a
 * caller holding views, sl3_step_view, and nothing else of HIP). */
int sl3_copy(void* dst_dev, const void* src_dev, int64_t bytes);

/* Kernel timing with device events on the library's stream (off by
 * default): enable, run, then read a kernel's total milliseconds and launch
 * count ("legendre_synthesis", "legendre_analysis", "ring", "force",
 * "band_solve", "strain", "anderson", "residual", ...); names lists them,
 * newline-separated. */
int sl3_profile_enable(int on);
/* Limits the timing to the named kernels (comma-separated; NULL or "": all),
 * so that a timed run carries events only where they are read. */
int sl3_profile_select(const char* kernels);
int sl3_profile_reset(void);
int sl3_profile_read(const char* kernel, double* total_ms, int64_t* launches);
int sl3_profile_names(char* buffer, int64_t capacity);

/* The tile-shape tuner (no counterpart in the reference): an operator times
 * every tile shape of its Legendre kernels on its own shape when it is
 * made (degrees per analysis workgroup 16, 32, 64 or 128; colatitudes per
 * synthesis workgroup 32, 64, 112 or 208) and keeps the fastest; a shape met
 * before in the process takes the choice made then. The 1D evolution
 * times the band panel of its batched Cholesky solve (rows staged per
 * panel: 4, 8 or 16) per flavour the same way. Every choice gives the same
 * bits. This call fixes the shapes of operators and flavours made from now
 * on (0: tune), for tests and timing studies; the exports "tuning" report
 * the choice and the timings. */
int sl3_tuning_set(int analysis_rows, int synthesis_theta, int band_panel_rows);

/* ===================================================================== *
 * Banded Cholesky and LU, FP64, LAPACK band storage (column-major bands, *
 * 0-based pivots). Replace sl3dpp::detail::BandCholeskyFactor,          *
 * BandCholeskySolve, BandLuFactor and BandLuSolve                       *
 * (include/sl3dpp/detail/BandSolvers.hpp:24-49, src/BandSolvers.cpp).   *
 * ===================================================================== */

/* One system, as BandCholeskyFactor (BandSolvers.hpp:24-32): factorises
 * the (kd+1) x n upper band at ab_dev in place; *info = 0 or the 1-based
 * index of the first non-positive leading minor. */
int sl3_band_cholesky_factor(double* ab_dev, int n, int kd, int* info);
/* As BandCholeskySolve (BandSolvers.hpp:34-37): B (n x nrhs column-major)
 * at b_dev overwritten by X. */
int sl3_band_cholesky_solve(const double* ab_dev, int n, int kd,
                            double* b_dev, int nrhs);
/* As BandLuFactor (BandSolvers.hpp:39-45): (2kl+ku+1) x n band, pivots
 * (0-based) to piv_dev; *info = 0 or the 1-based first zero pivot. */
int sl3_band_lu_factor(double* ab_dev, int n, int kl, int ku, int* piv_dev,
                       int* info);
/* As BandLuSolve (BandSolvers.hpp:47-49). */
int sl3_band_lu_solve(const double* ab_dev, int n, int kl, int ku,
                      const int* piv_dev, double* b_dev, int nrhs);

/* Batched forms: `count` independent systems in one launch each (the
 * per-degree loops of DegreeOperator::Factorise and Viscoelastic3D::
 * SolveForces, src/DegreeOperator.cpp:775-901, src/Viscoelastic3D.cpp:
 * 876-978). Per system s (host arrays): n[s], kd[s] (Cholesky) or kl[s],
 * ku[s] (LU), the offset in doubles of its band in ab_dev, of its pivots
 * in piv_dev, of its right-hand sides in b_dev, their count nrhs[s] and
 * their layout: element (i, k) at b_dev[b_off[s] + i*row_stride[s] +
 * k*col_stride[s]] (column-major: row_stride 1, col_stride n). info_dev
 * receives each system's info. */
int sl3_band_cholesky_factor_batched(int count, const int* n, const int* kd,
                                     const int64_t* ab_off, double* ab_dev,
                                     int* info_dev);
int sl3_band_cholesky_solve_batched(int count, const int* n, const int* kd,
                                    const int64_t* ab_off,
                                    const double* ab_dev, const int* nrhs,
                                    const int64_t* b_off,
                                    const int64_t* row_stride,
                                    const int64_t* col_stride, double* b_dev);
int sl3_band_lu_factor_batched(int count, const int* n, const int* kl,
                               const int* ku, const int64_t* ab_off,
                               double* ab_dev, const int64_t* piv_off,
                               int* piv_dev, int* info_dev);
int sl3_band_lu_solve_batched(int count, const int* n, const int* kl,
                              const int* ku, const int64_t* ab_off,
                              const double* ab_dev, const int64_t* piv_off,
                              const int* piv_dev, const int* nrhs,
                              const int64_t* b_off, const int64_t* row_stride,
                              const int64_t* col_stride, double* b_dev);
/* sl3_band_lu_solve_batched that allocates nothing. It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops). The per-call system
 * descriptors and work list go to a 64 KiB device array of the library's
 * own (one per GPU, held until the process ends) instead of buffers
 * allocated per call; the kernels and results are sl3_band_lu_solve_
 * batched's. Limits: at most 512 systems and 4096 blocks of 64 right-hand
 * sides per call (SL3_EINVAL beyond, before any launch). Calls that
 * overlap on one GPU (from several host threads) race on that array. */
int sl3_band_lu_solve_batched_noalloc(int count, const int* n, const int* kl,
                                      const int* ku, const int64_t* ab_off,
                                      const double* ab_dev, const int64_t* piv_off,
                                      const int* piv_dev, const int* nrhs,
                                      const int64_t* b_off, const int64_t* row_stride,
                                      const int64_t* col_stride, double* b_dev);

/* ===================================================================== *
 * The Earth and its radial set-up (host): a layered body with properties *
 * linear in radius per layer, meshed into GLL elements, one Maxwell       *
 * layer (include/sl3dpp/RadialMesh.hpp, Material.hpp, Rheology.hpp:       *
 * Maxwell; src/Viscoelastic.cpp:114-176 for the memory points' moduli).   *
 * SI inputs; everything inside is non-dimensional as the reference's.    *
 * ===================================================================== */
#define SL3_MAX_LAYERS 8
typedef struct {
  int layers;
  double r0[SL3_MAX_LAYERS], r1[SL3_MAX_LAYERS];         /* m */
  double rho0[SL3_MAX_LAYERS], rho1[SL3_MAX_LAYERS];     /* kg m^-3 */
  double kappa0[SL3_MAX_LAYERS], kappa1[SL3_MAX_LAYERS]; /* Pa */
  double mu0[SL3_MAX_LAYERS], mu1[SL3_MAX_LAYERS];       /* Pa; 0 = fluid */
  int elements[SL3_MAX_LAYERS];
  int ngll;
  int viscous_layer;
  double viscosity_boundary; /* m */
  double viscosity_upper;    /* Pa s above the boundary */
  double viscosity_lower;    /* Pa s at or below it */
} sl3_earth_spec;

typedef struct sl3_earth sl3_earth;
int sl3_earth_create(const sl3_earth_spec* spec, sl3_earth** out);
int sl3_earth_destroy(sl3_earth* earth);
/* The non-dimensional scales: length (m), time (s), mass (kg). */
int sl3_earth_scales(const sl3_earth* earth, double* length, double* time,
                     double* mass);
/* Mesh element count and nodes per element. */
int sl3_earth_mesh(const sl3_earth* earth, int* elements, int* ngll);

/* Host arrays of the set-up (no device work): "radius" (element x gll
 * node radii), "jac" (per element), "gll_weight", "gll_deriv" (ngll x ngll),
 * "point_value", "point_deriv" (points x ngll), "point_weight",
 * "point_radius" (element x point), "point_mu", "point_itau" (the Maxwell
 * branch at the points, element x point), "fluid" (per element),
 * "surface_gravity", "memory_elements" (first, count). dst NULL returns the
 * size through *count. */
int sl3_earth_export(const sl3_earth* earth, const char* name, double* dst,
                     int64_t* count);
/* One assembled degree system (host; DegreeOperator / ToroidalOperator,
 * src/DegreeOperator.cpp:24-474, src/ToroidalOperator.cpp:14-147): kind 0
 * spheroidal, 1 toroidal, with the memory points' shear of (scheme, dt)
 * (dt = 0 the elastic operators'; dt < 0 none). Returns ndof, kd, the first
 * element, the pinned dof (or -1); fills, where non-NULL, the unscaled
 * upper band ((kd+1) x ndof), the dof map (per global mesh node: U, V, phi
 * for spheroidal, W for toroidal; -1 absent), the unit load vector
 * (spheroidal, ndof; DegreeOperator::Load, src/DegreeOperator.cpp:960-970)
 * and the unit tide vector (spheroidal, ndof; the force of the potential
 * (r/a)^p, p = 2 at l = 0 and l above: DegreeOperator::Tide,
 * src/DegreeOperator.cpp:999-1038, as AddUnitChannels takes it,
 * src/Viscoelastic3D.cpp:1049). */
int sl3_earth_degree_system(const sl3_earth* earth, int l, int kind, int scheme,
                            double dt, double eps, int* ndof, int* kd,
                            int* first_element, int* pinned, double* band,
                            int* dofs, double* load, double* tide);

/* An Earth from its nodal material on a given mesh, as a caller holding the
 * reference's Material with its RadialMesh (include/sl3dpp/Material.hpp,
 * include/sl3dpp/RadialMesh.hpp) and Rheology::Maxwell's branch
 * (include/sl3dpp/Rheology.hpp:100-190) passes them. SI; arrays per
 * (element, GLL node), element-major, unless noted. At most one Maxwell
 * branch. */
typedef struct {
  int elements, ngll;
  const double* edges;       /* elements + 1 element boundaries (m), increasing */
  const double* rho;         /* kg m^-3 */
  const double* drho;        /* d rho / dr, kg m^-4 */
  const double* gravity;     /* m s^-2 */
  const double *A, *C, *F, *L, *N; /* transversely isotropic moduli, Pa */
  const int* fluid;          /* per element */
  int branches;              /* 0 (elastic) or 1 */
  const double* branch_mu;   /* the branch's shear, Pa */
  const double* branch_itau; /* mu / eta, 1/s */
  double length, density;    /* the scales (m, kg m^-3); 0 for 6371 km and 5515 */
} sl3_earth_nodal;
int sl3_earth_create_nodal(const sl3_earth_nodal* spec, sl3_earth** out);

/* DegreeOperator and ToroidalOperator (include/sl3dpp/DegreeOperator.hpp:
 * 94-384, include/sl3dpp/ToroidalOperator.hpp:58-176): one degree's radial
 * system, assembled on the host, factorised and solved on the GPU. */
enum { SL3_SPHEROIDAL = 0, SL3_TOROIDAL = 1 };
enum { SL3_MATERIAL_ELASTIC = 0, SL3_MATERIAL_LONG_TERM = 1 };
enum { SL3_POINTS_GAUSS = 0, SL3_POINTS_GLL = 1 };
/* DegreeOperator::Options (include/sl3dpp/DegreeOperator.hpp:101-131). */
typedef struct {
  double eps;       /* graded truncation level, 1e-8 */
  int gravitating;  /* 1; 0 for the non-gravitating spheroidal system */
  int material;     /* SL3_MATERIAL_ELASTIC: the unrelaxed material;
                       SL3_MATERIAL_LONG_TERM: the Maxwell branch's shear
                       removed (the step's systems) */
  const double* point_shear; /* NULL, or a shear modulus (Pa) added at the
                                memory points, (element, point) */
  int point_set;    /* SL3_POINTS_GAUSS (ngll - 1 points, the step's) or
                       SL3_POINTS_GLL (the ngll nodes) */
} sl3_degree_options;
void sl3_degree_options_default(sl3_degree_options* options);
typedef struct sl3_degree_operator sl3_degree_operator;
/* DegreeOperator(material, l, options) / ToroidalOperator(material, l,
 * options) (src/DegreeOperator.cpp:24-55, src/ToroidalOperator.cpp:16-30);
 * kind SL3_SPHEROIDAL or SL3_TOROIDAL. */
int sl3_degree_operator_create(const sl3_earth* earth, int l, int kind,
                               const sl3_degree_options* options,
                               sl3_degree_operator** out);
int sl3_degree_operator_destroy(sl3_degree_operator* op);
/* DofCount, the band half-width, FirstElement and the pinned dof (-1 for
 * none). */
int sl3_degree_operator_shape(const sl3_degree_operator* op, int* ndof, int* kd,
                              int* first_element, int* pinned);
/* Host arrays by name: "band" (the unscaled upper band, (kd+1) x ndof),
 * "scale" (Factorise's diagonal scaling), "dofs" (per global mesh node: U,
 * V, phi for spheroidal, W for toroidal, -1 where absent, as doubles), and
 * for the spheroidal system "load", "load_force", "load_potential",
 * "tangential", "traction", "traction_tangential" (DegreeOperator::Load
 * (Both, Force, Potential), Tangential, Traction, TractionTangential,
 * src/DegreeOperator.cpp:960-997) and "tide:<p>" (Tide(p), p = -1 for l,
 * :999-1038). dst NULL returns the size through *count. */
int sl3_degree_operator_export(const sl3_degree_operator* op, const char* name,
                               double* dst, int64_t* count);
/* DegreeOperator::Solve (src/DegreeOperator.cpp:861-920): B (ndof x nrhs,
 * column-major, device) overwritten by X. The factorisation (Cholesky, LU
 * where the system is not positive definite; src/DegreeOperator.cpp:
 * 775-852) is made on the first call and kept; used_cholesky (may be NULL)
 * says which. */
int sl3_degree_operator_solve(sl3_degree_operator* op, double* b_dev, int nrhs,
                              int* used_cholesky);
/* DegreeOperator::Apply (src/DegreeOperator.cpp:946-949): y = A x with the
 * unscaled system, nrhs column-major columns, host arrays. */
int sl3_degree_operator_apply(const sl3_degree_operator* op, const double* x,
                              int nrhs, double* y);

/* ===================================================================== *
 * Viscoelastic3D (include/sl3dpp/Viscoelastic3D.hpp:89-675): the 3D      *
 * lateral-viscosity step on the GPU. One Maxwell branch, Gauss memory    *
 * points; the step's system closed by the fixed point with adaptive      *
 * Anderson mixing, or by preconditioned conjugate gradients.             *
 * ===================================================================== */
/* Options::StepSolver (include/sl3dpp/Viscoelastic3D.hpp:122-123). */
enum { SL3_FIXED_POINT = 0, SL3_CONJUGATE_GRADIENTS = 1 };

#define SL3_MAX_TRANSFORM_DEGREES 32

typedef struct {
  double eps;           /* truncation level, 1e-8 */
  double oversampling;  /* 1.5 */
  int max_iterations;   /* 40 */
  double rtol;          /* 1e-11 */
  double damping;       /* 1 */
  int anderson_memory;  /* 5; at most 16 */
  double anderson_rate; /* 0.5; <= 0 for Anderson from the first pass */
  int solver;           /* SL3_FIXED_POINT */
  double solve_rtol;    /* 1e-11: the relative residual in the preconditioned
                           energy norm the conjugate gradients stop at */
  int solve_max_iterations; /* 200 */
  int graph_passes;     /* 0: a pass's kernels are launched one by one; 1:
                           each pass replays a hipGraph captured on its first
                           use (per input/output tensor pair, flavour and tide
                           channel), the same kernels and results */
  int transform_degree_count; /* 0 */
  int transform_degrees[SL3_MAX_TRANSFORM_DEGREES];
                        /* Options::transform_degrees (include/sl3dpp/
                           Viscoelastic3D.hpp:172-175, GridCoupling.hpp:
                           99-107): the degrees in [2, lmax] the transforms
                           may be cut to with depth. Each memory element's
                           nodes are transformed at the least listed degree
                           (or lmax) at or above the largest degree whose
                           subset reaches the element, on a grid of that
                           degree (src/GridCoupling.cpp:163-190): the nodes
                           form blocks, and every grid field over the nodes
                           lays the blocks' grids end to end
                           (sl3_operator_blocks). Fixed when the operator is
                           made. */
  int realisations;     /* 1; at most 64. Independent realisations of the
                           step on one Earth and degree (each with its own
                           lateral field and state; load and tide shared),
                           stepped together: realisation r's memory nodes are
                           nodes r N .. r N + N - 1 of every node-indexed
                           tensor (N = sl3_operator_realisations' radial
                           nodes), every kernel of a pass takes them all, and
                           each realisation converges on its own residual and
                           Anderson history, as if stepped alone
                           (sl3_step_realisation_infos). With the fixed point
                           only, no transform degrees and no guarded warm
                           start; fixed when the operator is made. No
                           reference counterpart: the reference steps
                           realisations one operator at a time. */
} sl3_step_options;

typedef struct {
  int iterations;
  double residual;
  int converged;
  int mixed_passes;
  float start_ms, passes_ms, finish_ms; /* device time of the phases */
  int solve_iterations; /* conjugate-gradient iterations (StepInfo::solveIterations) */
  int warm_fallback;    /* sl3_step_guarded: the warm start was dropped */
} sl3_step_info;

/* Schemes (include/sl3dpp/StepWeights.hpp, StepRecipe.hpp): Sdirk2 runs
 * two backward Euler stages of length gamma dt, gamma = 1 - 1/sqrt(2). */
enum { SL3_EXPONENTIAL_TRAPEZOID = 0, SL3_BACKWARD_EULER = 1, SL3_SDIRK2 = 2 };

typedef struct sl3_operator sl3_operator;
typedef struct sl3_state sl3_state;

void sl3_step_options_default(sl3_step_options* options);
/* delta_dev: delta(r, theta, phi) sampled on every memory node's grid,
 * [node][theta][phi] on the device (viscosity eta_ref exp(delta)), or NULL
 * for none (src/Viscoelastic3D.cpp:226-287). The operator shares the
 * Earth: the earth handle may be destroyed first. */
int sl3_operator_create(const sl3_earth* earth, int lmax, const double* delta_dev,
                        const sl3_step_options* options, sl3_operator** out);
int sl3_operator_destroy(sl3_operator* op);
int sl3_operator_set_options(sl3_operator* op, const sl3_step_options* options);
/* Replaces the lateral field: itau(x) = itau_ref(node) exp(-delta(x)) from
 * delta_dev ([node][theta][phi], device; NULL for delta = 0). */
int sl3_operator_set_delta(sl3_operator* op, const double* delta_dev);
/* Memory nodes (of every realisation), grid shape, doubles of a spectral
 * tensor (device layout). The grid is lmax's; with transform degrees some
 * nodes have smaller ones (sl3_operator_blocks). */
int sl3_operator_shape(const sl3_operator* op, int* nodes, int* ntheta,
                       int* nphi, int64_t* spectral_doubles);
/* The node blocks (GridCoupling's FieldOffset / GridOf, include/sl3dpp/
 * GridCoupling.hpp:176-230): *count blocks (at most capacity are written;
 * any array may be NULL), each its first memory node, node count, degree
 * and grid (ntheta x nphi); *grid_values, the values of a grid field over
 * every node, the blocks end to end (each block's nodes [node][theta][phi]
 * on its grid). */
int sl3_operator_blocks(const sl3_operator* op, int capacity, int* count, int* first_node,
                        int* nodes, int* lmax, int* ntheta, int* nphi, int64_t* grid_values);
/* One block's colatitudes and Gauss-Legendre weights (its ntheta each). */
int sl3_operator_block_geometry(const sl3_operator* op, int block, double* theta,
                                double* wtheta);
/* Host arrays: node radii (nodes), colatitudes and their weights (ntheta). */
int sl3_operator_geometry(const sl3_operator* op, double* radii, double* theta,
                          double* wtheta);
/* The realisations stepped side by side (sl3_step_options.realisations)
 * and the memory nodes of each (the shape's nodes = realisations x
 * radial_nodes). No reference counterpart. */
int sl3_operator_realisations(const sl3_operator* op, int* realisations, int* radial_nodes);
/* Per realisation, the last step's (or converge's) own sl3_step_info
 * (passes, residual, convergence, mixed passes; the timings zero): *count
 * realisations, at most capacity written. The info a step returns for
 * several realisations takes the most passes and mixed passes, the largest
 * residual, and converged when every realisation converged. No reference
 * counterpart. */
int sl3_step_realisation_infos(const sl3_operator* op, int capacity, int* count,
                               sl3_step_info* infos);

/* LoadCallable3D: write degree l's m >= 0 columns at time t as (re, im)
 * pairs into columns (2(l+1) doubles) and return 1, or return 0 for none.
 * The same type serves the load and the tide channel. A negative return
 * (the caller's own failure, an exception in the reference's callable)
 * ends the step with SL3_EINVAL: the channels are gathered before the
 * state changes, except in Sdirk2's second stage, after which the state
 * holds the stages' combination. */
typedef int (*sl3_load_fn)(int l, double t, double* columns, void* user);

int sl3_state_zero(sl3_operator* op, sl3_state** out);
/* A copy of a state (device memory of its own): Viscoelastic3D::Step
 * (include/sl3dpp/Viscoelastic3D.hpp:407-410) returns a new State and leaves
 * its argument unchanged, which a caller keeping both gets by stepping a
 * clone. */
int sl3_state_clone(const sl3_operator* op, const sl3_state* state, sl3_state** out);
int sl3_state_destroy(sl3_state* state);
int sl3_state_time(const sl3_state* state, double* t);
/* Step (src/Viscoelastic3D.cpp:1982-2030; Viscoelastic3D::Step(state, dt,
 * scheme, load, tide), include/sl3dpp/Viscoelastic3D.hpp:407-410), in place
 * on the state. load, tide: the channels (NULL for none), both called with
 * user; a tide whose degree-1 columns are not zero is an error
 * (src/Viscoelastic3D.cpp:142-154). start_dev: a warm start for the fixed
 * point (device layout spectral tensor), or NULL for the step's own start
 * strain; Sdirk2 takes NULL. For Sdirk2 the info sums both stages' passes
 * and phase times. */
int sl3_step(sl3_operator* op, sl3_state* state, double dt, int scheme,
             sl3_load_fn load, sl3_load_fn tide, void* user,
             const double* start_dev, sl3_step_info* info);

/* sl3_step with a guard on the warm start (no counterpart in the
 * reference; the learned warm start's fallback): when the first pass from
 * start_dev changes it by a relative residual above `guard` (or a non-finite
 * one), the step drops it and converges from its own start, the dropped pass
 * counted in info->iterations and info->warm_fallback set. guard 0: no
 * guard (sl3_step). */
int sl3_step_guarded(sl3_operator* op, sl3_state* state, double dt, int scheme,
                     sl3_load_fn load, sl3_load_fn tide, void* user,
                     const double* start_dev, double guard, sl3_step_info* info);
/* dst = sum_j coefficients[j] sources[j] over a spectral tensor (1 to 8
 * terms; dst may be one of the sources). With the states' coefficients as
 * sources this forms the extrapolated warm start of a trajectory's next
 * step (the `start` of OpenStep::Converge, src/Viscoelastic3D.cpp:
 * 1916-1921). Device pointers. */
int sl3_spectral_combine(const sl3_operator* op, int count, const double* coefficients,
                         const double* const* sources_dev, double* dst_dev);
/* Warm starts along a trajectory (a caller of Converge's `start`,
 * src/Viscoelastic3D.cpp:1916-1921; no counterpart in the reference's
 * library): the extrapolator keeps the last order + 1 converged strains of
 * the states it stepped, and sl3_extrapolated_step starts each step from
 * their Lagrange extrapolation in time to the step's end (order 1..5; the
 * step's own start while fewer than two are kept, for Sdirk2, and when the
 * state is not the last one it stepped or has moved since, which starts a
 * new history). The
 * fixed point then converges to the same tolerance. guard as in
 * sl3_step_guarded. */
typedef struct sl3_extrapolator sl3_extrapolator;
int sl3_extrapolator_create(sl3_operator* op, int order, sl3_extrapolator** out);
int sl3_extrapolator_destroy(sl3_extrapolator* ex);
int sl3_extrapolator_reset(sl3_extrapolator* ex);
int sl3_extrapolated_step(sl3_extrapolator* ex, sl3_state* state, double dt, int scheme,
                          sl3_load_fn load, sl3_load_fn tide, void* user, double guard,
                          sl3_step_info* info);
/* The state's strain coefficients (device layout, the next step's start),
 * borrowed until the state is stepped or destroyed; NULL before a step. */
int sl3_state_spectral(const sl3_state* state, const double** spectral_dev);

/* OpenStep (include/sl3dpp/Viscoelastic3D.hpp:461-647) at the boundary:
 * Step is sl3_step_start + sl3_step_close + sl3_step_finish, and a driver
 * (Evolution3D, the adjoint and tangent sweeps) calls the pieces itself.
 * The open step lives in the operator's buffers: one open step per
 * operator at a time, the state kept alive until it is finished. */
/* StartStep (src/Viscoelastic3D.cpp:1223-1378), one stage (exponential
 * trapezoid or backward Euler; Sdirk2 is opened stage by stage): the
 * state's memory becomes the decayed memory m~ in place; the lateral weight,
 * beta, the fixed memory force and the start strain (the iterate, view
 * "strain") are formed, the channels gathered at the end time. */
int sl3_step_start(sl3_operator* op, sl3_state* state, double dt, int scheme,
                   sl3_load_fn load, sl3_load_fn tide, void* user);
/* One pass of the within-step fixed point (Respond, AddChannels, StrainOf:
 * src/Viscoelastic3D.cpp:1758-1786, 1029-1073): the strain coefficients
 * (device layout) the step's solutions take for s_dev, into next_dev. */
int sl3_step_pass(sl3_operator* op, const double* s_dev, double* next_dev);
/* A driver's pass (OpenStep::Pass, :540-547): next = pass(s), device
 * layout; return 0, or non-zero for a failure that ends Converge with
 * SL3_EINVAL. It may call sl3_step_pass and add its own terms. */
typedef int (*sl3_pass_fn)(const double* s_dev, double* next_dev, void* user);
/* Converge (src/Viscoelastic3D.cpp:1916-1974): s = pass(s) from the start
 * strain (or start_dev), adaptively Anderson-mixed under the operator's
 * options; pass NULL is sl3_step_pass. The result is the iterate (view
 * "strain"). */
int sl3_step_converge(sl3_operator* op, sl3_pass_fn pass, void* user,
                      const double* start_dev, sl3_step_info* info);
/* Close (:2008-2015): the step under its channels as the operator's solver
 * says (the fixed point, or the conjugate gradients). */
int sl3_step_close(sl3_operator* op, const double* start_dev, sl3_step_info* info);
/* Finish (:1885-1914) with strain_dev (device coefficients; NULL for the
 * iterate): m+ = m~ + beta d, the strain grid and coefficients kept, the
 * state at the end time; the step is closed. */
int sl3_step_finish(sl3_operator* op, const double* strain_dev);
/* Drops the open step unfinished: the state keeps the decayed memory. */
int sl3_step_cancel(sl3_operator* op);
/* OpenStep::AddMemory (:576-583): increment_dev[3] (grid blocks as the
 * state's m0 m1 m2) joins the decayed memory and the fixed force. */
int sl3_step_add_memory(sl3_operator* op, const double* const* increment_dev);
/* OpenStep::AddForce's tensor (:585-598): an analysed tensor (device
 * layout) joins the fixed force. */
int sl3_step_add_force(sl3_operator* op, const double* analysed_dev);
/* The discrete 3D adjoint's transposed pass of the open step
 * (src/AdjointDiscrete3D.cpp:998-1037: passWith's map without the sea-level
 * solve and the measurement duals): out = Analyse(W x Synthesise(
 * StrainOf(SolveForces(Entry(z)))) + src), with Entry the forces of z
 * (SpheroidalForces, ToroidalForce), StrainOf without channels, W the open
 * step's lateral weight and src a grid tensor (src_dev[3], grid blocks as
 * the state's m0 m1 m2; NULL for none). Device-layout spectral tensors; z
 * and out distinct. The adjoint of sl3_step_pass's linear part in the
 * quadrature-weighted pairing of spectral tensors. */
int sl3_step_transposed_pass(sl3_operator* op, const double* z_dev,
                             const double* const* src_dev, double* out_dev);
/* The open step's device buffers, borrowed (OpenStep's friends weight_,
 * beta_, memory_, fixed_: include/sl3dpp/Viscoelastic3D.hpp:614-646):
 * "weight", "beta" (grid values), "fixed", "strain" (device-layout spectral
 * tensors: the fixed memory force, the iterate), "memory0".."memory2" (the
 * decayed memory m~, open step only); count in doubles. */
int sl3_step_view(sl3_operator* op, const char* name, const double** dev,
                  int64_t* count);
/* OpenStep::RelativeChange (:1636-1669) of two device spectral tensors. */
int sl3_step_relative_change(sl3_operator* op, const double* next_dev,
                             const double* now_dev, double* out);

/* GridCoupling (include/sl3dpp/GridCoupling.hpp:166-384) over the
 * operator's memory nodes. Solutions and forces are in the reference's
 * Solutions layout: every task's ndof x (l + 1) complex columns
 * (column-major), spheroidal l = 0..lmax then toroidal 1..lmax, end to
 * end (sl3_coupling_solution_doubles); spectral tensors in the device
 * layout. */
int sl3_coupling_solution_doubles(const sl3_operator* op, int64_t* count);
/* SpheroidalForce + ToroidalForce (src/GridCoupling.cpp:691-809): the
 * memory force of an analysed tensor on every degree system. */
int sl3_coupling_force(sl3_operator* op, const double* weighted_dev, double* forces_dev);
/* AddStrainOfSolution over every degree (src/GridCoupling.cpp:595-689;
 * StrainOf without channels): the strain coefficients of solutions. */
int sl3_coupling_strain(sl3_operator* op, const double* solutions_dev, double* spectral_dev);
/* Multiply (src/GridCoupling.cpp:497-509): a grid tensor (blocks as m0 m1
 * m2) times a weight per grid point, in place. */
int sl3_coupling_multiply(sl3_operator* op, double* c0_dev, double* c1_dev, double* r_dev,
                          const double* weight_dev);
/* ScaleByNode (src/GridCoupling.cpp:524-540): each node's coefficients
 * times factor[node] (host, one per memory node), in place. */
int sl3_coupling_scale_by_node(sl3_operator* op, const double* factor, double* spectral_dev);

/* Host copies for checks and outputs. Names:
 *   state:    "coefficients" (the strain coefficients, reference l-major
 *             layout per node, blocks spin -2, -1, 0 end to end, complex),
 *             "coefficients_device" (device layout), "m0" "m1" "m2"
 *             (memory grid components), "strain0" "strain1" "strain2";
 *   operator: "weight" "beta" "itau" (grid fields), "fixed" "analysed"
 *             (device layout spectral), "solution:<task>" (the last pass's
 *             solution of a task, ndof x (l+1) complex column-major),
 *             "tasks" (per task: l, kind, ndof, first memory node, memory
 *             dofs), "residuals" (last step), "solve_energy" (the last
 *             conjugate-gradient solve's <K^-1 f, f>, its residuals' scale),
 *             "tuning" (the Legendre tile shapes chosen, analysis rows and
 *             synthesis colatitudes, then the ms per launch timed for the
 *             analysis choices 16, 32, 64, 128 and the synthesis choices
 *             32, 64, 112, 208).
 * dst NULL returns the size in doubles through *count. */
int sl3_state_export(const sl3_operator* op, const sl3_state* state,
                     const char* name, double* dst, int64_t* count);
int sl3_operator_export(sl3_operator* op, const char* name, double* dst,
                        int64_t* count);
/* Copies a spectral tensor between the device layout and the reference's
 * (both device pointers). */
int sl3_spectral_to_reference(const sl3_operator* op, const double* device_dev,
                              double* reference_dev);
int sl3_spectral_from_reference(const sl3_operator* op, const double* reference_dev,
                                double* device_dev);

/* ===================================================================== *
 * ViscoelasticOperator (include/sl3dpp/Viscoelastic.hpp:163-301): the    *
 * radially symmetric Maxwell evolution in the degree basis, every        *
 * spheroidal degree 0..lmax with all its orders as columns (2 (l + 1)    *
 * real columns: Re of m = 0..l, then Im), one branch, Gauss memory       *
 * points, batched on the GPU. The load is one surface-mass channel: a    *
 * fixed pattern times a ramp sampled at the step times.                  *
 * ===================================================================== */
typedef struct sl3_maxwell1d sl3_maxwell1d;
/* ViscoelasticOperator(rheology, lmax, options) with ZeroState: every
 * degree 0..lmax, eps the graded truncation level (1e-8). Shares the
 * Earth. */
int sl3_maxwell1d_create(const sl3_earth* earth, int lmax, double eps, sl3_maxwell1d** out);
int sl3_maxwell1d_destroy(sl3_maxwell1d* h);
/* Back to ZeroState (t = 0, no memory, no strain). */
int sl3_maxwell1d_reset(sl3_maxwell1d* h);
/* The load channel (the LoadCallable of Step, include/sl3dpp/
 * Viscoelastic.hpp:287-291): pattern, (lmax+1)(lmax+2) doubles, complex
 * (Re, Im) per (l, m >= 0) at l(l+1)/2 + m, a non-dimensional surface mass;
 * ramp[k] its factor at step time t0 + k dt (ramp_count >= the steps to be
 * taken + 1). Host arrays. */
int sl3_maxwell1d_set_load(sl3_maxwell1d* h, const double* pattern, int64_t pattern_count,
                           const double* ramp, int64_t ramp_count);
/* steps calls of Step(state, dt, scheme, load) (src/Viscoelastic.cpp:
 * 1016-1075), scheme SL3_EXPONENTIAL_TRAPEZOID or SL3_BACKWARD_EULER; the
 * first from ZeroState takes its elastic start under the load at t0 (the
 * zero-length step). use_graph != 0 replays one captured hipGraph per
 * step; both give the same bits. */
int sl3_maxwell1d_evolve(sl3_maxwell1d* h, double dt, int scheme, int steps, int use_graph);
/* Sizes: memory (and strain) doubles, solution doubles, memory nodes, the
 * steps taken. Any may be NULL. */
int sl3_maxwell1d_shape(const sl3_maxwell1d* h, int64_t* memory_doubles,
                        int64_t* solution_doubles, int* nodes, int* steps_taken);
/* One degree's layout: its dof count, first memory node, and offsets into
 * the memory ([node][scalar M, R, S][col]) and the solutions ([dof][col],
 * row-major). */
int sl3_maxwell1d_degree(const sl3_maxwell1d* h, int l, int* ndof, int* first_node,
                         int64_t* memory_offset, int64_t* solution_offset);
/* Host copies by name: "memory", "strain" (the state's cache), "solutions"
 * (the last step's, unscaled, with the channel), "tuning" (4 doubles: the
 * last evolution's band panel rows, then the ms per solve timed for panels
 * of 4, 8 and 16 rows). count in doubles. */
int sl3_maxwell1d_export(const sl3_maxwell1d* h, const char* name, double* dst, int64_t count);
/* ViscoelasticOperator::Surface (include/sl3dpp/Viscoelastic.hpp:262-267):
 * component 0 U, 1 V, 2 phi of the last solutions at the surface per
 * degree and column, out_dev[l (l+1) + col], (lmax+1)(lmax+2) doubles on
 * the device. */
int sl3_maxwell1d_surface(const sl3_maxwell1d* h, int component, double* out_dev);

/* ===================================================================== *
 * The learned warm start (no counterpart in the reference: it feeds the  *
 * `start` of OpenStep::Converge, src/Viscoelastic3D.cpp:1916-1921).      *
 * ===================================================================== */
/* The rows of degree <= cut of a device-layout spectral tensor (spin -2,
 * -1 and 0 blocks; a row is one (l, m) over every memory node). */
int sl3_spectral_rows(sl3_operator* op, int cut, int64_t* count);
/* Per memory node, the rows of degree <= cut as float features:
 * out[node * ld + col0 + 2 e + c] = ((spectral - base)[row e][node].c -
 * mean[2 e + c]) / scale[2 e + c], c = re, im. base, mean/scale may be NULL.
 * Device pointers. */
int sl3_spectral_gather(sl3_operator* op, const double* spectral_dev, const double* base_dev,
                        int cut, const float* mean_dev, const float* scale_dev, float* out_dev,
                        int ld, int col0);
/* out = base (zero when NULL) with the rows of degree <= cut increased by
 * in[node * ld + 2 e + c] * scale + mean. Device pointers. */
int sl3_spectral_scatter(sl3_operator* op, const float* in_dev, int ld, int cut,
                         const float* mean_dev, const float* scale_dev, const double* base_dev,
                         double* out_dev);

/* A multilayer perceptron on the BF16 matrix cores (FP32 accumulation):
 * widths[0] -> ... -> widths[layers], GELU (erf form) between layers, the
 * last linear; weights W_k (widths[k] x widths[k-1], row-major) and biases
 * kept in FP32 for Adam and in BF16 for the products. Initialised
 * U(-1/sqrt(in), 1/sqrt(in)) from `seed`. Training rows per call <= max_rows;
 * forward takes any number of rows. Device pointers, row-major FP32. */
typedef struct sl3_mlp sl3_mlp;
int sl3_mlp_create(int layers, const int* widths, int max_rows, uint64_t seed, sl3_mlp** out);
int sl3_mlp_destroy(sl3_mlp* mlp);
int sl3_mlp_forward(sl3_mlp* mlp, int64_t rows, const float* x_dev, float* y_dev);
/* Gradients of mean((y - target)^2) into the network's gradient buffer;
 * loss (may be NULL) receives the loss. */
int sl3_mlp_backward(sl3_mlp* mlp, int rows, const float* x_dev, const float* target_dev,
                     double* loss);
/* One Adam step (bias-corrected) from the gradient buffer. */
int sl3_mlp_adam(sl3_mlp* mlp, double lr, double beta1, double beta2, double eps);
/* Host copies by name: "weight:k", "bias:k", "grad_weight:k", "grad_bias:k"
 * (k = 1..layers), and "chunks" (the tuner's choices: kind, M, N, K, slice
 * depth per product shape); dst NULL returns the size through *count. Import sets a
 * weight or bias. */
int sl3_mlp_export(sl3_mlp* mlp, const char* name, float* dst, int64_t* count);
int sl3_mlp_import(sl3_mlp* mlp, const char* name, const float* src, int64_t count);
/* The products' LDS slice depth: 0 (default) times 32 and 128 on a shape's
 * first use and keeps the faster; 32 or 128 fixes it. Both give the same
 * bits. */
int sl3_mlp_set_chunk(sl3_mlp* mlp, int chunk);
/* Column means and scales (population standard deviation, at least
 * floor x the largest) of a rows x cols matrix. Device pointers. */
int sl3_column_moments(int64_t rows, int cols, const float* x_dev, float floor, float* mean_dev,
                       float* scale_dev);
/* x = (x - mean) / scale per column, in place; inverse != 0: x scale + mean. */
int sl3_column_standardise(int64_t rows, int cols, float* x_dev, const float* mean_dev,
                           const float* scale_dev, int inverse);

/* ===================================================================== *
 * Batched generalised spherical-harmonic transforms over the operator's *
 * memory nodes on its grid: SurfaceGrid::SynthesiseSpinBatch /          *
 * ExpandSpinBatch (spin -2, -1) and SynthesiseBatch / ExpandBatch       *
 * (spin 0, real), include/sl3dpp/SurfaceGrid.hpp:164-253. Coefficients   *
 * in the reference layout per node, end to end; fields [node][theta]     *
 * [phi] (complex for spins -2, -1, real for 0). Device pointers.         *
 * ===================================================================== */
int sl3_gsh_synthesise(const sl3_operator* op, int spin, const double* coef_dev,
                       double* field_dev);
int sl3_gsh_analyse(const sl3_operator* op, int spin, const double* field_dev,
                    double* coef_dev);

#ifdef __cplusplus
}
#endif
#endif
