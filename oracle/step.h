/*
 * step.h -- the oracle's restatement of one 3D lateral-viscosity step
 * (Viscoelastic3D::Step with the fixed point, src/Viscoelastic3D.cpp:
 * 1223-1424, 1636-1669, 1758-1974, 1982-2030; GridCoupling,
 * src/GridCoupling.cpp:263-809; the degree solves, src/DegreeOperator.cpp:
 * 775-901, src/ToroidalOperator.cpp:190-272), FP64 on the CPU, in the
 * reference's own layouts. Test infrastructure only (see oracle.h).
 *
 * The problem's radial set-up (the assembled, unscaled degree bands, dof
 * maps, memory points) is an input: it is host set-up, not the path under
 * test. Everything from the factorisation on is restated here.
 */
#ifndef SL3_ORACLE_STEP_H
#define SL3_ORACLE_STEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int lmax, ntheta, nphi;
  const double* theta;   /* ntheta colatitudes */
  const double* wtheta;  /* their Gauss-Legendre weights */
  int nodes, np, ngll, elem0, elements;
  const double* radius;  /* per memory node */
  const double* jac;     /* per memory element */
  const double* pweight; /* np */
  const double* pvalue;  /* np x ngll */
  const double* pderiv;  /* np x ngll */
  const double* mu_node;
  const double* itau_ref_node;
  const double* itau;    /* nodes x ntheta x nphi */
  /* tasks: spheroidal l = 0..lmax, then toroidal l = 1..lmax */
  const int* ndof;
  const int* kd;
  const int* first_element;
  const int* pinned;        /* -1 or the pinned dof */
  const int64_t* band_off;  /* unscaled upper bands (kd+1) x ndof */
  const double* bands;
  const int64_t* dof_off;   /* memory elements x ngll x 2 */
  const int* dofs;
  const int64_t* load_off;  /* spheroidal tasks' unit loads (ndof) */
  const double* loads;
  const int64_t* tide_off;  /* spheroidal tasks' unit tides (ndof), or NULL */
  const double* tides;
  /* node blocks (GridCoupling's Options::transform_degrees,
   * src/GridCoupling.cpp:163-190): block_count 0 puts every node on the one
   * grid above; otherwise block k holds the nodes from block_first_node[k]
   * to the next block's first node, transformed at degree block_lmax[k] on
   * its own grid (block_ntheta[k] x block_nphi[k], colatitudes
   * block_theta[k], weights block_wtheta[k]); a node's coefficients of
   * degree above its block's are zero, and every grid field (itau, the
   * state's grids) lays the nodes' grids end to end. */
  int block_count;
  const int* block_first_node;
  const int* block_lmax;
  const int* block_ntheta;
  const int* block_nphi;
  const double* const* block_theta;
  const double* const* block_wtheta;
} sl3o_problem;

typedef struct {
  int max_iterations;
  double rtol, damping;
  int anderson_memory;
  double anderson_rate;
} sl3o_options;

typedef struct {
  int iterations;
  double residual;
  int converged;
  int mixed_passes;
  /* wall-clock milliseconds: the factorisations, tables and unit solves
   * (once per dt and scheme in the reference, which caches them), StartStep,
   * the fixed point's passes, Finish */
  double setup_ms, start_ms, passes_ms, finish_ms;
} sl3o_info;

/* The state (caller-owned arrays): m and strain grid components (c0, c1
 * complex as (re, im) pairs, r real; nodes x ntheta x nphi values), the
 * strain coefficients in the reference layout (blocks spin -2, -1, 0 end to
 * end, node-major, complex). Without coefficients the state must be the
 * zero state under a zero load (the elastic start strain is then zero). */
typedef struct {
  double t;
  double* m[3];
  double* strain[3];
  int has_strain;
  double* coef;
  int has_coef;
} sl3o_state;

/* Size in doubles of the reference-layout spectral tensor. */
int64_t sl3o_spectral_doubles(int lmax, int nodes);
/* Values of a grid field over every node (the nodes' grids end to end). */
int64_t sl3o_grid_values(const sl3o_problem* p);

/* One step in place; amp, tamp: the load and tide channels' columns at
 * the step's end time, complex per (l, m) at l(l+1)/2 + m (NULL for none;
 * a tide needs the problem's unit tides). scheme 0 the exponential
 * trapezoid, 1 backward Euler. Returns 0, non-zero when a factorisation
 * fails, 2 for a tide without unit tides. */
int sl3o_step(const sl3o_problem* p, sl3o_state* st, double dt, int scheme,
              const double* amp, const double* tamp, const sl3o_options* o,
              sl3o_info* info);

/* The residual history of the last sl3o_step (count passes). */
int sl3o_last_residuals(double* out, int capacity);

/* The last pass's solutions of the last sl3o_step for one task (spheroidal
 * l = 0..lmax, then toroidal 1..lmax): ndof x (l+1) complex, column-major,
 * unscaled, with the channels. Returns the size in doubles (out NULL for
 * the size only), or -1 for an unknown task. */
int64_t sl3o_last_solution(int task, double* out, int64_t capacity);

/* The pieces of a pass, one at a time (test infrastructure for the
 * reference's gates). Solutions and forces: every task's complex columns
 * (ndof x (l+1), column-major) end to end, spheroidal l = 0..lmax then
 * toroidal 1..lmax; coefficients in the reference layout (as sl3o_state). */
int64_t sl3o_solution_doubles(const sl3o_problem* p);
/* The memory force of a spectral tensor on every task
 * (src/GridCoupling.cpp:691-809). */
void sl3o_memory_force(const sl3o_problem* p, const double* weighted, double* forces);
/* The strain coefficients of the tasks' solutions (src/GridCoupling.cpp:
 * 635-689). */
void sl3o_strain_of(const sl3o_problem* p, const double* solutions, double* coef);
/* Synthesis and analysis of a whole tensor (grids as in sl3o_state). */
void sl3o_synthesise_tensor(const sl3o_problem* p, const double* coef, double* const grid[3]);
void sl3o_analyse_tensor(const sl3o_problem* p, double* const grid[3], double* coef);

/* The degree-basis (1D) Maxwell stepper at the memory points, spheroidal
 * degrees, one branch, no lateral field (ViscoelasticOperator::Step,
 * src/Viscoelastic.cpp:914-1076). m and d: the memory and the last
 * strain, three deviatoric scalars (M, R, S) complex per (degree, node,
 * order m = 0..l), degrees end to end (sl3o_scalar_doubles). The itau of
 * the problem is not read (only itau_ref_node). */
typedef struct {
  double t;
  double* m;
  double* d;
  int has_strain;
} sl3o_state1d;
int64_t sl3o_scalar_doubles(const sl3o_problem* p);
/* One step; solutions: the spheroidal tasks' (l = 0..lmax) solutions end
 * to end with the load channel (amp as in sl3o_step). Returns 0, or
 * non-zero when a factorisation fails. */
int sl3o_step_degree_basis(const sl3o_problem* p, sl3o_state1d* st, double dt, int scheme,
                           const double* amp, double* solutions);

#ifdef __cplusplus
}
#endif
#endif
