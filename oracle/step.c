/*
 * step.c -- the oracle's 3D step (see step.h). Test infrastructure only.
 */
#include "step.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct sl3o_mixer sl3o_mixer;
sl3o_mixer* sl3o_mixer_create(int memory);
void sl3o_mixer_destroy(sl3o_mixer* mx);
int sl3o_mixer_update(sl3o_mixer* mx, size_t n, double* x, const double* gx);

static double g_residuals[4096];
static int g_residual_count = 0;

int sl3o_last_residuals(double* out, int capacity) {
  const int n = g_residual_count < capacity ? g_residual_count : capacity;
  memcpy(out, g_residuals, sizeof(double) * n);
  return g_residual_count;
}

/* The last pass's solutions of the last step (StepResult::solutions), kept
 * until the next step. */
static double** g_solutions = NULL;
static int64_t* g_solution_size = NULL;
static int g_solution_tasks = 0;

static void keep_solutions(double** sol, int64_t* size, int ntasks) {
  for (int ti = 0; ti < g_solution_tasks; ++ti) free(g_solutions[ti]);
  free(g_solutions);
  free(g_solution_size);
  g_solutions = sol;
  g_solution_size = size;
  g_solution_tasks = ntasks;
}

int64_t sl3o_last_solution(int task, double* out, int64_t capacity) {
  if (task < 0 || task >= g_solution_tasks) return -1;
  const int64_t n = g_solution_size[task];
  if (out != NULL) memcpy(out, g_solutions[task], sizeof(double) * (n < capacity ? n : capacity));
  return n;
}

/* ---- layouts (include/sl3dpp/SurfaceGrid.hpp:116-118, 191-200) ---------- */

static int64_t block_size(int b, int lmax) {
  const int64_t L1 = lmax + 1;
  return b == 2 ? L1 * (lmax + 2) / 2 : L1 * L1 - (b == 0 ? 4 : 1);
}
static int spin_of(int b) { return b == 0 ? -2 : (b == 1 ? -1 : 0); }
static int64_t block_offset(int b, int lmax, int nodes) {
  int64_t off = 0;
  for (int k = 0; k < b; ++k) off += 2 * block_size(k, lmax) * nodes;
  return off;
}
int64_t sl3o_spectral_doubles(int lmax, int nodes) {
  return block_offset(3, lmax, nodes);
}
static int64_t index_of(int b, int l, int m) {
  if (b == 2) return (int64_t)l * (l + 1) / 2 + m;
  const int N = spin_of(b);
  return (int64_t)l * l - N * N + m + l;
}
/* pointer to the (re, im) pair of (block, node, l, m) */
static double* coef_at(double* s, int lmax, int nodes, int b, int node, int l, int m) {
  return s + block_offset(b, lmax, nodes) + 2 * (node * block_size(b, lmax) + index_of(b, l, m));
}

/* ---- transforms of a whole tensor, node by node ------------------------- */
/* A table-driven form of sht.c's direct sums (same formulas, the Wigner d
 * values and the longitude exponentials computed once per step): per node
 * and block, F_m(theta) = sum_l c_l d^l_{mN}(theta) f_lm and
 * f(phi_j) = sum_m F_m e^{i m phi_j}; analysis the adjoint with the
 * Gauss-Legendre and uniform weights. */
typedef struct {
  int lmax, ntheta, nphi;
  double* d[3];       /* [m - mlo][theta][l - lmin] */
  int64_t* moff[3];   /* per m - mlo */
  double* cosT;       /* [j][m + lmax] */
  double* sinT;
} tables_t;

static int mlo_of(int b, int lmax) { return b == 2 ? 0 : -lmax; }
static int lmin_of(int b, int m) {
  const int am = abs(m), an = abs(spin_of(b));
  return am > an ? am : an;
}

static void tables_build(tables_t* t, int L, int nt, int nph, const double* theta) {
  t->lmax = L;
  t->ntheta = nt;
  t->nphi = nph;
  for (int b = 0; b < 3; ++b) {
    const int mlo = mlo_of(b, L), orders = L - mlo + 1;
    t->moff[b] = (int64_t*)malloc(sizeof(int64_t) * (orders + 1));
    int64_t off = 0;
    for (int m = mlo; m <= L; ++m) {
      t->moff[b][m - mlo] = off;
      const int lmin = lmin_of(b, m);
      if (lmin <= L) off += (int64_t)nt * (L - lmin + 1);
    }
    t->moff[b][orders] = off;
    t->d[b] = (double*)malloc(sizeof(double) * (off > 0 ? off : 1));
#pragma omp parallel for schedule(dynamic)
    for (int m = mlo; m <= L; ++m) {
      const int lmin = lmin_of(b, m);
      if (lmin > L) continue;
      for (int i = 0; i < nt; ++i) {
        double* col = t->d[b] + t->moff[b][m - mlo] + (int64_t)i * (L - lmin + 1);
        sl3o_wigner_d_column(L, spin_of(b), m, theta[i], col);
        for (int l = lmin; l <= L; ++l) col[l - lmin] *= sqrt((2.0 * l + 1.0) / (4.0 * M_PI));
      }
    }
  }
  t->cosT = (double*)malloc(sizeof(double) * nph * (2 * L + 1));
  t->sinT = (double*)malloc(sizeof(double) * nph * (2 * L + 1));
  for (int j = 0; j < nph; ++j) {
    for (int m = -L; m <= L; ++m) {
      const double a = 2.0 * M_PI * (double)((int64_t)j * m % nph) / nph;
      t->cosT[(int64_t)j * (2 * L + 1) + m + L] = cos(a);
      t->sinT[(int64_t)j * (2 * L + 1) + m + L] = sin(a);
    }
  }
}

static void tables_free(tables_t* t) {
  for (int b = 0; b < 3; ++b) {
    free(t->d[b]);
    free(t->moff[b]);
  }
  free(t->cosT);
  free(t->sinT);
}

/* Every node's grid: its block's tables and weights, and where its values
 * start in a grid field. */
typedef struct {
  int count;
  tables_t* tables;       /* per block */
  const double** wtheta;  /* per block */
  int* node_block;        /* per node */
  size_t* node_offset;    /* per node, in grid values */
  size_t* node_size;      /* per node: its grid's values */
  size_t values;          /* over every node */
} grids_t;

static void grids_layout(const sl3o_problem* p, int* count, int* node_block,
                         size_t* node_offset, size_t* node_size, size_t* values) {
  const int blocks = p->block_count > 0 ? p->block_count : 1;
  size_t off = 0;
  for (int node = 0; node < p->nodes; ++node) {
    int k = 0;
    if (p->block_count > 0) {
      while (k + 1 < blocks && p->block_first_node[k + 1] <= node) ++k;
    }
    const size_t fs = p->block_count > 0 ? (size_t)p->block_ntheta[k] * p->block_nphi[k]
                                         : (size_t)p->ntheta * p->nphi;
    if (node_block) node_block[node] = k;
    if (node_offset) node_offset[node] = off;
    if (node_size) node_size[node] = fs;
    off += fs;
  }
  if (count) *count = blocks;
  if (values) *values = off;
}

int64_t sl3o_grid_values(const sl3o_problem* p) {
  size_t values = 0;
  grids_layout(p, NULL, NULL, NULL, NULL, &values);
  return (int64_t)values;
}

static void grids_build(grids_t* g, const sl3o_problem* p) {
  g->node_block = (int*)malloc(sizeof(int) * (p->nodes > 0 ? p->nodes : 1));
  g->node_offset = (size_t*)malloc(sizeof(size_t) * (p->nodes > 0 ? p->nodes : 1));
  g->node_size = (size_t*)malloc(sizeof(size_t) * (p->nodes > 0 ? p->nodes : 1));
  grids_layout(p, &g->count, g->node_block, g->node_offset, g->node_size, &g->values);
  g->tables = (tables_t*)calloc(g->count, sizeof(tables_t));
  g->wtheta = (const double**)calloc(g->count, sizeof(double*));
  for (int k = 0; k < g->count; ++k) {
    if (p->block_count > 0) {
      tables_build(&g->tables[k], p->block_lmax[k], p->block_ntheta[k], p->block_nphi[k],
                   p->block_theta[k]);
      g->wtheta[k] = p->block_wtheta[k];
    } else {
      tables_build(&g->tables[k], p->lmax, p->ntheta, p->nphi, p->theta);
      g->wtheta[k] = p->wtheta;
    }
  }
}

static void grids_free(grids_t* g) {
  for (int k = 0; k < g->count; ++k) tables_free(&g->tables[k]);
  free(g->tables);
  free(g->wtheta);
  free(g->node_block);
  free(g->node_offset);
  free(g->node_size);
}

static grids_t* g_grids = NULL;

static void synthesise_block(const tables_t* t, int b, const double* c, double* f) {
  const int L = t->lmax, nt = t->ntheta, nph = t->nphi, mlo = mlo_of(b, L);
  const int real = b == 2;
  double* F = (double*)malloc(sizeof(double) * 2 * (2 * L + 1));
  for (int i = 0; i < nt; ++i) {
    for (int m = mlo; m <= L; ++m) {
      double re = 0.0, im = 0.0;
      const int lmin = lmin_of(b, m);
      if (lmin <= L) {
        const double* col = t->d[b] + t->moff[b][m - mlo] + (int64_t)i * (L - lmin + 1);
        for (int l = lmin; l <= L; ++l) {
          const int64_t k = index_of(b, l, m);
          re += col[l - lmin] * c[2 * k];
          im += col[l - lmin] * c[2 * k + 1];
        }
      }
      F[2 * (m + L)] = re;
      F[2 * (m + L) + 1] = im;
    }
    for (int j = 0; j < nph; ++j) {
      const double* cs = t->cosT + (int64_t)j * (2 * L + 1) + L;
      const double* sn = t->sinT + (int64_t)j * (2 * L + 1) + L;
      double re = 0.0, im = 0.0;
      for (int m = mlo; m <= L; ++m) {
        const double fr = F[2 * (m + L)], fi = F[2 * (m + L) + 1];
        double tr = fr * cs[m] - fi * sn[m], ti = fr * sn[m] + fi * cs[m];
        if (real && m > 0) {
          tr *= 2.0;
          ti = 0.0;
        }
        re += tr;
        im += ti;
      }
      if (real) {
        f[(int64_t)i * nph + j] = re;
      } else {
        f[2 * ((int64_t)i * nph + j)] = re;
        f[2 * ((int64_t)i * nph + j) + 1] = im;
      }
    }
  }
  free(F);
}

static void analyse_block(const tables_t* t, int b, const double* wtheta, const double* f,
                          double* c) {
  const int L = t->lmax, nt = t->ntheta, nph = t->nphi, mlo = mlo_of(b, L);
  const int real = b == 2;
  memset(c, 0, sizeof(double) * 2 * block_size(b, L));
  for (int i = 0; i < nt; ++i) {
    for (int m = mlo; m <= L; ++m) {
      const int lmin = lmin_of(b, m);
      if (lmin > L) continue;
      double re = 0.0, im = 0.0;
      for (int j = 0; j < nph; ++j) {
        const double cs = t->cosT[(int64_t)j * (2 * L + 1) + m + L];
        const double sn = -t->sinT[(int64_t)j * (2 * L + 1) + m + L];
        if (real) {
          const double v = f[(int64_t)i * nph + j];
          re += v * cs;
          im += v * sn;
        } else {
          const double fr = f[2 * ((int64_t)i * nph + j)], fi = f[2 * ((int64_t)i * nph + j) + 1];
          re += fr * cs - fi * sn;
          im += fr * sn + fi * cs;
        }
      }
      const double w = 2.0 * M_PI / nph * wtheta[i];
      re *= w;
      im *= w;
      const double* col = t->d[b] + t->moff[b][m - mlo] + (int64_t)i * (L - lmin + 1);
      for (int l = lmin; l <= L; ++l) {
        const int64_t k = index_of(b, l, m);
        c[2 * k] += col[l - lmin] * re;
        c[2 * k + 1] += col[l - lmin] * im;
      }
    }
  }
}

/* Each node on its block's grid, its coefficients (a prefix of lmax's in
 * the reference's l-major order) to the block's degree. */
static void synthesise_tensor(const sl3o_problem* p, const double* s, double* const grid[3]) {
#pragma omp parallel for schedule(dynamic) collapse(2)
  for (int node = 0; node < p->nodes; ++node) {
    for (int b = 0; b < 3; ++b) {
      const double* c = s + block_offset(b, p->lmax, p->nodes) + 2 * node * block_size(b, p->lmax);
      double* f = grid[b] + (b == 2 ? 1 : 2) * g_grids->node_offset[node];
      synthesise_block(&g_grids->tables[g_grids->node_block[node]], b, c, f);
    }
  }
}

/* The coefficients above a node's block degree are zero. */
static void analyse_tensor(const sl3o_problem* p, double* const grid[3], double* s) {
#pragma omp parallel for schedule(dynamic) collapse(2)
  for (int node = 0; node < p->nodes; ++node) {
    for (int b = 0; b < 3; ++b) {
      double* c = s + block_offset(b, p->lmax, p->nodes) + 2 * node * block_size(b, p->lmax);
      const double* f = grid[b] + (b == 2 ? 1 : 2) * g_grids->node_offset[node];
      const int k = g_grids->node_block[node];
      memset(c, 0, sizeof(double) * 2 * block_size(b, p->lmax));
      analyse_block(&g_grids->tables[k], b, g_grids->wtheta[k], f, c);
    }
  }
}

/* ---- GridCoupling::Gather / Scatter (src/GridCoupling.cpp:595-633) ------- */

static void gather(const sl3o_problem* p, const double* s, int node, int l, int m, double A[10]) {
  const double sign = (m % 2 == 0) ? 1.0 : -1.0;
  double* w = (double*)s;
  memset(A, 0, sizeof(double) * 10);
  const double* r = coef_at(w, p->lmax, p->nodes, 2, node, l, m);
  A[4] = r[0];
  A[5] = r[1];
  if (l >= 1) {
    const double* a = coef_at(w, p->lmax, p->nodes, 1, node, l, m);
    const double* b = coef_at(w, p->lmax, p->nodes, 1, node, l, -m);
    A[2] = a[0];
    A[3] = a[1];
    A[6] = sign * b[0];
    A[7] = -sign * b[1];
  }
  if (l >= 2) {
    const double* a = coef_at(w, p->lmax, p->nodes, 0, node, l, m);
    const double* b = coef_at(w, p->lmax, p->nodes, 0, node, l, -m);
    A[0] = a[0];
    A[1] = a[1];
    A[8] = sign * b[0];
    A[9] = -sign * b[1];
  }
}

static void scatter(const sl3o_problem* p, double* s, int node, int l, int m, const double A[10]) {
  const double sign = (m % 2 == 0) ? 1.0 : -1.0;
  for (int c = 0; c < 2; ++c) {
    const int N = spin_of(c);
    if (l < abs(N)) continue;
    double* a = coef_at(s, p->lmax, p->nodes, c, node, l, m);
    a[0] += A[2 * c];
    a[1] += A[2 * c + 1];
    if (m > 0) {
      double* b = coef_at(s, p->lmax, p->nodes, c, node, l, -m);
      b[0] += sign * A[2 * (4 - c)];
      b[1] += -sign * A[2 * (4 - c) + 1];
    }
  }
  double* r = coef_at(s, p->lmax, p->nodes, 2, node, l, m);
  r[0] += A[4];
  r[1] += A[5];
}

static int first_node(const sl3o_problem* p, int first_element) {
  for (int node = 0; node < p->nodes; ++node) {
    if (p->elem0 + node / p->np >= first_element) return node;
  }
  return p->nodes;
}

/* ---- the degree systems --------------------------------------------------- */

typedef struct {
  int l, kind, ndof, kd, cholesky, pinned, first_element;
  double* scale;
  double* factor;
  int* piv;
  double* unit; /* spheroidal: the unit load response, then the unit tide's */
  const int* dofs;
} task_t;

/* src/DegreeOperator.cpp:775-852, src/ToroidalOperator.cpp:190-239 */
static int factorise(task_t* t, const double* band) {
  const int n = t->ndof, kd = t->kd, ld = kd + 1;
  t->scale = (double*)malloc(sizeof(double) * n);
  int positive = 1;
  for (int j = 0; j < n; ++j) {
    const double d = band[(size_t)j * ld + kd];
    positive = positive && d > 0.0;
    t->scale[j] = 1.0 / sqrt(fabs(d) > 0.0 ? fabs(d) : 1.0);
    if (t->kind == 1 && !(d > 0.0)) return 1;
  }
  t->piv = NULL;
  t->cholesky = 0;
  if (positive) {
    t->factor = (double*)calloc((size_t)ld * n, sizeof(double));
    for (int j = 0; j < n; ++j)
      for (int i = (j - kd > 0 ? j - kd : 0); i <= j; ++i)
        t->factor[(size_t)j * ld + kd + i - j] = band[(size_t)j * ld + kd + i - j] * t->scale[i] * t->scale[j];
    if (t->pinned >= 0) {
      const int pp = t->pinned;
      for (int j = (pp - kd > 0 ? pp - kd : 0); j <= (pp + kd < n - 1 ? pp + kd : n - 1); ++j) {
        const int i = pp < j ? pp : j, k = pp < j ? j : pp;
        t->factor[(size_t)k * ld + kd + i - k] = i == k ? 1.0 : 0.0;
      }
    }
    if (sl3o_band_cholesky_factor(t->factor, n, kd) == 0) {
      t->cholesky = 1;
      return 0;
    }
    free(t->factor);
    if (t->kind == 1) return 1;
  }
  const int ldlu = 3 * kd + 1;
  t->factor = (double*)calloc((size_t)ldlu * n, sizeof(double));
  for (int j = 0; j < n; ++j) {
    for (int i = (j - kd > 0 ? j - kd : 0); i <= (j + kd < n - 1 ? j + kd : n - 1); ++i) {
      const int r = i <= j ? i : j, c = i <= j ? j : i;
      t->factor[(size_t)j * ldlu + 2 * kd + i - j] =
          band[(size_t)c * ld + kd + r - c] * t->scale[r] * t->scale[c];
    }
  }
  t->piv = (int*)malloc(sizeof(int) * n);
  return sl3o_band_lu_factor(t->factor, n, kd, kd, t->piv) != 0;
}

/* BeginSolve + solve + FinishSolve on nrhs real columns (column-major). */
static void solve(const task_t* t, double* x, int nrhs) {
  const int n = t->ndof;
  for (int k = 0; k < nrhs; ++k) {
    for (int j = 0; j < n; ++j) x[(size_t)k * n + j] *= t->scale[j];
    if (t->pinned >= 0) x[(size_t)k * n + t->pinned] = 0.0;
  }
  if (t->cholesky) {
    sl3o_band_cholesky_solve(t->factor, n, t->kd, x, nrhs);
  } else {
    sl3o_band_lu_solve(t->factor, n, t->kd, t->kd, t->piv, x, nrhs);
  }
  for (int k = 0; k < nrhs; ++k)
    for (int j = 0; j < n; ++j) x[(size_t)k * n + j] *= t->scale[j];
}

/* ---- force (src/GridCoupling.cpp:691-809) and solves --------------------- */

/* The memory force of one task (src/GridCoupling.cpp:691-809): complex
 * columns m = 0..l, ndof each (column-major), into b (zeroed first). */
static void task_force(const sl3o_problem* p, const task_t* t, const double* weighted, double* b) {
  const int l = t->l, cols = l + 1, nd = t->ndof, ngll = p->ngll;
  memset(b, 0, sizeof(double) * 2 * (size_t)nd * cols);
  const int fn = first_node(p, t->first_element);
  for (int node = fn; node < p->nodes; ++node) {
    const int q = node % p->np;
    const int le = node / p->np;
    const double r = p->radius[node], jac = p->jac[le];
    for (int m = 0; m < cols; ++m) {
      double A[10], rows[12];
      gather(p, weighted, node, l, m, A);
      sl3o_force_rows(l, A, r, rows);
      const double point = jac * p->pweight[q] * r;
      for (int i = 0; i < ngll; ++i) {
        const double v = point * p->pvalue[q * ngll + i];
        const double deriv = p->pweight[q] * p->pderiv[q * ngll + i] * r;
        const int* d = t->dofs + (le * ngll + i) * 2;
        /* value rows then derivative rows: U (or W) and V */
        const int vrow[2] = {t->kind == 1 ? 4 : 0, 2};
        const int drow[2] = {t->kind == 1 ? 10 : 6, 8};
        for (int c = 0; c < (t->kind == 1 ? 1 : 2); ++c) {
          if (d[c] < 0) continue;
          double* z = b + 2 * ((size_t)m * nd + d[c]);
          z[0] += v * rows[vrow[c]] + deriv * rows[drow[c]];
          z[1] += v * rows[vrow[c] + 1] + deriv * rows[drow[c] + 1];
        }
      }
    }
  }
}

/* SolveForces (src/Viscoelastic3D.cpp:912-939) of one task's complex force
 * columns as [Re | Im] real columns, the channels added (AddUnitChannels,
 * :1029-1073: the load's unit response in column 0 of unit, the tide's in
 * column 1), into sol (complex, column-major). */
static void task_solve(const task_t* t, const double* b, const double* amp, const double* tamp,
                       double* sol) {
  const int l = t->l, cols = l + 1, nd = t->ndof;
  double* x = (double*)malloc(sizeof(double) * 2 * nd * cols);
  for (int c = 0; c < cols; ++c) {
    for (int d = 0; d < nd; ++d) {
      x[(size_t)c * nd + d] = b[2 * ((size_t)c * nd + d)];
      x[(size_t)(cols + c) * nd + d] = b[2 * ((size_t)c * nd + d) + 1];
    }
  }
  solve(t, x, 2 * cols);
  for (int c = 0; c < cols; ++c) {
    for (int d = 0; d < nd; ++d) {
      double re = x[(size_t)c * nd + d], im = x[(size_t)(cols + c) * nd + d];
      const size_t at = (size_t)l * (l + 1) / 2 + c;
      if (t->kind == 0 && amp != NULL) {
        re += amp[2 * at] * t->unit[d];
        im += amp[2 * at + 1] * t->unit[d];
      }
      if (t->kind == 0 && tamp != NULL) {
        re += tamp[2 * at] * t->unit[nd + d];
        im += tamp[2 * at + 1] * t->unit[nd + d];
      }
      sol[2 * ((size_t)c * nd + d)] = re;
      sol[2 * ((size_t)c * nd + d) + 1] = im;
    }
  }
  free(x);
}

static void force_and_solve(const sl3o_problem* p, task_t* tasks, int ntasks,
                            const double* weighted, const double* amp, const double* tamp,
                            double** solutions) {
#pragma omp parallel for schedule(dynamic)
  for (int ti = 0; ti < ntasks; ++ti) {
    task_t* t = &tasks[ti];
    double* b = (double*)malloc(sizeof(double) * 2 * (size_t)t->ndof * (t->l + 1));
    task_force(p, t, weighted, b);
    task_solve(t, b, amp, tamp, solutions[ti]);
    free(b);
  }
}

/* ---- strain (src/GridCoupling.cpp:635-689; Viscoelastic3D.cpp:832-865) --- */

static void strain_of(const sl3o_problem* p, const task_t* tasks, double** solutions, double* s) {
  memset(s, 0, sizeof(double) * sl3o_spectral_doubles(p->lmax, p->nodes));
  const int L = p->lmax, ngll = p->ngll;
#pragma omp parallel for schedule(dynamic)
  for (int i0 = 0; i0 <= L; ++i0) {
    const int l = L - i0;
    const task_t* ts = &tasks[l];
    const task_t* tt = l >= 1 ? &tasks[L + l] : NULL;
    const double* x = solutions[l];
    const double* w = l >= 1 ? solutions[L + l] : NULL;
    const int fn = first_node(p, ts->first_element);
    for (int node = fn; node < p->nodes; ++node) {
      const int e = p->elem0 + node / p->np, q = node % p->np, le = e - p->elem0;
      const double r = p->radius[node], jac = p->jac[le];
      for (int c = 0; c <= l; ++c) {
        double v[12] = {0};
        for (int i = 0; i < ngll; ++i) {
          const int* d = ts->dofs + (le * ngll + i) * 2;
          double U[2] = {0, 0}, V[2] = {0, 0}, W[2] = {0, 0};
          if (d[0] >= 0) {
            U[0] = x[2 * ((size_t)c * ts->ndof + d[0])];
            U[1] = x[2 * ((size_t)c * ts->ndof + d[0]) + 1];
          }
          if (d[1] >= 0) {
            V[0] = x[2 * ((size_t)c * ts->ndof + d[1])];
            V[1] = x[2 * ((size_t)c * ts->ndof + d[1]) + 1];
          }
          if (tt) {
            const int dw = tt->dofs[(le * ngll + i) * 2];
            if (dw >= 0) {
              W[0] = w[2 * ((size_t)c * tt->ndof + dw)];
              W[1] = w[2 * ((size_t)c * tt->ndof + dw) + 1];
            }
          }
          const double val = p->pvalue[q * ngll + i];
          for (int k = 0; k < 2; ++k) {
            v[0 + k] += val * U[k];
            v[2 + k] += val * V[k];
            v[4 + k] += val * W[k];
          }
          const double bb = r * p->pderiv[q * ngll + i] / jac;
          for (int k = 0; k < 2; ++k) {
            v[6 + k] += bb * U[k];
            v[8 + k] += bb * V[k];
            v[10 + k] += bb * W[k];
          }
        }
        double A[10];
        sl3o_deviatoric_components(l, v, r, A);
        scatter(p, s, node, l, c, A);
      }
    }
  }
}

/* ---- RelativeChange (src/Viscoelastic3D.cpp:1636-1669) ------------------- */

static double relative_change(const double* next, const double* now, int64_t n) {
  double scale = 0.0, diff = 0.0;
  int finite = 1;
  for (int64_t i = 0; i < n / 2; ++i) {
    const double nr = next[2 * i], ni = next[2 * i + 1];
    finite = finite && isfinite(nr) && isfinite(ni);
    const double a = nr * nr + ni * ni;
    const double dr = nr - now[2 * i], di = ni - now[2 * i + 1];
    const double b = dr * dr + di * di;
    if (a > scale) scale = a;
    if (b > diff) diff = b;
  }
  if (!finite) return NAN;
  return scale > 0.0 ? sqrt(diff) / sqrt(scale) : 0.0;
}

/* ---- the step -------------------------------------------------------------- */

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

int sl3o_step(const sl3o_problem* p, sl3o_state* st, double dt, int scheme,
              const double* amp, const double* tamp, const sl3o_options* o,
              sl3o_info* info) {
  const int L = p->lmax;
  const int ntasks = 2 * L + 1;
  const int64_t S = sl3o_spectral_doubles(L, p->nodes);
  int status = 0;
  double clock0 = now_ms();
  info->setup_ms = info->start_ms = info->passes_ms = info->finish_ms = 0.0;
  if (dt == 0.0) scheme = 1;
  if (tamp != NULL && p->tides == NULL) return 2;
  task_t* tasks = (task_t*)calloc(ntasks, sizeof(task_t));
  for (int ti = 0; ti < ntasks; ++ti) {
    task_t* t = &tasks[ti];
    t->kind = ti <= L ? 0 : 1;
    t->l = ti <= L ? ti : ti - L;
    t->ndof = p->ndof[ti];
    t->kd = p->kd[ti];
    t->first_element = p->first_element[ti];
    t->pinned = p->pinned[ti];
    t->dofs = p->dofs + p->dof_off[ti];
    if (factorise(t, p->bands + p->band_off[ti]) != 0) status = 1;
  }
  if (status != 0) goto done;
  grids_t grids;
  grids_build(&grids, p);
  g_grids = &grids;
  const size_t G = grids.values;
  for (int l = 0; l <= L; ++l) {
    task_t* t = &tasks[l];
    /* the unit load and tide solved together (:1042-1055) */
    t->unit = (double*)calloc((size_t)2 * t->ndof, sizeof(double));
    memcpy(t->unit, p->loads + p->load_off[l], sizeof(double) * t->ndof);
    if (p->tides != NULL) {
      memcpy(t->unit + t->ndof, p->tides + p->tide_off[l], sizeof(double) * t->ndof);
    }
    solve(t, t->unit, 2);
  }
  info->setup_ms = now_ms() - clock0;
  clock0 = now_ms();
  {
    double* s = (double*)calloc(S, sizeof(double));
    double* sNext = (double*)calloc(S, sizeof(double));
    double* fixedS = (double*)calloc(S, sizeof(double));
    double* analysed = (double*)calloc(S, sizeof(double));
    double* grid[3];
    double* d0[3];
    for (int b = 0; b < 3; ++b) {
      grid[b] = (double*)calloc((b == 2 ? 1 : 2) * G, sizeof(double));
      d0[b] = st->has_strain ? st->strain[b] : (double*)calloc((b == 2 ? 1 : 2) * G, sizeof(double));
    }
    double* beta = (double*)malloc(sizeof(double) * G);
    double* weight = (double*)malloc(sizeof(double) * G);
    double** sol = (double**)malloc(sizeof(double*) * ntasks);
    for (int ti = 0; ti < ntasks; ++ti) {
      sol[ti] = (double*)calloc((size_t)2 * tasks[ti].ndof * (tasks[ti].l + 1), sizeof(double));
    }
    /* s0 and d0 (:1252-1263) */
    if (st->has_coef) memcpy(s, st->coef, sizeof(double) * S);
    if (!st->has_strain) synthesise_tensor(p, s, d0);
    /* StartStep's sweep (:1283-1332), one branch */
    for (int node = 0; node < p->nodes; ++node) {
      const double mu = p->mu_node[node];
      const double href = dt * p->itau_ref_node[node];
      double refBeta;
      if (scheme == 1) {
        refBeta = 1.0 - sl3o_backward_euler_weight(href);
      } else {
        double e, a, bb;
        sl3o_trapezoid_weights(href, &e, &a, &bb);
        refBeta = bb;
      }
      const size_t x0 = grids.node_offset[node], x1 = x0 + grids.node_size[node];
      for (size_t x = x0; x < x1; ++x) {
        const double h = dt * p->itau[x];
        double ew, aw, bw;
        if (scheme == 1) {
          const double w = sl3o_backward_euler_weight(h);
          ew = w;
          aw = 0.0;
          bw = 1.0 - w;
        } else {
          sl3o_trapezoid_weights(h, &ew, &aw, &bw);
        }
        beta[x] = bw;
        weight[x] = 0.0 + mu * (bw - refBeta);
        for (int c = 0; c < 2; ++c) {
          for (int k = 0; k < 2; ++k) {
            double* mm = &st->m[c][2 * x + k];
            *mm = ew * *mm + aw * d0[c][2 * x + k];
            grid[c][2 * x + k] = 0.0 + mu * *mm;
          }
        }
        st->m[2][x] = ew * st->m[2][x] + aw * d0[2][x];
        grid[2][x] = 0.0 + mu * st->m[2][x];
      }
    }
    analyse_tensor(p, grid, fixedS);
    for (size_t x = 0; x < G; ++x) {
      for (int c = 0; c < 2; ++c) {
        grid[c][2 * x] = weight[x] * d0[c][2 * x];
        grid[c][2 * x + 1] = weight[x] * d0[c][2 * x + 1];
      }
      grid[2][x] = weight[x] * d0[2][x];
    }
    analyse_tensor(p, grid, analysed);
    for (int64_t i = 0; i < S; ++i) analysed[i] += fixedS[i];

    /* Converge (:1916-1974) */
    sl3o_mixer* mixer = NULL;
    const int adaptive = o->anderson_rate > 0.0;
    if (o->anderson_memory > 0 && !adaptive) mixer = sl3o_mixer_create(o->anderson_memory);
    double previous = 0.0;
    info->iterations = 0;
    info->residual = 0.0;
    info->converged = 0;
    info->mixed_passes = 0;
    g_residual_count = 0;
    info->start_ms = now_ms() - clock0;
    clock0 = now_ms();
    for (int it = 0; it < o->max_iterations; ++it) {
      if (it > 0) {
        synthesise_tensor(p, s, grid);
        for (size_t x = 0; x < G; ++x) {
          for (int c = 0; c < 2; ++c) {
            grid[c][2 * x] *= weight[x];
            grid[c][2 * x + 1] *= weight[x];
          }
          grid[2][x] *= weight[x];
        }
        analyse_tensor(p, grid, analysed);
        for (int64_t i = 0; i < S; ++i) analysed[i] += fixedS[i];
      }
      force_and_solve(p, tasks, ntasks, analysed, amp, tamp, sol);
      strain_of(p, tasks, sol, sNext);
      const double residual = relative_change(sNext, s, S);
      info->iterations = it + 1;
      info->residual = residual;
      if (g_residual_count < 4096) g_residuals[g_residual_count++] = residual;
      if (!isfinite(residual)) break;
      if (residual <= o->rtol) {
        info->converged = 1;
        memcpy(s, sNext, sizeof(double) * S);
        break;
      }
      if (o->damping != 1.0) {
        for (int64_t i = 0; i < S; ++i) sNext[i] = s[i] + o->damping * (sNext[i] - s[i]);
      }
      if (adaptive && !mixer && o->anderson_memory > 0 && it > 0 &&
          residual > o->anderson_rate * previous) {
        mixer = sl3o_mixer_create(o->anderson_memory);
      }
      previous = residual;
      if (mixer) {
        sl3o_mixer_update(mixer, (size_t)S, s, sNext);
        ++info->mixed_passes;
      } else {
        double* t = s;
        s = sNext;
        sNext = t;
      }
    }
    sl3o_mixer_destroy(mixer);
    info->passes_ms = now_ms() - clock0;
    clock0 = now_ms();
    /* Finish (:1900-1914) */
    synthesise_tensor(p, s, grid);
    for (size_t x = 0; x < G; ++x) {
      for (int c = 0; c < 2; ++c) {
        st->m[c][2 * x] += beta[x] * grid[c][2 * x];
        st->m[c][2 * x + 1] += beta[x] * grid[c][2 * x + 1];
      }
      st->m[2][x] += beta[x] * grid[2][x];
    }
    for (int b = 0; b < 3; ++b) {
      memcpy(st->strain[b], grid[b], sizeof(double) * (b == 2 ? 1 : 2) * G);
    }
    memcpy(st->coef, s, sizeof(double) * S);
    st->has_strain = 1;
    st->has_coef = 1;
    st->t += dt;
    info->finish_ms = now_ms() - clock0;
    for (int b = 0; b < 3; ++b) {
      free(grid[b]);
      if (d0[b] != st->strain[b]) free(d0[b]);
    }
    {
      int64_t* size = (int64_t*)malloc(sizeof(int64_t) * ntasks);
      for (int ti = 0; ti < ntasks; ++ti) size[ti] = (int64_t)2 * tasks[ti].ndof * (tasks[ti].l + 1);
      keep_solutions(sol, size, ntasks);
    }
    free(beta);
    free(weight);
    free(s);
    free(sNext);
    free(fixedS);
    free(analysed);
  }
  grids_free(&grids);
  g_grids = NULL;
done:
  for (int ti = 0; ti < ntasks; ++ti) {
    free(tasks[ti].scale);
    free(tasks[ti].factor);
    free(tasks[ti].piv);
    free(tasks[ti].unit);
  }
  free(tasks);
  return status;
}

/* ---- the pieces of a pass, one at a time --------------------------------- */

static void tasks_init(const sl3o_problem* p, task_t* tasks, int ntasks) {
  for (int ti = 0; ti < ntasks; ++ti) {
    task_t* t = &tasks[ti];
    memset(t, 0, sizeof(*t));
    t->kind = ti <= p->lmax ? 0 : 1;
    t->l = ti <= p->lmax ? ti : ti - p->lmax;
    t->ndof = p->ndof[ti];
    t->kd = p->kd[ti];
    t->first_element = p->first_element[ti];
    t->pinned = p->pinned[ti];
    t->dofs = p->dofs + p->dof_off[ti];
  }
}

static void tasks_free(task_t* tasks, int ntasks) {
  for (int ti = 0; ti < ntasks; ++ti) {
    free(tasks[ti].scale);
    free(tasks[ti].factor);
    free(tasks[ti].piv);
    free(tasks[ti].unit);
  }
  free(tasks);
}

int64_t sl3o_solution_doubles(const sl3o_problem* p) {
  int64_t n = 0;
  for (int ti = 0; ti < 2 * p->lmax + 1; ++ti) {
    const int l = ti <= p->lmax ? ti : ti - p->lmax;
    n += (int64_t)2 * p->ndof[ti] * (l + 1);
  }
  return n;
}

void sl3o_memory_force(const sl3o_problem* p, const double* weighted, double* forces) {
  const int ntasks = 2 * p->lmax + 1;
  task_t* tasks = (task_t*)calloc(ntasks, sizeof(task_t));
  tasks_init(p, tasks, ntasks);
  int64_t off = 0;
  for (int ti = 0; ti < ntasks; ++ti) {
    task_force(p, &tasks[ti], weighted, forces + off);
    off += (int64_t)2 * tasks[ti].ndof * (tasks[ti].l + 1);
  }
  tasks_free(tasks, ntasks);
}

void sl3o_strain_of(const sl3o_problem* p, const double* solutions, double* coef) {
  const int ntasks = 2 * p->lmax + 1;
  task_t* tasks = (task_t*)calloc(ntasks, sizeof(task_t));
  tasks_init(p, tasks, ntasks);
  double** sol = (double**)malloc(sizeof(double*) * ntasks);
  int64_t off = 0;
  for (int ti = 0; ti < ntasks; ++ti) {
    sol[ti] = (double*)solutions + off;
    off += (int64_t)2 * tasks[ti].ndof * (tasks[ti].l + 1);
  }
  strain_of(p, tasks, sol, coef);
  free(sol);
  tasks_free(tasks, ntasks);
}

void sl3o_synthesise_tensor(const sl3o_problem* p, const double* coef, double* const grid[3]) {
  grids_t grids;
  grids_build(&grids, p);
  g_grids = &grids;
  synthesise_tensor(p, coef, grid);
  g_grids = NULL;
  grids_free(&grids);
}

void sl3o_analyse_tensor(const sl3o_problem* p, double* const grid[3], double* coef) {
  grids_t grids;
  grids_build(&grids, p);
  g_grids = &grids;
  analyse_tensor(p, grid, coef);
  g_grids = NULL;
  grids_free(&grids);
}

/* ---- the degree-basis stepper --------------------------------------------- */
/* ViscoelasticOperator's 1D Maxwell step (src/Viscoelastic.cpp:914-1076:
 * StartStep, SolveElastic, the memory update), one branch, spheroidal
 * degrees only, at the memory points: the memory is held as the three
 * deviatoric scalars per (degree, order, point),
 *   M = V,  R = r U' - U + k2 V / 2,  S = U + r V' - V,
 * the strain the same scalars of the solution, and the memory force is the
 * point-shear form mu (k2 (k2 - 2) M M' + 4/3 R R' + k2 S S') with the test
 * function's scalars M', R', S' (the 1D state's layout:
 * tests/TestViscoelastic3D.cpp:154-170). No transform is involved: this is
 * the path the 3D step must reproduce when delta = 0. */

int64_t sl3o_scalar_doubles(const sl3o_problem* p) {
  int64_t n = 0;
  for (int l = 0; l <= p->lmax; ++l) n += (int64_t)p->nodes * (l + 1) * 3 * 2;
  return n;
}

static int64_t scalar_offset(const sl3o_problem* p, int l) {
  int64_t n = 0;
  for (int k = 0; k < l; ++k) n += (int64_t)p->nodes * (k + 1) * 3 * 2;
  return n;
}

/* The scalars of basis function i of a node's element at memory point q:
 * coefficients of (M, R, S) for its U and for its V. */
static void scalar_rows(const sl3o_problem* p, int q, int le, int i, double k2, double r,
                        double rowU[3], double rowV[3]) {
  const double v = p->pvalue[q * p->ngll + i];
  const double b = r * p->pderiv[q * p->ngll + i] / p->jac[le];
  rowU[0] = 0.0;
  rowU[1] = b - v;
  rowU[2] = v;
  rowV[0] = v;
  rowV[1] = 0.5 * k2 * v;
  rowV[2] = b - v;
}

int sl3o_step_degree_basis(const sl3o_problem* p, sl3o_state1d* st, double dt, int scheme,
                           const double* amp, double* solutions) {
  const int L = p->lmax;
  task_t* tasks = (task_t*)calloc(L + 1, sizeof(task_t));
  tasks_init(p, tasks, L + 1);
  int status = 0;
  for (int l = 0; l <= L; ++l) {
    if (factorise(&tasks[l], p->bands + p->band_off[l]) != 0) status = 1;
  }
  if (status != 0) {
    tasks_free(tasks, L + 1);
    return status;
  }
  if (dt == 0.0) scheme = 1;
  int64_t soff = 0;
  for (int l = 0; l <= L; ++l) {
    task_t* t = &tasks[l];
    const int cols = l + 1, nd = t->ndof, ngll = p->ngll;
    const double k2 = l * (l + 1.0);
    const double D[3] = {k2 * (k2 - 2.0), 4.0 / 3.0, k2};
    double* m = st->m + scalar_offset(p, l);
    double* d = st->d + scalar_offset(p, l);
    t->unit = (double*)calloc((size_t)2 * nd, sizeof(double));
    memcpy(t->unit, p->loads + p->load_off[l], sizeof(double) * nd);
    solve(t, t->unit, 2);
    const int fn = first_node(p, t->first_element);
    double* beta = (double*)calloc(p->nodes, sizeof(double));
    double* b = (double*)calloc((size_t)2 * nd * cols, sizeof(double));
    /* StartStep: m~ = e m + alpha d0; the force of mu m~ */
    for (int node = fn; node < p->nodes; ++node) {
      const double h = dt * p->itau_ref_node[node];
      double e, a, bb;
      if (scheme == 1) {
        e = sl3o_backward_euler_weight(h);
        a = 0.0;
        bb = 1.0 - e;
      } else {
        sl3o_trapezoid_weights(h, &e, &a, &bb);
      }
      beta[node] = bb;
      const int q = node % p->np, le = node / p->np;
      const double r = p->radius[node];
      const double w = p->jac[le] * p->pweight[q] * p->mu_node[node];
      for (int c = 0; c < cols; ++c) {
        double* mm = m + ((int64_t)node * cols + c) * 6;
        const double* d0 = d + ((int64_t)node * cols + c) * 6;
        for (int k = 0; k < 6; ++k) mm[k] = e * mm[k] + (st->has_strain ? a * d0[k] : 0.0);
        for (int i = 0; i < ngll; ++i) {
          double rowU[3], rowV[3];
          scalar_rows(p, q, le, i, k2, r, rowU, rowV);
          const int* dof = t->dofs + (le * ngll + i) * 2;
          for (int part = 0; part < 2; ++part) {
            double fu = 0.0, fv = 0.0;
            for (int k = 0; k < 3; ++k) {
              fu += rowU[k] * D[k] * mm[2 * k + part];
              fv += rowV[k] * D[k] * mm[2 * k + part];
            }
            if (dof[0] >= 0) b[2 * ((size_t)c * nd + dof[0]) + part] += w * fu;
            if (dof[1] >= 0) b[2 * ((size_t)c * nd + dof[1]) + part] += w * fv;
          }
        }
      }
    }
    double* x = solutions + soff;
    task_solve(t, b, amp, NULL, x);
    soff += (int64_t)2 * nd * cols;
    /* the strain scalars and the memory update m = m~ + beta d */
    for (int node = fn; node < p->nodes; ++node) {
      const int q = node % p->np, le = node / p->np;
      const double r = p->radius[node];
      for (int c = 0; c < cols; ++c) {
        double* dd = d + ((int64_t)node * cols + c) * 6;
        double* mm = m + ((int64_t)node * cols + c) * 6;
        for (int k = 0; k < 6; ++k) dd[k] = 0.0;
        for (int i = 0; i < ngll; ++i) {
          double rowU[3], rowV[3];
          scalar_rows(p, q, le, i, k2, r, rowU, rowV);
          const int* dof = t->dofs + (le * ngll + i) * 2;
          for (int part = 0; part < 2; ++part) {
            const double U = dof[0] >= 0 ? x[2 * ((size_t)c * nd + dof[0]) + part] : 0.0;
            const double V = dof[1] >= 0 ? x[2 * ((size_t)c * nd + dof[1]) + part] : 0.0;
            for (int k = 0; k < 3; ++k) dd[2 * k + part] += rowU[k] * U + rowV[k] * V;
          }
        }
        for (int k = 0; k < 6; ++k) mm[k] += beta[node] * dd[k];
      }
    }
    free(b);
    free(beta);
  }
  st->has_strain = 1;
  st->t += dt;
  tasks_free(tasks, L + 1);
  return 0;
}
