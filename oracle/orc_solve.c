/*
 * orc_solve.c -- oracle Krylov solve for the pressure-Poisson system:
 * CSR SpMV, restarted right-preconditioned GMRES(m) with classical
 * Gram-Schmidt plus one reorthogonalisation, Jacobi and Jacobi-scaled
 * Chebyshev right preconditioners, and a dense LU reference solve.
 * Test infrastructure only; see orc.h.
 *
 * The reference delegates this to Belos/Ifpack2 (configured in
 * src/LinearSolverConfig.C:37-140, driven by src/LinearSolver.C:209-254):
 * GMRES with "Num Blocks" = kspace, ICGS orthogonalisation, right
 * preconditioning, x0 = 0 and convergence on the residual scaled by the
 * initial residual norm.  This file restates that published algorithm.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_spmv(const orc_csr* A, const double* x, double* y)
{
  for (int64_t r = 0; r < A->nrows; ++r) {
    double s = 0.0;
    for (int64_t k = A->row_ptr[r]; k < A->row_ptr[r + 1]; ++k)
      s += A->vals[k] * x[A->cols[k]];
    y[r] = s;
  }
}

static double dot(int64_t n, const double* a, const double* b)
{
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static void diag_of(const orc_csr* A, double* d)
{
  for (int64_t r = 0; r < A->nrows; ++r) {
    d[r] = 0.0;
    for (int64_t k = A->row_ptr[r]; k < A->row_ptr[r + 1]; ++k)
      if (A->cols[k] == r) d[r] = A->vals[k];
  }
}

static uint64_t splitmix(uint64_t* s)
{
  uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

double orc_jacobi_lambda_max(const orc_csr* A, int iters, uint64_t seed)
{
  const int64_t n = A->nrows;
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  double* x = (double*)malloc(sizeof(double) * (size_t)n);
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  diag_of(A, d);
  for (int64_t i = 0; i < n; ++i)
    x[i] = (double)(splitmix(&seed) >> 11) * (1.0 / 9007199254740992.0) + 0.5;
  double lam = 0.0;
  for (int it = 0; it < iters; ++it) {
    const double nx = sqrt(dot(n, x, x));
    for (int64_t i = 0; i < n; ++i) x[i] /= nx;
    orc_spmv(A, x, y);
    for (int64_t i = 0; i < n; ++i) y[i] /= d[i];
    lam = dot(n, x, y);
    memcpy(x, y, sizeof(double) * (size_t)n);
  }
  free(d);
  free(x);
  free(y);
  return lam;
}

typedef struct {
  const orc_csr* A;
  const orc_gmres_opts* o;
  double* dinv;
  double lmax, lmin;
  double *r, *d; /* Chebyshev work */
} prec_t;

/* z = M^-1 v */
static void prec_apply(prec_t* P, const double* v, double* z)
{
  const int64_t n = P->A->nrows;
  if (P->o->precond == ORC_PREC_NONE) {
    memcpy(z, v, sizeof(double) * (size_t)n);
    return;
  }
  if (P->o->precond == ORC_PREC_JACOBI) {
    for (int64_t i = 0; i < n; ++i) z[i] = P->dinv[i] * v[i];
    return;
  }
  /* Chebyshev iteration on D^-1 A x = D^-1 v, x0 = 0, degree k: the
   * three-term recurrence of Saad, Iterative Methods for Sparse Linear
   * Systems, 2nd ed., Algorithm 12.1, on [lmin, lmax]. */
  const double theta = 0.5 * (P->lmax + P->lmin);
  const double delta = 0.5 * (P->lmax - P->lmin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  for (int64_t i = 0; i < n; ++i) {
    P->d[i] = P->dinv[i] * v[i] / theta;
    z[i] = P->d[i];
  }
  for (int k = 1; k < P->o->cheb_degree; ++k) {
    orc_spmv(P->A, z, P->r);
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    const double c1 = rho_new * rho, c2 = 2.0 * rho_new / delta;
    for (int64_t i = 0; i < n; ++i) {
      P->d[i] = c1 * P->d[i] + c2 * P->dinv[i] * (v[i] - P->r[i]);
      z[i] += P->d[i];
    }
    rho = rho_new;
  }
}

int orc_gmres(const orc_csr* A, const double* b, double* x,
              const orc_gmres_opts* o, orc_gmres_stats* st)
{
  const int64_t n = A->nrows;
  const int m = o->kspace;
  double* V = (double*)malloc(sizeof(double) * (size_t)n * (size_t)(m + 1));
  double* Z = (double*)malloc(sizeof(double) * (size_t)n);
  double* w = (double*)malloc(sizeof(double) * (size_t)n);
  double* H = (double*)calloc((size_t)(m + 1) * (size_t)m, sizeof(double));
  double* cs = (double*)malloc(sizeof(double) * (size_t)m);
  double* sn = (double*)malloc(sizeof(double) * (size_t)m);
  double* g = (double*)malloc(sizeof(double) * (size_t)(m + 1));
  double* y = (double*)malloc(sizeof(double) * (size_t)m);
  double* h = (double*)malloc(sizeof(double) * (size_t)(m + 1));
  prec_t P = {A, o, NULL, 0, 0, NULL, NULL};
  P.dinv = (double*)malloc(sizeof(double) * (size_t)n);
  diag_of(A, P.dinv);
  for (int64_t i = 0; i < n; ++i) P.dinv[i] = 1.0 / P.dinv[i];
  if (o->precond == ORC_PREC_CHEBYSHEV) {
    P.r = (double*)malloc(sizeof(double) * (size_t)n);
    P.d = (double*)malloc(sizeof(double) * (size_t)n);
    P.lmax = orc_jacobi_lambda_max(A, 30, 12345) * 1.1;
    P.lmin = P.lmax / o->cheb_ratio;
  }

  const double bnorm = sqrt(dot(n, b, b));
  st->bnorm = bnorm;
  st->iters = 0;
  st->converged = 0;
  double rel = 1.0;
  const double scale = bnorm > 0 ? bnorm : 1.0;
  while (1) {
    orc_spmv(A, x, w);
    for (int64_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
    const double beta = sqrt(dot(n, w, w));
    rel = beta / scale;
    if (rel <= o->tol) {
      st->converged = 1;
      break;
    }
    if (st->iters >= o->max_iters) break;
    for (int64_t i = 0; i < n; ++i) V[i] = w[i] / beta;
    memset(g, 0, sizeof(double) * (size_t)(m + 1));
    g[0] = beta;
    int j = 0;
    for (; j < m && st->iters < o->max_iters; ++j) {
      prec_apply(&P, V + (int64_t)j * n, Z);
      orc_spmv(A, Z, w);
      /* CGS with one reorthogonalisation pass */
      for (int k = 0; k <= j; ++k) h[k] = 0.0;
      for (int pass = 0; pass < 2; ++pass) {
        for (int k = 0; k <= j; ++k) {
          H[(int64_t)k * m + j] = dot(n, V + (int64_t)k * n, w);
        }
        for (int k = 0; k <= j; ++k) {
          const double c = H[(int64_t)k * m + j];
          const double* vk = V + (int64_t)k * n;
          for (int64_t i = 0; i < n; ++i) w[i] -= c * vk[i];
          h[k] += c;
        }
      }
      for (int k = 0; k <= j; ++k) H[(int64_t)k * m + j] = h[k];
      const double hn = sqrt(dot(n, w, w));
      H[(int64_t)(j + 1) * m + j] = hn;
      for (int64_t i = 0; i < n; ++i)
        V[(int64_t)(j + 1) * n + i] = hn > 0 ? w[i] / hn : 0.0;
      /* Givens rotations */
      for (int k = 0; k < j; ++k) {
        const double a = H[(int64_t)k * m + j], c = H[(int64_t)(k + 1) * m + j];
        H[(int64_t)k * m + j] = cs[k] * a + sn[k] * c;
        H[(int64_t)(k + 1) * m + j] = -sn[k] * a + cs[k] * c;
      }
      {
        const double a = H[(int64_t)j * m + j], c = H[(int64_t)(j + 1) * m + j];
        const double r = sqrt(a * a + c * c);
        cs[j] = r > 0 ? a / r : 1.0;
        sn[j] = r > 0 ? c / r : 0.0;
        H[(int64_t)j * m + j] = r;
        H[(int64_t)(j + 1) * m + j] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
      }
      st->iters++;
      st->implicit_rel = fabs(g[j + 1]) / scale;
      if (st->implicit_rel <= o->tol || hn == 0.0) {
        ++j;
        break;
      }
    }
    /* back substitution, x += M^-1 V y */
    for (int k = j - 1; k >= 0; --k) {
      double s = g[k];
      for (int l = k + 1; l < j; ++l) s -= H[(int64_t)k * m + l] * y[l];
      y[k] = s / H[(int64_t)k * m + k];
    }
    memset(w, 0, sizeof(double) * (size_t)n);
    for (int k = 0; k < j; ++k) {
      const double* vk = V + (int64_t)k * n;
      for (int64_t i = 0; i < n; ++i) w[i] += y[k] * vk[i];
    }
    prec_apply(&P, w, Z);
    for (int64_t i = 0; i < n; ++i) x[i] += Z[i];
  }
  orc_spmv(A, x, w);
  for (int64_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
  st->final_abs = sqrt(dot(n, w, w));
  free(V);
  free(Z);
  free(w);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(y);
  free(h);
  free(P.dinv);
  free(P.r);
  free(P.d);
  return st->converged ? 0 : 1;
}

int orc_dense_solve(int n, double* A, double* b)
{
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[(int64_t)i * n + k]) > fabs(A[(int64_t)p * n + k])) p = i;
    if (A[(int64_t)p * n + k] == 0.0) return -1;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        const double t = A[(int64_t)k * n + j];
        A[(int64_t)k * n + j] = A[(int64_t)p * n + j];
        A[(int64_t)p * n + j] = t;
      }
      const double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[(int64_t)i * n + k] / A[(int64_t)k * n + k];
      for (int j = k; j < n; ++j) A[(int64_t)i * n + j] -= f * A[(int64_t)k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A[(int64_t)k * n + j] * b[j];
    b[k] = s / A[(int64_t)k * n + k];
  }
  return 0;
}
