/*
 * orc_edge.c -- oracle restatement of the three interior edge solver
 * algorithms, the Peclet blending functions, the Tpetra-style CSR graph and
 * sum_into scatter, the row-reset / Dirichlet / fix-node operations and the
 * serial edge-loop assembly drivers.  Test infrastructure only; see orc.h.
 */
#include "orc.h"

/* Threads of the CSR assembly loops (orc_set_threads); above 1 the loops
 * run in parallel over edges and the scatter adds are atomic, as the
 * reference's Kokkos-OpenMP build does (src/TpetraLinearSystem.C:1304-1305). */
static int g_threads = 1;

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

#define ORC_ADD(p, v)                                   \
  do {                                                  \
    if (g_threads > 1) {                                \
      _Pragma("omp atomic")(p) += (v);                  \
    } else {                                            \
      (p) += (v);                                       \
    }                                                   \
  } while (0)


#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- */
/* Peclet functions and limiter                                      */
/* ---------------------------------------------------------------- */

double orc_peclet_classic(double hf, double pe)
{
  /* src/PecletFunction.C:41-45 (A is carried but unused by execute) */
  const double m = hf * pe;
  return m * m / (5.0 + m * m);
}

double orc_peclet_tanh(double c1, double c2, double x)
{
  /* src/PecletFunction.C:68-71 */
  return 0.50 * (1.0 + tanh((x - c1) / c2));
}

double orc_van_leer(double dqm, double dqp, double eps)
{
  /* include/edge_kernels/EdgeKernelUtils.h:18-24 */
  return (2.0 * (dqm * dqp + fabs(dqm * dqp))) /
         ((dqm + dqp) * (dqm + dqp) + eps);
}

double orc_peclet_eval(const orc_peclet* p, double pe)
{
  return p->form == 0 ? orc_peclet_classic(p->hybrid, pe)
                      : orc_peclet_tanh(p->c1, p->c2, pe);
}

void orc_momentum_edge_peclet(int64_t nedge, const int32_t* edges,
                              const double* coords, const double* vel,
                              const double* rho, const double* visc,
                              const orc_peclet* pf, double* pecnum,
                              double* pecfac)
{
  /* src/edge_kernels/MomentumEdgePecletAlg.C:74-101, eps_ = 1e-16 (.h:40) */
  const double eps = 1.0e-16;
  /* independent edges: over orc_set_threads threads, like the assembly */
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int64_t e = 0; e < nedge; ++e) {
    const int32_t L = edges[2 * e], R = edges[2 * e + 1];
    const double diffIp = 0.5 * (visc[L] / rho[L] + visc[R] / rho[R]);
    double udotx = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double dxj = coords[3 * R + d] - coords[3 * L + d];
      udotx += 0.5 * dxj * (vel[3 * R + d] + vel[3 * L + d]);
    }
    const double pn = fabs(udotx) / (diffIp + eps);
    if (pecnum) pecnum[e] = pn;
    pecfac[e] = orc_peclet_eval(pf, pn);
  }
}

/* ---------------------------------------------------------------- */
/* Per-edge blocks                                                   */
/* ---------------------------------------------------------------- */

void orc_scalar_edge(const orc_scalar_opts* o, int64_t e, int32_t nL,
                     int32_t nR, const double* coords, const double* vrtm,
                     const double* q, const double* dqdx, const double* rho,
                     const double* dflux, const double* areav,
                     const double* mdotf, double lhs[4], double rhs[2])
{
  /* src/edge_kernels/ScalarEdgeSolverAlg.C:56-205 */
  const double eps = 1.0e-16;
  const double alpha = o->alpha, alphaUpw = o->alpha_upw;
  const double hoUpwind = o->upw, relaxFac = o->relax;
  const double om_alpha = 1.0 - alpha, om_alphaUpw = 1.0 - alphaUpw;

  double av[3];
  for (int d = 0; d < 3; ++d) av[d] = areav[3 * e + d];
  const double mdot = mdotf[e];
  const double densityL = rho[nL], densityR = rho[nR];
  const double qNp1L = q[nL], qNp1R = q[nR];
  const double viscosityL = dflux[nL], viscosityR = dflux[nR];
  const double viscIp = 0.5 * (viscosityL + viscosityR);
  const double diffIp = 0.5 * (viscosityL / densityL + viscosityR / densityR);

  double axdx = 0.0, asq = 0.0, udotx = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double dxj = coords[3 * nR + d] - coords[3 * nL + d];
    asq += av[d] * av[d];
    axdx += av[d] * dxj;
    udotx += 0.5 * dxj * (vrtm[3 * nR + d] + vrtm[3 * nL + d]);
  }
  const double inv_axdx = 1.0 / axdx;

  double dqL = 0.0, dqR = 0.0, nonOrth = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double dxj = (coords[3 * nR + d] - coords[3 * nL + d]);
    dqL += 0.5 * dxj * dqdx[3 * nL + d];
    dqR += 0.5 * dxj * dqdx[3 * nR + d];
    const double kxj = av[d] - asq * inv_axdx * dxj;
    nonOrth += -viscIp * kxj * 0.5 * (dqdx[3 * nR + d] + dqdx[3 * nL + d]);
  }

  const double pecnum = fabs(udotx) / (diffIp + eps);
  const double pecfac = orc_peclet_eval(&o->peclet, pecnum);
  const double om_pecfac = 1.0 - pecfac;

  double limitL = 1.0, limitR = 1.0;
  if (o->use_limiter) {
    const double dq = qNp1R - qNp1L;
    const double dqML = 4.0 * dqL - dq;
    const double dqMR = 4.0 * dqR - dq;
    limitL = orc_van_leer(dqML, dq, eps);
    limitR = orc_van_leer(dqMR, dq, eps);
  }

  const double qIpL = qNp1L + dqL * hoUpwind * limitL;
  const double qIpR = qNp1R - dqR * hoUpwind * limitR;

  const double lhsfac = -viscIp * asq * inv_axdx;
  const double diffFlux = lhsfac * (qNp1R - qNp1L) + nonOrth;

  lhs[0] = -lhsfac / relaxFac;
  lhs[1] = lhsfac;
  rhs[0] = -diffFlux;
  lhs[2] = lhsfac;
  lhs[3] = -lhsfac / relaxFac;
  rhs[1] = diffFlux;

  const double qIp = 0.5 * (qNp1R + qNp1L);
  const double qUpw = (mdot > 0) ? (alphaUpw * qIpL + om_alphaUpw * qIp)
                                 : (alphaUpw * qIpR + om_alphaUpw * qIp);
  const double qHatL = (alpha * qIpL + om_alpha * qIp);
  const double qHatR = (alpha * qIpR + om_alpha * qIp);
  const double qCds = 0.5 * (qHatL + qHatR);

  const double adv_flux = mdot * (pecfac * qUpw + om_pecfac * qCds);
  rhs[0] -= adv_flux;
  rhs[1] += adv_flux;

  double alhsfac = 0.5 * (mdot + fabs(mdot)) * pecfac * alphaUpw +
                   0.5 * alpha * om_pecfac * mdot;
  lhs[0] += alhsfac / relaxFac;
  lhs[2] -= alhsfac;

  alhsfac = 0.5 * (mdot - fabs(mdot)) * pecfac * alphaUpw +
            0.5 * alpha * om_pecfac * mdot;
  lhs[3] -= alhsfac / relaxFac;
  lhs[1] += alhsfac;

  alhsfac = 0.5 * mdot * (pecfac * om_alphaUpw + om_pecfac * om_alpha);
  lhs[0] += alhsfac / relaxFac;
  lhs[1] += alhsfac;
  lhs[2] -= alhsfac;
  lhs[3] -= alhsfac / relaxFac;
}

void orc_momentum_edge(const orc_momentum_opts* o, int64_t e, int32_t nL,
                       int32_t nR, const double* coords, const double* vel,
                       const double* dudx, const double* visc,
                       const double* rho, const double* areav,
                       const double* mdotf, const double* pecfacf,
                       const double* wallmask, double lhs[36], double rhs[6])
{
  /* src/edge_kernels/MomentumEdgeSolverAlg.C:71-312 (has_vof = 0) */
  const double eps = 1.0e-16;
  const double includeDivU = o->include_divU;
  const double alpha = o->alpha, alphaUpw = o->alpha_upw;
  const double hoUpwind = o->upw, relaxFacU = o->relax;
  const double om_alpha = 1.0 - alpha, om_alphaUpw = 1.0 - alphaUpw;
  (void)rho;
  memset(lhs, 0, sizeof(double) * 36);
  memset(rhs, 0, sizeof(double) * 6);
#define LHS(i, j) lhs[6 * (i) + (j)]
#define X(n, d) coords[3 * (n) + (d)]
#define U(n, d) vel[3 * (n) + (d)]
#define G(n, k) dudx[9 * (int64_t)(n) + (k)]

  double av[3];
  for (int d = 0; d < 3; ++d) av[d] = areav[3 * e + d];
  const double mdot = mdotf[e];
  const double viscIp = 0.5 * (visc[nL] + visc[nR]);

  double axdx = 0.0, asq = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double dxj = X(nR, d) - X(nL, d);
    asq += av[d] * av[d];
    axdx += av[d] * dxj;
  }
  const double inv_axdx = 1.0 / axdx;

  double duL[3], duR[3];
  for (int i = 0; i < 3; ++i) {
    const int offset = i * 3;
    duL[i] = 0.0;
    duR[i] = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double dxj = 0.5 * (X(nR, j) - X(nL, j));
      duL[i] += dxj * G(nL, offset + j);
      duR[i] += dxj * G(nR, offset + j);
    }
  }

  double limitL[3] = {1.0, 1.0, 1.0}, limitR[3] = {1.0, 1.0, 1.0};
  if (o->use_limiter) {
    for (int d = 0; d < 3; ++d) {
      const double du = U(nR, d) - U(nL, d);
      const double duML = 4.0 * duL[d] - du;
      const double duMR = 4.0 * duR[d] - du;
      limitL[d] = orc_van_leer(duML, du, eps);
      limitR[d] = orc_van_leer(duMR, du, eps);
    }
  }

  const double pecfac = pecfacf[e];
  const double om_pecfac = 1.0 - pecfac;
  const double density_upwinding_factor = 1.0;

  double uIpL[3], uIpR[3];
  for (int d = 0; d < 3; ++d) {
    uIpL[d] = U(nL, d) + duL[d] * hoUpwind * limitL[d] * density_upwinding_factor;
    uIpR[d] = U(nR, d) - duR[d] * hoUpwind * limitR[d] * density_upwinding_factor;
  }

  double duidxj[3][3];
  for (int i = 0; i < 3; ++i) {
    const double dui = U(nR, i) - U(nL, i);
    const int offset = i * 3;
    double gjuidx = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double dxj = X(nR, j) - X(nL, j);
      const double gjui = 0.5 * (G(nR, offset + j) + G(nL, offset + j));
      gjuidx += gjui * dxj;
    }
    for (int j = 0; j < 3; ++j) {
      const double gjui = 0.5 * (G(nR, offset + j) + G(nL, offset + j));
      duidxj[i][j] = gjui + (dui - gjuidx) * av[j] * inv_axdx;
    }
  }

  const double dlhsfac = -viscIp * asq * inv_axdx;

  for (int i = 0; i < 3; ++i) {
    const int rowL = i, rowR = i + 3;
    const double uiIp = 0.5 * (U(nR, i) + U(nL, i));
    const double uiUpw = (mdot > 0.0) ? (alphaUpw * uIpL[i] + om_alphaUpw * uiIp)
                                      : (alphaUpw * uIpR[i] + om_alphaUpw * uiIp);
    const double uiHatL = (alpha * uIpL[i] + om_alpha * uiIp);
    const double uiHatR = (alpha * uIpR[i] + om_alpha * uiIp);
    const double uiCds = 0.5 * (uiHatL + uiHatR);
    const double adv_flux = mdot * (pecfac * uiUpw + om_pecfac * uiCds);

    double diff_flux = 0.0;
    for (int j = 0; j < 3; ++j) diff_flux += duidxj[j][j];
    diff_flux *= 2.0 / 3.0 * viscIp * av[i] * includeDivU;
    for (int j = 0; j < 3; ++j)
      diff_flux += -viscIp * (duidxj[i][j] + duidxj[j][i]) * av[j];

    const double mL = wallmask[nL], mR = wallmask[nR];
    const double maskNode = mL < mR ? mL : mR;
    const double total_flux = adv_flux + diff_flux * maskNode;

    rhs[rowL] -= total_flux;
    rhs[rowR] += total_flux;

    double alhsfac = 0.5 * (mdot + fabs(mdot)) * pecfac * alphaUpw +
                     0.5 * alpha * om_pecfac * mdot;
    LHS(rowL, rowL) += alhsfac / relaxFacU;
    LHS(rowR, rowL) -= alhsfac;

    alhsfac = 0.5 * (mdot - fabs(mdot)) * pecfac * alphaUpw +
              0.5 * alpha * om_pecfac * mdot;
    LHS(rowR, rowR) -= alhsfac / relaxFacU;
    LHS(rowL, rowR) += alhsfac;

    alhsfac = 0.5 * mdot * (pecfac * om_alphaUpw + om_pecfac * om_alpha);
    LHS(rowL, rowL) += alhsfac / relaxFacU;
    LHS(rowL, rowR) += alhsfac;
    LHS(rowR, rowL) -= alhsfac;
    LHS(rowR, rowR) -= alhsfac / relaxFacU;

    LHS(rowL, rowL) -= dlhsfac / relaxFacU;
    LHS(rowL, rowR) += dlhsfac;
    LHS(rowR, rowL) += dlhsfac;
    LHS(rowR, rowR) -= dlhsfac / relaxFacU;

    for (int j = 0; j < 3; ++j) {
      const double lhsfacNS = -viscIp * av[i] * av[j] * inv_axdx;
      const int colL = j, colR = j + 3;
      LHS(rowL, colL) -= lhsfacNS / relaxFacU;
      LHS(rowL, colR) += lhsfacNS;
      LHS(rowR, colL) += lhsfacNS;
      LHS(rowR, colR) -= lhsfacNS / relaxFacU;
    }
  }
#undef LHS
#undef X
#undef U
#undef G
}

void orc_continuity_edge(const orc_continuity_opts* o, int64_t e,
                         int32_t nL, int32_t nR, const double* coords,
                         const double* vel, const double* dpdx,
                         const double* rho, const double* p,
                         const double* udiag, const double* areav,
                         double lhs[4], double rhs[2])
{
  /* src/edge_kernels/ContinuityEdgeSolverAlg.C:38-194 */
  const double nocFac = o->noc;
  const double tauScale = o->dt / o->gamma1;
  const double interpTogether = o->interp_together;
  const double om_interpTogether = 1.0 - interpTogether;
  const double solveIncompressibleEqn = o->incompressible;
  const double om_solveIncompressibleEqn = 1.0 - solveIncompressibleEqn;

  double av[3];
  for (int d = 0; d < 3; ++d) av[d] = areav[3 * e + d];
  const double pressureL = p[nL], pressureR = p[nR];
  const double densityL = rho[nL], densityR = rho[nR];
  const double udiagL = udiag[nL], udiagR = udiag[nR];
  const double projTimeScale = 0.5 * (1.0 / udiagL + 1.0 / udiagR);
  const double rhoIp = 0.5 * (densityL + densityR);
  const double denScale =
    (1.0 / rhoIp) * solveIncompressibleEqn + om_solveIncompressibleEqn;

  double axdx = 0.0, asq = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double dxj = coords[3 * nR + d] - coords[3 * nL + d];
    asq += av[d] * av[d];
    axdx += av[d] * dxj;
  }
  const double inv_axdx = 1.0 / axdx;

  double tmdot = -projTimeScale * (pressureR - pressureL) * asq * inv_axdx;

  /* ContinuityEdgeSolverAlg.C:147-152 */
  const orc_buoyancy* bf = o->buoy;
  if (bf) {
    const double masked_weights = 0.5 * (bf->mask[nL] + bf->mask[nR]);
    for (int d = 0; d < 3; ++d)
      tmdot += projTimeScale * av[d] * bf->gravity[d] * rhoIp * masked_weights;
  }

  for (int d = 0; d < 3; ++d) {
    const double dxj = coords[3 * nR + d] - coords[3 * nL + d];
    const double kxj = av[d] - asq * inv_axdx * dxj;
    const double rhoUjIp =
      0.5 * (densityR * vel[3 * nR + d] + densityL * vel[3 * nL + d]);
    const double ujIp = 0.5 * (vel[3 * nR + d] + vel[3 * nL + d]);
    double GjIp =
      0.5 * (dpdx[3 * nR + d] / (udiagR) + dpdx[3 * nL + d] / (udiagL));
    if (bf) /* ContinuityEdgeSolverAlg.C:168-173 */
      GjIp -= 0.5 * ((bf->mask[nR] * bf->source[3 * nR + d]) / (udiagR) +
                     (bf->mask[nL] * bf->source[3 * nL + d]) / (udiagL));
    tmdot += (interpTogether * rhoUjIp + om_interpTogether * rhoIp * ujIp +
              GjIp) * av[d] -
             kxj * GjIp * nocFac;
  }
  tmdot /= tauScale;
  tmdot *= denScale;
  const double lhsfac = -asq * inv_axdx * projTimeScale * denScale / tauScale;

  lhs[0] = -lhsfac;
  lhs[1] = +lhsfac;
  rhs[0] = -tmdot;
  lhs[2] = +lhsfac;
  lhs[3] = -lhsfac;
  rhs[1] = tmdot;
}

/* ---------------------------------------------------------------- */
/* CSR graph and scatter                                             */
/* ---------------------------------------------------------------- */

static int cmp_i32(const void* a, const void* b)
{
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int orc_csr_build_edge_graph(int64_t nnode, int64_t nedge,
                             const int32_t* edges, int ndof, orc_csr* A)
{
  /* self + neighbours per node, sorted ascending (Tpetra graph is sorted
   * in finalizeLinearSystem, src/TpetraLinearSystem.C:1137-1272) */
  int64_t* deg = (int64_t*)calloc((size_t)nnode + 1, sizeof(int64_t));
  if (!deg) return -1;
  for (int64_t e = 0; e < nedge; ++e) {
    deg[edges[2 * e]]++;
    deg[edges[2 * e + 1]]++;
  }
  int64_t* nptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nnode + 1));
  nptr[0] = 0;
  for (int64_t n = 0; n < nnode; ++n) nptr[n + 1] = nptr[n] + deg[n] + 1;
  int32_t* nbr = (int32_t*)malloc(sizeof(int32_t) * (size_t)nptr[nnode]);
  for (int64_t n = 0; n < nnode; ++n) {
    deg[n] = nptr[n] + 1;
    nbr[nptr[n]] = (int32_t)n;
  }
  for (int64_t e = 0; e < nedge; ++e) {
    const int32_t a = edges[2 * e], b = edges[2 * e + 1];
    nbr[deg[a]++] = b;
    nbr[deg[b]++] = a;
  }
  for (int64_t n = 0; n < nnode; ++n)
    qsort(nbr + nptr[n], (size_t)(nptr[n + 1] - nptr[n]), sizeof(int32_t),
          cmp_i32);
  A->nnode = nnode;
  A->ndof = ndof;
  A->skip = NULL;
  A->skip_check = 1;
  A->nrows = nnode * ndof;
  A->nnz = nptr[nnode] * (int64_t)ndof * ndof;
  A->row_ptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)A->nrows + 1));
  A->cols = (int32_t*)malloc(sizeof(int32_t) * (size_t)A->nnz);
  A->vals = (double*)calloc((size_t)A->nnz, sizeof(double));
  A->rhs = (double*)calloc((size_t)A->nrows, sizeof(double));
  int64_t k = 0;
  A->row_ptr[0] = 0;
  for (int64_t n = 0; n < nnode; ++n) {
    for (int d = 0; d < ndof; ++d) {
      for (int64_t j = nptr[n]; j < nptr[n + 1]; ++j)
        for (int c = 0; c < ndof; ++c)
          A->cols[k++] = nbr[j] * ndof + c;
      A->row_ptr[n * ndof + d + 1] = k;
    }
  }
  free(deg);
  free(nptr);
  free(nbr);
  return 0;
}

void orc_csr_free(orc_csr* A)
{
  free(A->row_ptr);
  free(A->cols);
  free(A->vals);
  free(A->rhs);
  memset(A, 0, sizeof(*A));
}

/* sum_into_row / sum_into_row_vec_3, src/TpetraLinearSystem.C:1293-1379 */
static void sum_into_row(orc_csr* A, int64_t row, int nent, int ndof,
                         const int32_t* localIds, const int32_t* perm,
                         const double* input)
{
  const int64_t base = A->row_ptr[row];
  const int64_t length = A->row_ptr[row + 1] - base;
  const int32_t* colidx = A->cols + base;
  double* value = A->vals + base;
  int64_t offset = 0;
  if (ndof == 3) {
    for (int j = 0; j < nent; ++j) {
      const int id_index = 3 * j;
      const int32_t cur = localIds[id_index];
      while (colidx[offset] != cur) {
        offset += 3;
        if (offset >= length) return;
      }
      const int entry_offset = perm[id_index];
      ORC_ADD(value[offset + 0], input[entry_offset + 0]);
      ORC_ADD(value[offset + 1], input[entry_offset + 1]);
      ORC_ADD(value[offset + 2], input[entry_offset + 2]);
      offset += 3;
    }
    return;
  }
  const int numCols = nent * ndof;
  for (int j = 0; j < numCols; ++j) {
    const int perm_index = perm[j];
    const int32_t cur = localIds[j];
    while (offset < length && colidx[offset] != cur) ++offset;
    if (offset < length) ORC_ADD(value[offset], input[perm_index]);
  }
}

void orc_csr_sum_into(orc_csr* A, int nent, const int32_t* nodes,
                      const double* lhs, const double* rhs)
{
  /* sum_into, src/TpetraLinearSystem.C:1389-1459 (single rank: every row
   * is owned, column id == row id of the node) */
  const int ndof = A->ndof;
  const int numRows = nent * ndof;
  int32_t localIds[64], perm[64];
  for (int i = 0; i < nent; ++i)
    for (int d = 0; d < ndof; ++d) localIds[i * ndof + d] = nodes[i] * ndof + d;
  for (int i = 0; i < numRows; ++i) perm[i] = i;
  /* sort keys with values (keys are distinct, so any sort is the same) */
  for (int i = 1; i < numRows; ++i) {
    const int32_t k = localIds[i], v = perm[i];
    int j = i - 1;
    while (j >= 0 && localIds[j] > k) {
      localIds[j + 1] = localIds[j];
      perm[j + 1] = perm[j];
      --j;
    }
    localIds[j + 1] = k;
    perm[j + 1] = v;
  }
  for (int r = 0; r < numRows; ++r) {
    const int i = perm[r] / ndof;
    if (A->skip && A->skip_check && A->skip[nodes[i]]) continue;
    const int64_t rowLid = (int64_t)nodes[i] * ndof + perm[r] % ndof;
    const double* cur_lhs = lhs + (int64_t)perm[r] * numRows;
    sum_into_row(A, rowLid, nent, ndof, localIds, perm, cur_lhs);
    ORC_ADD(A->rhs[rowLid], rhs[perm[r]]);
  }
}

void orc_csr_reset_rows(orc_csr* A, int64_t nnodes, const int32_t* nodes,
                        int begin, int end, double diag_value,
                        double rhs_residual)
{
  /* reset_rows / reset_row, src/TpetraLinearSystem.C:1461-1520 */
  A->skip_check = 0;
  for (int64_t nn = 0; nn < nnodes; ++nn)
    for (int d = begin; d < end; ++d) {
      const int64_t row = (int64_t)nodes[nn] * A->ndof + d;
      for (int64_t k = A->row_ptr[row]; k < A->row_ptr[row + 1]; ++k)
        A->vals[k] = (A->cols[k] == row) ? diag_value : 0.0;
      A->rhs[row] = rhs_residual;
    }
}

void orc_csr_zero(orc_csr* A)
{
  memset(A->vals, 0, sizeof(double) * (size_t)A->nnz);
  memset(A->rhs, 0, sizeof(double) * (size_t)A->nrows);
  A->skip_check = 1;
}

void orc_csr_apply_dirichlet_hypre(orc_csr* A, int64_t nnodes,
                                   const int32_t* nodes, const double* sol,
                                   const double* bc, int ncomp)
{
  /* src/HypreLinearSystem.C:2446-2455: vals(row start) = 1, the row start
   * of a skipped row being its diagonal entry */
  for (int64_t nn = 0; nn < nnodes; ++nn) {
    const int32_t n = nodes[nn];
    for (int d = 0; d < A->ndof; ++d) {
      const int64_t row = (int64_t)n * A->ndof + d;
      for (int64_t k = A->row_ptr[row]; k < A->row_ptr[row + 1]; ++k)
        if (A->cols[k] == row) A->vals[k] = 1.0;
      A->rhs[row] = bc[(int64_t)n * ncomp + d] - sol[(int64_t)n * ncomp + d];
    }
  }
}

void orc_csr_apply_dirichlet(orc_csr* A, int64_t nnodes,
                             const int32_t* nodes, int begin, int end,
                             const double* sol, const double* bc, int ncomp)
{
  /* applyDirichletBCs + adjust_lhs_row, src/TpetraLinearSystem.C:1673-1757 */
  for (int64_t nn = 0; nn < nnodes; ++nn) {
    const int32_t n = nodes[nn];
    for (int d = begin; d < end; ++d) {
      const int64_t row = (int64_t)n * A->ndof + d;
      for (int64_t k = A->row_ptr[row]; k < A->row_ptr[row + 1]; ++k)
        A->vals[k] = (A->cols[k] == row) ? 1.0 : 0.0;
      A->rhs[row] = bc[(int64_t)n * ncomp + d] - sol[(int64_t)n * ncomp + d];
    }
  }
}

void orc_csr_fix_node(orc_csr* A, int32_t node, double pref, double pnode)
{
  /* src/FixPressureAtNodeAlgorithm.C:103-119: resetRows(1, node, 0, 1)
   * with the default diag 0 / rhs 0, then sum lhs = 1, rhs = pref - p.
   * Pressure is a 1-dof system. */
  if (A->ndof != 1) return;
  orc_csr_reset_rows(A, 1, &node, 0, 1, 0.0, 0.0);
  const double l = 1.0, r = pref - pnode;
  orc_csr_sum_into(A, 1, &node, &l, &r);
}

void orc_dense_sum_into(int64_t nnode, int ndof, double* lhs_dense,
                        double* rhs_dense, int nent, const int32_t* nodes,
                        const double* lhs, const double* rhs)
{
  /* edgeSumInto, unit_tests/UnitTestLinearSystem.h:42-72 */
  const int64_t n = nnode * ndof;
  const int nr = nent * ndof;
  for (int i = 0; i < nent; ++i) {
    const int64_t ioff = (int64_t)nodes[i] * ndof;
    for (int d = 0; d < ndof; ++d) rhs_dense[ioff + d] += rhs[i * ndof + d];
  }
  for (int i = 0; i < nent; ++i) {
    const int64_t ioff = (int64_t)nodes[i] * ndof;
    for (int j = 0; j < nent; ++j) {
      const int64_t joff = (int64_t)nodes[j] * ndof;
      for (int d = 0; d < ndof; ++d) {
        const int ii = i * ndof + d, jj = j * ndof + d;
        lhs_dense[(ioff + d) * n + joff + d] += lhs[ii * nr + jj];
      }
    }
  }
}

/* ---------------------------------------------------------------- */
/* Edge-loop drivers (AssembleEdgeSolverAlgorithm::run_algorithm,     */
/* include/AssembleEdgeSolverAlgorithm.h:47-100: per edge, zero the    */
/* scratch, run the physics, hand the block to the coefficient applier) */
/* ---------------------------------------------------------------- */

static void apply_block(orc_csr* A, double* ld, double* rd, int64_t nnode,
                        int ndof, int32_t nL, int32_t nR, const double* lhs,
                        const double* rhs)
{
  const int32_t nodes[2] = {nL, nR};
  if (A)
    orc_csr_sum_into(A, 2, nodes, lhs, rhs);
  else
    orc_dense_sum_into(nnode, ndof, ld, rd, 2, nodes, lhs, rhs);
}

void orc_assemble_scalar(const orc_scalar_opts* o, int64_t nnode,
                         int64_t nedge, const int32_t* edges,
                         const double* coords, const double* vrtm,
                         const double* q, const double* dqdx,
                         const double* rho, const double* dflux,
                         const double* areav, const double* mdot, orc_csr* A,
                         double* lhs_dense, double* rhs_dense)
{
#pragma omp parallel for schedule(static) num_threads(g_threads) if (A && g_threads > 1)
  for (int64_t e = 0; e < nedge; ++e) {
    double lhs[4], rhs[2];
    orc_scalar_edge(o, e, edges[2 * e], edges[2 * e + 1], coords, vrtm, q,
                    dqdx, rho, dflux, areav, mdot, lhs, rhs);
    apply_block(A, lhs_dense, rhs_dense, nnode, 1, edges[2 * e],
                edges[2 * e + 1], lhs, rhs);
  }
}

void orc_assemble_momentum(const orc_momentum_opts* o, int64_t nnode,
                           int64_t nedge, const int32_t* edges,
                           const double* coords, const double* vel,
                           const double* dudx, const double* visc,
                           const double* rho, const double* areav,
                           const double* mdot, const double* pecfac,
                           const double* wallmask, orc_csr* A,
                           double* lhs_dense, double* rhs_dense)
{
#pragma omp parallel for schedule(static) num_threads(g_threads) if (A && g_threads > 1)
  for (int64_t e = 0; e < nedge; ++e) {
    double lhs[36], rhs[6];
    orc_momentum_edge(o, e, edges[2 * e], edges[2 * e + 1], coords, vel, dudx,
                      visc, rho, areav, mdot, pecfac, wallmask, lhs, rhs);
    apply_block(A, lhs_dense, rhs_dense, nnode, 3, edges[2 * e],
                edges[2 * e + 1], lhs, rhs);
  }
}

void orc_assemble_continuity(const orc_continuity_opts* o, int64_t nnode,
                             int64_t nedge, const int32_t* edges,
                             const double* coords, const double* vel,
                             const double* dpdx, const double* rho,
                             const double* p, const double* udiag,
                             const double* areav, orc_csr* A,
                             double* lhs_dense, double* rhs_dense)
{
#pragma omp parallel for schedule(static) num_threads(g_threads) if (A && g_threads > 1)
  for (int64_t e = 0; e < nedge; ++e) {
    double lhs[4], rhs[2];
    orc_continuity_edge(o, e, edges[2 * e], edges[2 * e + 1], coords, vel,
                        dpdx, rho, p, udiag, areav, lhs, rhs);
    apply_block(A, lhs_dense, rhs_dense, nnode, 1, edges[2 * e],
                edges[2 * e + 1], lhs, rhs);
  }
}
