// nw_amg_setup.cpp -- smoothed-aggregation AMG setup on the host.
//
// Algorithm (Vanek, Mandel, Brezina, "Algebraic multigrid by smoothed
// aggregation for second and fourth order elliptic problems", Computing
// 56, 1996), with the settings the reference uses for the pressure system
// through MueLu (reg_tests/xml/milestone.xml):
//   strength     |a_ij|^2 > theta^2 |a_ii a_jj|, theta = drop tol 0.02;
//                weak entries are lumped into the diagonal (filtered A)
//   aggregation  "uncoupled": phase 1 roots whose strong neighbourhood is
//                free, phase 2 joins leftovers to the most strongly
//                connected aggregate, phase 3 makes aggregates of the rest
//   P            (I - omega D_F^-1 A_F) P0, P0 = normalised aggregate
//                indicator, omega = 4/3 / lambda_max(D_F^-1 A_F)
//   coarse A     R A P with R = P^T
// Rows whose off-diagonals are all zero (Dirichlet rows) stay out of the
// coarse space.  Work is split over OpenMP threads by rows; every row is
// computed by one thread in a fixed order, so the result does not depend on
// the thread count.
#include "nw_amg_setup.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace nw {
namespace amg {

namespace {

}  // namespace

std::vector<double> power_start(int64_t n)
{
  std::vector<double> x(n);
  uint64_t z = 0x2545f4914f6cdd1dULL;
  for (int64_t i = 0; i < n; ++i) {
    z ^= z << 13;
    z ^= z >> 7;
    z ^= z << 17;
    x[i] = 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
  return x;
}

namespace {

// Sums run over fixed chunks of kChunk elements in parallel and then over
// the chunks in order, so a reduction gives the same bits for any thread
// count.
constexpr int64_t kChunk = 1 << 14;

double dot(const std::vector<double>& a, const std::vector<double>& b)
{
  const int64_t n = (int64_t)a.size(), nc = (n + kChunk - 1) / kChunk;
  std::vector<double> part(nc, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    double s = 0.0;
    for (int64_t i = c * kChunk; i < std::min(n, (c + 1) * kChunk); ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0.0;
  for (double v : part) s += v;
  return s;
}

// lambda_max(D^-1 A) by power iteration from a fixed start vector.  Each
// step normalises x, then forms y = D^-1 A x together with the chunked
// partial sums of x.y (the Rayleigh quotient) and y.y (the next norm), in
// the chunks of dot(), so the result equals the unfused sequence bitwise.
double power_lmax(const Csr& A, const std::vector<double>& dinv, int iters)
{
  const int64_t n = A.nrows, nc = (n + kChunk - 1) / kChunk;
  std::vector<double> x = power_start(n), y(n);
  std::vector<double> pxy(nc), pyy(nc);
  double lam = 1.0, nx2 = dot(x, x);
  for (int it = 0; it < iters; ++it) {
    const double nx = std::sqrt(nx2);
    if (!(nx > 0)) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] /= nx;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
      double sxy = 0.0, syy = 0.0;
      for (int64_t i = c * kChunk; i < std::min(n, (c + 1) * kChunk); ++i) {
        double s = 0.0;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.v[k] * x[A.ci[k]];
        y[i] = s * dinv[i];
        sxy += x[i] * y[i];
        syy += y[i] * y[i];
      }
      pxy[c] = sxy;
      pyy[c] = syy;
    }
    lam = 0.0;
    nx2 = 0.0;
    for (int64_t c = 0; c < nc; ++c) {
      lam += pxy[c];
      nx2 += pyy[c];
    }
    x.swap(y);
  }
  return lam;
}

// 1 / a_ii per row (0 where the diagonal is 0 or absent)
std::vector<double> inv_diagonal(const Csr& A)
{
  std::vector<double> d(A.nrows, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < A.nrows; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
      if (A.ci[k] == i && A.v[k] != 0.0) d[i] = 1.0 / A.v[k];
  return d;
}

// R = P^T with sorted columns: entries counted and placed per column in
// parallel, then each row of R sorted by column (the rows of P), which
// gives the same arrays as a serial row-order fill.
Csr transpose(const Csr& P)
{
  Csr R;
  R.nrows = P.ncols;
  R.ncols = P.nrows;
  R.rp.assign(R.nrows + 1, 0);
  const int64_t nnz = P.nnz();
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nnz; ++k) {
#pragma omp atomic
    R.rp[P.ci[k] + 1]++;
  }
  for (int64_t i = 0; i < R.nrows; ++i) R.rp[i + 1] += R.rp[i];
  R.ci.resize(nnz);
  R.v.resize(nnz);
  uvec<int64_t> pos(R.rp.begin(), R.rp.end() - 1);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < P.nrows; ++i)
    for (int64_t k = P.rp[i]; k < P.rp[i + 1]; ++k) {
      int64_t p;
#pragma omp atomic capture
      p = pos[P.ci[k]]++;
      R.ci[p] = (int32_t)i;
      R.v[p] = P.v[k];
    }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t r = 0; r < R.nrows; ++r)  // insertion sort: rows are short
    for (int64_t k = R.rp[r] + 1; k < R.rp[r + 1]; ++k) {
      const int32_t c = R.ci[k];
      const double v = R.v[k];
      int64_t q = k;
      for (; q > R.rp[r] && R.ci[q - 1] > c; --q) {
        R.ci[q] = R.ci[q - 1];
        R.v[q] = R.v[q - 1];
      }
      R.ci[q] = c;
      R.v[q] = v;
    }
  return R;
}

// C = A B (Gustavson, sorted output rows); each row built by one thread
Csr spgemm(const Csr& A, const Csr& B)
{
  Csr C;
  C.nrows = A.nrows;
  C.ncols = B.ncols;
  C.rp.assign(A.nrows + 1, 0);
  std::vector<int64_t> cnt(A.nrows, 0);
#pragma omp parallel
  {
    std::vector<int32_t> mark(B.ncols, -1);
    std::vector<int32_t> cols;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < A.nrows; ++i) {
      int64_t c = 0;
      for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int32_t j = A.ci[k];
        for (int64_t l = B.rp[j]; l < B.rp[j + 1]; ++l)
          if (mark[B.ci[l]] != (int32_t)i) {
            mark[B.ci[l]] = (int32_t)i;
            ++c;
          }
      }
      cnt[i] = c;
    }
  }
  for (int64_t i = 0; i < A.nrows; ++i) C.rp[i + 1] = C.rp[i] + cnt[i];
  C.ci.resize(C.rp[A.nrows]);
  C.v.resize(C.rp[A.nrows]);
#pragma omp parallel
  {
    std::vector<double> acc(B.ncols, 0.0);
    std::vector<int32_t> mark(B.ncols, -1);
    std::vector<int32_t> cols;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < A.nrows; ++i) {
      cols.clear();
      for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int32_t j = A.ci[k];
        const double a = A.v[k];
        for (int64_t l = B.rp[j]; l < B.rp[j + 1]; ++l) {
          const int32_t c = B.ci[l];
          if (mark[c] != (int32_t)i) {
            mark[c] = (int32_t)i;
            acc[c] = 0.0;
            cols.push_back(c);
          }
          acc[c] += a * B.v[l];
        }
      }
      std::sort(cols.begin(), cols.end());
      int64_t p = C.rp[i];
      for (int32_t c : cols) {
        C.ci[p] = c;
        C.v[p++] = acc[c];
      }
    }
  }
  return C;
}

int invert_dense(int64_t n, const Csr& A, std::vector<double>& inv)
{
  std::vector<double> M(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) M[i * n + A.ci[k]] += A.v[k];
  inv.assign(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  double scale = 0.0;
  for (double v : M) scale = std::max(scale, std::fabs(v));
  for (int64_t c = 0; c < n; ++c) {
    int64_t p = c;
    for (int64_t r = c + 1; r < n; ++r)
      if (std::fabs(M[r * n + c]) > std::fabs(M[p * n + c])) p = r;
    if (std::fabs(M[p * n + c]) <= 1e-14 * scale) return -1;
    if (p != c)
      for (int64_t j = 0; j < n; ++j) {
        std::swap(M[p * n + j], M[c * n + j]);
        std::swap(inv[p * n + j], inv[c * n + j]);
      }
    const double d = 1.0 / M[c * n + c];
    for (int64_t j = 0; j < n; ++j) {
      M[c * n + j] *= d;
      inv[c * n + j] *= d;
    }
    for (int64_t r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = M[r * n + c];
      if (f == 0.0) continue;
      for (int64_t j = 0; j < n; ++j) {
        M[r * n + j] -= f * M[c * n + j];
        inv[r * n + j] -= f * inv[c * n + j];
      }
    }
  }
  return 0;
}

// one coarsening step: P and R for level A (lf_given > 0: the damping's
// eigenvalue estimate, else estimated here)
void coarsen(const Csr& A, const Options& o, Csr& P, double lf_given)
{
  const int64_t n = A.nrows;
  std::vector<double> diag(n, 0.0);
  std::vector<char> dirichlet(n, 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      if (A.ci[k] == i)
        diag[i] = A.v[k];
      else if (A.v[k] != 0.0)
        dirichlet[i] = 0;
    }
  // strength and filtered matrix (weak entries lumped to the diagonal):
  // rows counted, then filled, in parallel
  const double th2 = o.drop_tol * o.drop_tol;
  auto keep = [&](int64_t i, int64_t k) {
    const int32_t j = A.ci[k];
    const double a = A.v[k];
    return j == i || (!dirichlet[i] && !dirichlet[j] &&
                      a * a > th2 * std::fabs(diag[i] * diag[j]));
  };
  Csr F;
  F.nrows = F.ncols = n;
  F.rp.assign(n + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int64_t c = 0;
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) c += keep(i, k);
    F.rp[i + 1] = c;
  }
  for (int64_t i = 0; i < n; ++i) F.rp[i + 1] += F.rp[i];
  F.ci.resize(F.rp[n]);
  F.v.resize(F.rp[n]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double lump = 0.0;
    int64_t dpos = -1, p = F.rp[i];
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      if (keep(i, k)) {
        if (A.ci[k] == i) dpos = p;
        F.ci[p] = A.ci[k];
        F.v[p++] = A.v[k];
      } else {
        lump += A.v[k];
      }
    }
    if (dpos >= 0) F.v[dpos] += lump;
  }
  // aggregation
  std::vector<int32_t> agg(n, -1);
  for (int64_t i = 0; i < n; ++i)
    if (dirichlet[i]) agg[i] = -2;
  int32_t na = 0;
  for (int64_t i = 0; i < n; ++i) {  // phase 1
    if (agg[i] != -1) continue;
    bool free_nbhd = true;
    for (int64_t k = F.rp[i]; k < F.rp[i + 1] && free_nbhd; ++k)
      if (F.ci[k] != i && agg[F.ci[k]] != -1) free_nbhd = false;
    if (!free_nbhd) continue;
    agg[i] = na;
    for (int64_t k = F.rp[i]; k < F.rp[i + 1]; ++k) agg[F.ci[k]] = na;
    ++na;
  }
  std::vector<int32_t> agg1 = agg;
  for (int64_t i = 0; i < n; ++i) {  // phase 2
    if (agg[i] != -1) continue;
    double best = 0.0;
    int32_t ba = -1;
    for (int64_t k = F.rp[i]; k < F.rp[i + 1]; ++k) {
      const int32_t j = F.ci[k];
      if (j != i && agg1[j] >= 0 && std::fabs(F.v[k]) > best) {
        best = std::fabs(F.v[k]);
        ba = agg1[j];
      }
    }
    if (ba >= 0) agg[i] = ba;
  }
  for (int64_t i = 0; i < n; ++i) {  // phase 3
    if (agg[i] != -1) continue;
    agg[i] = na;
    for (int64_t k = F.rp[i]; k < F.rp[i + 1]; ++k)
      if (agg[F.ci[k]] == -1) agg[F.ci[k]] = na;
    ++na;
  }
  std::vector<int32_t> size(na, 0);
  for (int64_t i = 0; i < n; ++i)
    if (agg[i] >= 0) size[agg[i]]++;
  std::vector<double> sq(na);  // the aggregate indicator's norm
  for (int32_t a = 0; a < na; ++a) sq[a] = std::sqrt((double)size[a]);
  // omega from lambda_max(D_F^-1 A_F)
  const std::vector<double> dfinv = inv_diagonal(F);
  const double lf = lf_given > 0 ? lf_given : power_lmax(F, dfinv, o.power_iters);
  const double omega = (4.0 / 3.0) / (lf > 0 ? lf : 1.0);
  // P = P0 - omega D_F^-1 A_F P0, one row per node: a row's terms (its
  // aggregate's indicator and a smoothing term per F entry into an
  // aggregate) are merged by aggregate; the rows are formed twice, once to
  // count and once to write, which costs less than staging them
  P.nrows = n;
  P.ncols = na;
  P.rp.assign(n + 1, 0);
  using Term = std::pair<int32_t, double>;
  auto row_terms = [&](int64_t i, std::vector<Term>& r) {
    r.clear();
    r.push_back({agg[i], 1.0 / sq[agg[i]]});
    for (int64_t k = F.rp[i]; k < F.rp[i + 1]; ++k) {
      const int32_t j = F.ci[k];
      if (agg[j] < 0) continue;
      r.push_back({agg[j], -omega * dfinv[i] * F.v[k] / sq[agg[j]]});
    }
    std::sort(r.begin(), r.end(), [](const Term& a, const Term& b) { return a.first < b.first; });
    size_t w = 0;
    for (size_t q = 0; q < r.size(); ++q) {
      if (w > 0 && r[w - 1].first == r[q].first)
        r[w - 1].second += r[q].second;
      else
        r[w++] = r[q];
    }
    r.resize(w);
  };
#pragma omp parallel
  {
    std::vector<Term> r;
#pragma omp for schedule(dynamic, 4096)
    for (int64_t i = 0; i < n; ++i) {
      if (agg[i] < 0) continue;
      row_terms(i, r);
      P.rp[i + 1] = (int64_t)r.size();
    }
  }
  for (int64_t i = 0; i < n; ++i) P.rp[i + 1] += P.rp[i];
  P.ci.resize(P.rp[n]);
  P.v.resize(P.rp[n]);
#pragma omp parallel
  {
    std::vector<Term> r;
#pragma omp for schedule(dynamic, 4096)
    for (int64_t i = 0; i < n; ++i) {
      if (agg[i] < 0) continue;
      row_terms(i, r);
      for (size_t q = 0; q < r.size(); ++q) {
        P.ci[P.rp[i] + q] = r[q].first;
        P.v[P.rp[i] + q] = r[q].second;
      }
    }
  }
}

}  // namespace

Csr csr_with_structure(int64_t n, const int32_t* rp, const int32_t* ci)
{
  Csr A;
  A.nrows = A.ncols = n;
  A.rp.resize(n + 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i <= n; ++i) A.rp[i] = rp[i];
  const int64_t nnz = rp[n];
  A.ci.resize(nnz);
  A.v.resize(nnz);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nnz; ++k) A.ci[k] = ci[k];
  return A;
}

int build(Csr A0, const Options& o, Hierarchy& h)
{
  const auto t0 = std::chrono::steady_clock::now();
  h.levels.clear();
  h.levels.push_back(Level());
  h.levels[0].A = std::move(A0);
  for (int l = 0;; ++l) {
    Level& L = h.levels[l];
    const int64_t n = L.A.nrows;
    L.dinv = inv_diagonal(L.A);
    if (o.smoother_lmax) L.lmax = power_lmax(L.A, L.dinv, o.power_iters);
    if (n <= o.coarse_max || l + 1 >= o.max_levels) break;
    Csr P;
    coarsen(L.A, o, P, l == 0 ? o.filtered_lmax0 : 0.0);
    if (P.ncols == 0 || P.ncols > 0.9 * n) break;
    Csr R = transpose(P);
    Csr AP = spgemm(L.A, P);
    Level next;
    next.A = spgemm(R, AP);
    L.P = std::move(P);
    L.R = std::move(R);
    h.levels.push_back(std::move(next));
  }
  Level& last = h.levels.back();
  if (last.A.nrows > 8 * o.coarse_max) return -1;
  if (invert_dense(last.A.nrows, last.A, h.coarse_inv) != 0) return -2;
  h.setup_seconds = std::chrono::duration<double>(
                      std::chrono::steady_clock::now() - t0)
                      .count();
  return 0;
}

}  // namespace amg
}  // namespace nw
