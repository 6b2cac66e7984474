// nw_amg_setup.h -- host construction of a smoothed-aggregation AMG
// hierarchy for the pressure-Poisson matrix (the role MueLu plays in the
// reference: src/LinearSolver.C:130-164 with reg_tests/xml/milestone.xml:
// smoothed aggregation, "uncoupled" aggregation, drop tolerance 0.02,
// Chebyshev(2) smoothing with eigenvalue ratio 20, coarse max size 1000).
#pragma once

#include <cstdint>
#include <vector>

namespace nw {
namespace amg {

struct Csr {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
  int64_t nnz() const { return rp.empty() ? 0 : rp.back(); }
};

struct Level {
  Csr A;                    // operator on this level
  std::vector<double> dinv; // 1 / diag(A) (0 on empty rows)
  double lmax = 0.0;        // lambda_max(D^-1 A) estimate
  Csr P, R;                 // to / from the next level (empty on the last)
};

struct Options {
  double drop_tol = 0.02;     // strength threshold
  int coarse_max = 1000;      // direct solve below this size
  int max_levels = 12;
  double cheb_ratio = 20.0;   // smoother interval [lmax/ratio, lmax]
  int cheb_degree = 2;
  int power_iters = 15;
};

struct Hierarchy {
  std::vector<Level> levels;
  std::vector<double> coarse_inv;  // dense inverse of the last A, row-major
  double setup_seconds = 0.0;
};

// Builds the hierarchy from the fine matrix (CSR with sorted columns and a
// diagonal entry in every row), which becomes levels[0].A (pass it with
// std::move to avoid a copy).  Returns 0 or a negative error.
int build(Csr A, const Options& o, Hierarchy& h);

}  // namespace amg
}  // namespace nw
