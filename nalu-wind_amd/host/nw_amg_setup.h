// nw_amg_setup.h -- host construction of a smoothed-aggregation AMG
// hierarchy for the pressure-Poisson matrix (the role MueLu plays in the
// reference: src/LinearSolver.C:130-164 with reg_tests/xml/milestone.xml:
// smoothed aggregation, "uncoupled" aggregation, drop tolerance 0.02,
// Chebyshev(2) smoothing with eigenvalue ratio 20, coarse max size 1000).
#pragma once

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

namespace nw {
namespace amg {

// Storage whose resize leaves new elements uninitialised: the setup's
// arrays are filled by parallel loops right after they are sized, so a
// value-initialising resize would only add a serial pass over memory.
template <class T>
struct DefaultInit : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = DefaultInit<U>;
  };
  DefaultInit() = default;
  template <class U>
  DefaultInit(const DefaultInit<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept
  {
    ::new ((void*)p) U;
  }
  template <class U, class... Args>
  void construct(U* p, Args&&... args)
  {
    ::new ((void*)p) U(std::forward<Args>(args)...);
  }
};
template <class T>
using uvec = std::vector<T, DefaultInit<T>>;

struct Csr {
  int64_t nrows = 0, ncols = 0;
  uvec<int64_t> rp;
  uvec<int32_t> ci;
  uvec<double> v;
  int64_t nnz() const { return rp.empty() ? 0 : rp.back(); }
};

// The start vector of the setup's power iterations (a fixed xorshift
// sequence in [0.5, 1.5)).
std::vector<double> power_start(int64_t n);

// A CSR matrix of n rows with the given structure (rp [n + 1], ci); its
// values are left for the caller to fill.  Copies in parallel.
Csr csr_with_structure(int64_t n, const int32_t* rp, const int32_t* ci);

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
  // lambda_max(D^-1 A) of each level for its Chebyshev smoother, by power
  // iteration on the host; false leaves Level::lmax at 0 for the caller to
  // estimate where the operators live (the prolongator damping still uses
  // the host's estimate on the filtered operator)
  bool smoother_lmax = true;
  // > 0: lambda_max(D_F^-1 F) of the fine level's filtered operator, as
  // estimated by the caller, used for the fine prolongator's damping
  double filtered_lmax0 = 0.0;
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
