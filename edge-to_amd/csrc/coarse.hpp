// Dense FP64 coarsest-grid solver (blocked Cholesky + explicit inverse).
#pragma once
#include "common.hpp"

namespace et {

struct CoarseSolver {
  int n = 0;    // matrix order
  int ld = 0;   // padded order (multiple of 64)
  int nbt = 0;  // tiles per side
  double* A = nullptr;  // after factor(): A^-1 (ld x ld)
  double* Y = nullptr;  // L^-1 workspace
  int* info = nullptr;
  int* info_h = nullptr;

  int init(int n, hipStream_t s);
  void release();
  // Expand the block stencil S on grid g, factor, and form the inverse.
  int factor(const Grid& g, int dof, const double* S, hipStream_t s);
  // Synchronise and report a non-positive pivot, if any.
  int check_info(hipStream_t s);
  template <typename TI, typename TO>
  int solve(const TI* b, TO* x, hipStream_t s) const;
};

}  // namespace et
