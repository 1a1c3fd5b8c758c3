// Dense FP64 coarsest-grid solver (Cholesky + explicit inverse).
#pragma once
#include "common.hpp"

namespace et {

struct CoarseSolver {
  int n = 0;    // matrix order
  int ld = 0;   // padded order (multiple of 64)
  int nbt = 0;  // tiles per side
  double* A = nullptr;  // after factor(): A^-1 (ld x ld)
  double* Y = nullptr;  // L^-1 workspace
  int* info = nullptr;    // [0] factor, [1] inverse (device)
  int* info_h = nullptr;  // pinned host copy
  bool use_lib = true;    // rocSOLVER potrf + potri, else the in-house kernels
  void* blas = nullptr;   // rocblas_handle

  int init(int n, hipStream_t s);
  void release();
  // Expand the block stencil S on grid g, factor, and form the inverse.
  int factor(const Grid& g, int dof, const double* S, hipStream_t s);
  // Synchronise and report a non-positive pivot, if any.
  int check_info(hipStream_t s);
  template <typename TI, typename TO>
  int solve(const TI* b, TO* x, hipStream_t s) const;
};

// Factor path of coarse solvers created afterwards: 1 = rocSOLVER (default).
void set_coarse_lib(int on);
int get_coarse_lib();

}  // namespace et
