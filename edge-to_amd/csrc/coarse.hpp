// Dense FP64 coarsest-grid solver (Cholesky + explicit inverse).
#pragma once
#include "common.hpp"

namespace et {

struct CoarseSolver {
  int n = 0;    // matrix order
  int ld = 0;   // padded order (multiple of 64)
  int nbt = 0;  // tiles per side
  double* A = nullptr;  // after factor(): A^-1 (lib_mode 0, 1) or L^-1 (lib_mode 2), ld x ld
  double* Y = nullptr;  // lib_mode 0: L^-1 workspace
  int* info = nullptr;    // [0] factor, [1] inverse (device)
  int* info_h = nullptr;  // pinned host copy
  bool use_lib = true;    // rocBLAS / rocSOLVER handle needed (lib_mode 1, 2)
  int lib_mode = 2;       // 2: band Cholesky + trtri, 1: rocSOLVER potrf + potri, 0: in-house dense
  double* Wd = nullptr;   // lib_mode 2: inverse of the current diagonal factor tile (64x64)
  double* tv = nullptr;   // lib_mode 2: intermediate vector of the coarse solve
  double* Z = nullptr;    // lib_mode 2: L^-1 row-major (A holds it column-major)
  float* Zf = nullptr;    // lib_mode 2 with an FP32 copy: Z and A rounded to FP32, applied by FP32 solves
  float* Af = nullptr;
  void* blas = nullptr;   // rocblas_handle

  // f32_copy: FP32 solves (mixed-precision V-cycles) apply FP32 copies of the
  // FP64 factors (lib_mode 2)
  int init(int n, hipStream_t s, bool f32_copy = false);
  void release();
  // Expand the block stencil S on grid g, factor, and form the inverse.
  int factor(const Grid& g, int dof, const double* S, hipStream_t s);
  // Synchronise and report a non-positive pivot, if any.
  int check_info(hipStream_t s);
  template <typename TI, typename TO>
  int solve(const TI* b, TO* x, hipStream_t s) const;
};

// Factor path of coarse solvers created afterwards: 2 = band Cholesky (in-house
// tile kernels, rocBLAS band updates) + rocSOLVER trtri (default), 1 = rocSOLVER
// potrf + potri, 0 = in-house dense Cholesky and inverse.
void set_coarse_lib(int mode);
int get_coarse_lib();

}  // namespace et
