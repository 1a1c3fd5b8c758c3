// GMG-PCG solver state (see solver.hip).
#pragma once
#include <vector>

#include "coarse.hpp"
#include "common.hpp"

namespace et {

// device scalar slots of the PCG loop
enum { SC_BB = 0, SC_RHO, SC_RHO_NEW, SC_PQ, SC_RR, SC_OMEGA, SC_COUNT };

// operator data shared by all levels
struct MGShared {
  double* K0d = nullptr;
  float* K0f = nullptr;
  const double* Ed = nullptr;  // caller's E (FP64)
  float* Ef = nullptr;         // FP32 copy (mixed precision)
  const uint8_t* cons = nullptr;
  double* scal = nullptr;      // device scalars (SC_*)
  template <typename T>
  const T* K0() const;
  template <typename T>
  const T* E() const;
};
template <>
inline const double* MGShared::K0<double>() const { return K0d; }
template <>
inline const float* MGShared::K0<float>() const { return K0f; }
template <>
inline const double* MGShared::E<double>() const { return Ed; }
template <>
inline const float* MGShared::E<float>() const { return Ef; }

template <typename T>
struct MGLevel {
  int level = 0;
  Grid g{};
  int64_t n = 0;  // DOFs on this level
  T *b = nullptr, *x = nullptr, *r = nullptr, *e = nullptr, *t = nullptr;
  T* Dinv = nullptr;
  T* S = nullptr;  // block stencil (coarse smoothing levels only)
};

template <typename T>
struct MGHierarchy {
  std::vector<MGLevel<T>> lv;
  int dof = 0, n_level = 0, n_smooth = 0, cycle_w = 0, w_level_mask = 0;
  CoarseSolver* coarse = nullptr;
  int alloc(const Grid& g0, int dof, int n_level, hipStream_t s);
  void release();
  int vcycle(int l, MGShared& sh, hipStream_t s);
};

struct MGSolver {
  et_mg_params prm{};
  hipStream_t stream = nullptr;
  Grid g0{}, coarse_grid{};
  int dof = 0;
  int64_t n_dof = 0, n_elem = 0;
  MGShared sh;
  MGHierarchy<double> h64;
  MGHierarchy<float> h32;
  CoarseSolver coarse;
  std::vector<double*> S64;  // FP64 Galerkin stencils, level 0..n_level
  double *r = nullptr, *z = nullptr, *p_ = nullptr, *q = nullptr, *diag64 = nullptr;
  double* partial = nullptr;
  double* host_scal = nullptr;
  double max_omega = 0, min_omega = 0, d_omega = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  double t_setup_ms = 0, t_pcg_ms = 0;

  int create(const et_mg_params& p, const double* K0_host, hipStream_t s);
  void destroy();
  int solve(const double* E, const uint8_t* cons, const double* b, double* x, double tol,
            int maxiter, int* iters, double* residual);

 private:
  template <typename T>
  int setup_levels(MGHierarchy<T>& h);
  template <typename T>
  int precondition(MGHierarchy<T>& h);
  int apply_A64(const double* x, double* y, int mode, const double* b);
  int read_scalars();
};

}  // namespace et

struct et_mg {
  et::MGSolver impl;
};
