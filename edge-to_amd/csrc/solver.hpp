// GMG-PCG solver state (see solver.hip).
#pragma once
#include <vector>

#include "coarse.hpp"
#include "common.hpp"
#include "level1.hpp"
#include "stiffness.hpp"

namespace et {

// device scalar slots of the PCG loop
// SC_RHO, SC_RHO + 1: r.z of consecutive PCG iterations (slots alternate)
enum { SC_BB = 0, SC_RHO, SC_RHO_ALT, SC_PQ, SC_RR, SC_OMEGA, SC_R, SC_STATUS, SC_STOP, SC_ALPHA, SC_COUNT };

// Event-based per-kernel-class timing (enabled on demand).
struct Prof {
  unsigned mask = 0;  // bit k: time kernel class k
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;  // recorded this window
  std::vector<hipEvent_t> pool;
  size_t next = 0;
  int open_kind = -1;
  hipEvent_t ev() {
    if (next == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[next++];
  }
  void begin(int kind, hipStream_t s) {
    if (!((mask >> kind) & 1u) || recs.size() > 200000) return;
    Rec r{kind, ev(), ev()};
    if (!r.a || !r.b) return;
    (void)hipEventRecord(r.a, s);
    recs.push_back(r);
    open_kind = kind;
  }
  void end(hipStream_t s) {
    if (open_kind < 0) return;
    (void)hipEventRecord(recs.back().b, s);
    open_kind = -1;
  }
  void reset() {
    recs.clear();
    next = 0;
    open_kind = -1;
  }
  void release() {
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
    recs.clear();
    next = 0;
  }
};

// operator data shared by all levels
struct MGShared {
  Prof* prof = nullptr;
  bool tiled = false;  // fine grid uses the LDS-tiled H8 operator
  bool mfma = false;   // FP32 fine-grid sweeps use the matrix-core operator
  float kscale = 1.f;  // power-of-two scale of K0 for the FP16 split
  WhtTable wht{};      // transformed-basis coefficients of K0 (when tiled)
  double* K0d = nullptr;
  float* K0f = nullptr;
  const double* Ed = nullptr;  // caller's E (FP64)
  float* Ef = nullptr;         // FP32 copy (mixed precision)
  const uint8_t* cons = nullptr;
  double* scal = nullptr;      // device scalars (SC_*)
  const Level1MF* l1 = nullptr;  // matrix-free first coarse level, when enabled
  // When set, the last fine-grid sweep of the V-cycle also writes the
  // per-workgroup partial sums of b.z (b the fine-grid right-hand side in the
  // V-cycle's precision, z its output) to dot_part and sets dot_done.
  bool want_dot = false;
  double* dot_part = nullptr;
  bool dot_done = false;
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
  T* S = nullptr;  // centre + forward-offset half of the symmetric block stencil (coarse smoothing levels)
};

template <typename T>
struct MGHierarchy {
  std::vector<MGLevel<T>> lv;
  int dof = 0, n_level = 0, n_smooth = 0, cycle_w = 0, w_level_mask = 0;
  CoarseSolver* coarse = nullptr;
  // skip_s1: level 1 is applied matrix-free (no stencil copy needed)
  int alloc(const Grid& g0, int dof, int n_level, hipStream_t s, bool skip_s1 = false);
  void release();
  // x0_ready: level l's x already holds w D^-1 b (written by the restriction)
  int vcycle(int l, MGShared& sh, hipStream_t s, bool x0_ready = false);
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
  bool elem_galerkin = false;  // level 1 built from elements (level-0 stencil not formed)
  double* Mt = nullptr;        // child tables W_t^T K0 W_t
  uint8_t* eflag = nullptr;    // per-element "touches a constrained DOF"
  double* gal_work = nullptr;  // intermediates of the three-pass 3-D Galerkin product (null: one-pass)
  Level1MF l1;                 // matrix-free level 1 (3-D, even counts, n_level >= 2)
  bool use_l1 = false;
  int k0_flags = 0;            // k0_structure() of the element matrix (fast 3-D kernels only if it allows)

  double *r = nullptr, *z = nullptr, *p_ = nullptr, *q = nullptr, *diag64 = nullptr;
  double* partial = nullptr;  // 2 x kRedBlocks first-stage reduction sums
  double* ring = nullptr;      // pinned host slots {R, status, stop} of the last two iterations
  double* ring_dev = nullptr;  // device address of ring
  hipEvent_t ring_ev[2] = {nullptr, nullptr};
  double max_omega = 0, min_omega = 0, d_omega = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  double t_setup_ms = 0, t_pcg_ms = 0;
  Prof prof;
  int profile_read(double* ms, int64_t* launches);
  int apply_level(int level, const void* x, void* y);
  int coarse_solve(const void* b, void* x);
  bool built = false;  // a solve has set up the hierarchy

  int create(const et_mg_params& p, const double* K0_host, hipStream_t s);
  void destroy();
  int solve(const double* E, const uint8_t* cons, const double* b, double* x, double tol,
            int maxiter, int* iters, double* residual);

 private:
  template <typename T>
  int setup_levels(MGHierarchy<T>& h);
  template <typename T>
  int precondition(MGHierarchy<T>& h, bool rhs_ready, const T** z);
  template <typename T>
  int pcg(MGHierarchy<T>& h, const double* b, double* x, double tol, int maxiter, int* iters,
          double* residual);
  int apply_A64(const double* x, double* y, int mode, const double* b);
};

void set_row_stencil_nodes(int64_t n);
void set_fused_rz(int on);
int get_fused_rz();
int64_t get_row_stencil_nodes();

}  // namespace et

struct et_mg {
  et::MGSolver impl;
};
