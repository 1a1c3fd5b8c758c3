// Matrix-free operator of the first coarse multigrid level (3-D, even element
// counts): A_1 = R A_hat R^T applied element by element from the fine element
// densities, plus a stored correction at the coarse nodes that constrained
// fine DOFs reach.  See level1.hip.
#pragma once
#include "common.hpp"

namespace et {

struct Level1MF {
  Grid g0{}, g1{};
  bool mixed = false;         // FP32 apply (else FP64)
  void* cf = nullptr;         // element coefficients (kL1Terms, working precision)
  void* Eh = nullptr;         // per coarse element: transform of its 8 children's E [n_elem1][8]
  void* scratch = nullptr;    // per coarse element: bottom-face node sums [12][n_elem1], unmoved top results [12][n_elem1]
  uint8_t* special = nullptr; // per coarse element: a child touches a constrained DOF
  int* aidx = nullptr;        // per coarse node: row in the correction list, or -1
  int* alist = nullptr;       // corrected coarse nodes
  int* acount = nullptr;      // device counter
  int* acount_h = nullptr;    // pinned copy
  int n_aff = 0;
  double* dA64 = nullptr;     // correction blocks [27][3][3][n_aff] (FP64)
  void* dA = nullptr;         // the same in working precision
  int dA_cap = 0;

  int init(const Grid& fine, const double* K0_host, bool mixed_, hipStream_t s);
  // After the level-1 Galerkin stencil S1 (true operator, FP64) has been built
  // from E, cons and the element flags: refresh E transforms and corrections.
  int setup(const double* E, const double* Mt, const uint8_t* eflag, const double* S1, hipStream_t s);
  // Level-1 operator in the smoother modes (0 y = Ax, 1 b - Ax, 2 Jacobi, 3 Jacobi + add).
  template <typename T>
  int apply(int mode, const T* x, const T* b, const T* Dinv, const double* omega, double wscale,
            const T* xadd, T* y, hipStream_t s) const;
  void release();
};

bool level1_supported(const Grid& fine, int n_level);
// whether the fixed term list of the element product holds every entry of
// H M_0 H / 64 above rtol * max for this element matrix (host)
bool level1_terms_cover(const double* K0_host, double rtol);
void set_level1_mf(int on);
int get_level1_mf();

}  // namespace et
