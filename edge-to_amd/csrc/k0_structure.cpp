// Host check of the element-matrix structure that the fast 3-D kernels use.
//
// The reference accepts any element matrix (geom/_mesh.py:113-135 take a
// user `stiffness=` callable).  The transformed-basis and LDS-tiled fine-grid
// operators and the matrix-free first coarse level instead derive K0's blocks
// from reflections of the brick:
//   * K0 symmetric;
//   * reflection invariance: with corners in bit order beta = x + 2y + 4z and
//     R_a = diag((-1)^{a_x}, (-1)^{a_y}, (-1)^{a_z}), every 3x3 block satisfies
//     K0[a][b] = R_a K0[0][a ^ b] R_a (true for isotropic and axis-aligned
//     orthotropic materials on any brick; false e.g. for a material rotated
//     about an axis);
//   * in the corner Walsh-Hadamard basis G = H K0 H / 64, component c of
//     index M couples only with component d of index M ^ e_c ^ e_d, and only
//     the entries the transformed-basis kernel keeps (wht_nz) are non-zero.
// Callers route matrices that fail any test to the general node-gather kernels.
#include <cmath>

#include "level1.hpp"
#include "op_tiled.hpp"
#include "stiffness.hpp"

namespace et {

namespace {

// wht_nz of op_wht.hip: structural non-zeros of G[M][c][d], M = 1..7
constexpr bool g_kept(int M, int c, int d) {
  if (M == 0) return false;
  if (c == d) return true;
  if (M == 7) return false;
  const int pair = (1 << c) | (1 << d);
  const int bits = __builtin_popcount(M);
  return bits == 1 ? (pair & M) != 0 : pair != M;
}

inline int corner(int beta) { return h8_corner(beta & 1, (beta >> 1) & 1, beta >> 2); }

}  // namespace

int k0_structure(int dim, const double* K0, double rtol) {
  const int nd = dim == 3 ? 24 : 8;
  double kmax = 0.0;
  for (int t = 0; t < nd * nd; ++t) kmax = std::fmax(kmax, std::fabs(K0[t]));
  if (!(kmax > 0.0) || !std::isfinite(kmax)) return 0;
  const double tol = rtol * kmax;
  int flags = 0;
  bool sym = true;
  for (int i = 0; i < nd && sym; ++i)
    for (int j = 0; j < i; ++j)
      if (std::fabs(K0[i * nd + j] - K0[j * nd + i]) > tol) {
        sym = false;
        break;
      }
  if (sym) flags |= K0S_SYM;
  if (dim != 3) return flags;
  auto K = [&](int a, int r, int b, int c) { return K0[(3 * corner(a) + r) * 24 + 3 * corner(b) + c]; };
  bool refl = true;
  for (int a = 0; a < 8 && refl; ++a)
    for (int b = 0; b < 8 && refl; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          const double sr = ((a >> r) & 1) ? -1.0 : 1.0, sc = ((a >> c) & 1) ? -1.0 : 1.0;
          if (std::fabs(K(a, r, b, c) - sr * sc * K(0, r, a ^ b, c)) > tol) refl = false;
        }
  if (refl) flags |= K0S_REFLECT;
  bool wht = true;
  for (int M = 0; M < 8 && wht; ++M)
    for (int c = 0; c < 3; ++c)
      for (int N = 0; N < 8; ++N)
        for (int d = 0; d < 3; ++d) {
          double s = 0.0;
          for (int b = 0; b < 8; ++b)
            for (int b2 = 0; b2 < 8; ++b2) {
              const int sg = (__builtin_popcount(M & b) + __builtin_popcount(N & b2)) & 1;
              s += sg ? -K(b, c, b2, d) : K(b, c, b2, d);
            }
          const bool kept = N == (M ^ (1 << c) ^ (1 << d)) && g_kept(M, c, d);
          if (!kept && std::fabs(s / 64.0) > tol) wht = false;
        }
  if (wht) flags |= K0S_WHT;
  if (k0_fast3d(flags) && level1_terms_cover(K0, rtol)) flags |= K0S_LEVEL1;
  return flags;
}

}  // namespace et
