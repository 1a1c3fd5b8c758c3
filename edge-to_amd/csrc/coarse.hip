// Coarsest-grid direct solve in FP64: the coarsest Galerkin stencil is
// expanded to a dense SPD matrix, Cholesky-factored and inverted once per
// solve.  Three factor paths: the band Cholesky (in-house tile kernels,
// rocBLAS band updates) with L^-1 from rocSOLVER trtri (default; each
// V-cycle's coarse solve is two triangular products), rocSOLVER potrf +
// potri, or the in-house dense blocked Cholesky with A^-1 = L^-T L^-1 (one
// dense symmetric product per coarse solve for the last two).
// The matrix is padded to a multiple of 64 with an identity block.
#include <algorithm>
#include <type_traits>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.hpp"
#include "coarse.hpp"
#include "chol_tile.hpp"

namespace et {

constexpr int NB = 64;

// ------------------------------------------------ dense from block stencil ---

template <int DIM, int DOF>
__global__ void dense_from_stencil_kernel(const Grid g, const double* __restrict__ S, int ld,
                                          double* __restrict__ A) {
  constexpr int NO = DIM == 3 ? 27 : 9;
  const int64_t N = g.n_nodes();
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= N * NO) return;
  const int o = (int)(tid / N);
  const int64_t n = tid % N;
  const int oi = DIM == 3 ? o / 9 - 1 : o / 3 - 1;
  const int oj = DIM == 3 ? (o / 3) % 3 - 1 : o % 3 - 1;
  const int ok = DIM == 3 ? o % 3 - 1 : 0;
  const int nz = g.nz(), nx = g.nx();
  const int k = DIM == 3 ? (int)(n % nz) : 0;
  const int64_t t = DIM == 3 ? n / nz : n;
  const int i = (int)(t % nx), j = (int)(t / nx);
  const int mi = i + oi, mj = j + oj, mk = k + ok;
  if (mi < 0 || mi >= nx || mj < 0 || mj >= g.ny() || mk < 0 || mk >= nz) return;
  const int64_t m = DIM == 3 ? ((int64_t)mj * nx + mi) * nz + mk : (int64_t)mj * nx + mi;
  for (int r = 0; r < DOF; ++r)
    for (int c = 0; c < DOF; ++c)
      A[(n * DOF + r) * ld + m * DOF + c] = S[((int64_t)(o * DOF + r) * DOF + c) * N + n];
}

__global__ void pad_identity_kernel(int n, int ld, double* A) {
  const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ld) A[(int64_t)i * ld + i] = 1.0;
}

// ------------------------------------------------------ blocked Cholesky ---

// Factor the diagonal tile (kb, kb) in place (lower triangle).  256 threads.
__global__ __launch_bounds__(256) void potrf_diag_kernel(double* A, int ld, int kb, int* info) {
  __shared__ double T[NB][NB + 1];
  const int base = kb * NB;
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int r = t / NB, c = t % NB;
    T[r][c] = A[(int64_t)(base + r) * ld + base + c];
  }
  __syncthreads();
  for (int j = 0; j < NB; ++j) {
    if (threadIdx.x == 0) {
      const double d = T[j][j];
      if (!(d > 0.0)) *info = base + j + 1;
      T[j][j] = sqrt(d > 0.0 ? d : 1.0);
    }
    __syncthreads();
    const double piv = T[j][j];
    for (int r = j + 1 + threadIdx.x; r < NB; r += blockDim.x) T[r][j] /= piv;
    __syncthreads();
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
      const int r = t / NB, c = t % NB;
      if (c > j && r >= c) T[r][c] -= T[r][j] * T[c][j];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int r = t / NB, c = t % NB;
    A[(int64_t)(base + r) * ld + base + c] = r >= c ? T[r][c] : 0.0;
  }
}

// Panel: A_ik <- A_ik L_kk^-T for tiles i > kb.  64 threads, one row each.
__global__ __launch_bounds__(64) void trsm_panel_kernel(double* A, int ld, int kb) {
  __shared__ double L[NB][NB + 1];
  __shared__ double X[NB][NB + 1];
  const int ib = kb + 1 + blockIdx.x;
  const int base = kb * NB, rbase = ib * NB;
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int r = t / NB, c = t % NB;
    L[r][c] = A[(int64_t)(base + r) * ld + base + c];
    X[r][c] = A[(int64_t)(rbase + r) * ld + base + c];
  }
  __syncthreads();
  const int r = threadIdx.x;
  for (int j = 0; j < NB; ++j) {
    double v = X[r][j];
    for (int m = 0; m < j; ++m) v -= X[r][m] * L[j][m];
    X[r][j] = v / L[j][j];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int rr = t / NB, c = t % NB;
    A[(int64_t)(rbase + rr) * ld + base + c] = X[rr][c];
  }
}

__device__ __forceinline__ void tri_index(int t, int& i, int& j) {
  // t -> (i, j) with 0 <= j <= i
  i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  while (i * (i + 1) / 2 > t) --i;
  j = t - i * (i + 1) / 2;
}

// 64x64 tile product helper: C(4x4 per thread) += Atile * Btile^T, tiles in LDS.
__device__ __forceinline__ void tile_mma_nt(const double (*Ta)[NB + 1], const double (*Tb)[NB + 1],
                                            double (&acc)[4][4], int r0, int c0) {
  for (int m = 0; m < NB; ++m) {
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = Ta[r0 + q][m];
      b[q] = Tb[c0 + q][m];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
  }
}

// Trailing update: A_ij -= L_ik L_jk^T for kb < j <= i.  256 threads per tile.
__global__ __launch_bounds__(256) void syrk_update_kernel(double* A, int ld, int kb) {
  __shared__ double Ta[NB][NB + 1];
  __shared__ double Tb[NB][NB + 1];
  int ti, tj;
  tri_index(blockIdx.x, ti, tj);
  const int ib = kb + 1 + ti, jb = kb + 1 + tj;
  const int kbase = kb * NB;
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int r = t / NB, c = t % NB;
    Ta[r][c] = A[(int64_t)(ib * NB + r) * ld + kbase + c];
    Tb[r][c] = A[(int64_t)(jb * NB + r) * ld + kbase + c];
  }
  __syncthreads();
  const int r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
  double acc[4][4] = {};
  tile_mma_nt(Ta, Tb, acc, r0, c0);
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) A[(int64_t)(ib * NB + r0 + p) * ld + jb * NB + c0 + q] -= acc[p][q];
}

// ------------------------------------------------------- inverse of L ---

// Y = L^-1 by block rows: Y_ij = L_ii^-1 (delta_ij I - sum_{m=j}^{i-1} L_im Y_mj).
// Launch per block row ib with ib+1 blocks (one per column tile jb <= ib).
__global__ __launch_bounds__(256) void trtri_row_kernel(const double* L, double* Y, int ld, int ib) {
  __shared__ double Ta[NB][NB + 1];
  __shared__ double Tb[NB][NB + 1];
  const int jb = blockIdx.x;
  const int r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
  double acc[4][4] = {};
  for (int mb = jb; mb < ib; ++mb) {
    __syncthreads();
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
      const int r = t / NB, c = t % NB;
      Ta[r][c] = L[(int64_t)(ib * NB + r) * ld + mb * NB + c];
      // store Y_mj transposed so tile_mma_nt computes L_im * Y_mj
      Tb[c][r] = Y[(int64_t)(mb * NB + r) * ld + jb * NB + c];
    }
    __syncthreads();
    tile_mma_nt(Ta, Tb, acc, r0, c0);
  }
  __syncthreads();
  // B = delta I - acc, stored in Tb; L_ii in Ta
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int r = t / NB, c = t % NB;
    Ta[r][c] = L[(int64_t)(ib * NB + r) * ld + ib * NB + c];
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      Tb[r0 + p][c0 + q] = ((ib == jb && r0 + p == c0 + q) ? 1.0 : 0.0) - acc[p][q];
  __syncthreads();
  // forward substitution per column: threads 0..63 own one column each
  if (threadIdx.x < NB) {
    const int c = threadIdx.x;
    for (int r = 0; r < NB; ++r) {
      double v = Tb[r][c];
      for (int m = 0; m < r; ++m) v -= Ta[r][m] * Tb[m][c];
      Tb[r][c] = v / Ta[r][r];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int r = t / NB, c = t % NB;
    Y[(int64_t)(ib * NB + r) * ld + jb * NB + c] = Tb[r][c];
  }
}

// Ainv_ij = sum_{m >= i} Y_mi^T Y_mj for i >= j (tiles), mirrored to j > i.
__global__ __launch_bounds__(256) void yty_kernel(const double* Y, double* Ainv, int ld, int nbt) {
  __shared__ double Ta[NB][NB + 1];
  __shared__ double Tb[NB][NB + 1];
  int ib, jb;
  tri_index(blockIdx.x, ib, jb);
  const int r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
  double acc[4][4] = {};
  for (int mb = ib; mb < nbt; ++mb) {
    __syncthreads();
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
      const int r = t / NB, c = t % NB;
      Ta[c][r] = Y[(int64_t)(mb * NB + r) * ld + ib * NB + c];  // (Y_mi)^T
      Tb[c][r] = Y[(int64_t)(mb * NB + r) * ld + jb * NB + c];  // (Y_mj)^T
    }
    __syncthreads();
    tile_mma_nt(Ta, Tb, acc, r0, c0);
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gi = ib * NB + r0 + p, gj = jb * NB + c0 + q;
      Ainv[(int64_t)gi * ld + gj] = acc[p][q];
      Ainv[(int64_t)gj * ld + gi] = acc[p][q];
    }
}

// ------------------------------ band Cholesky + rocSOLVER triangular inverse ---
// Column-major view (the matrix is symmetric, so the dense row-major matrix
// is also its column-major self), 64x64 tiles, L lower.  The coarse matrix
// is banded: with node order (j nx + i) nz + k its half-bandwidth is
// b = dof (nx nz + nz + 1) + dof - 1, so tile (i, m) of A and of L is zero
// for i - m > nbd = ceil(b / 64), and each factor step k touches only the nr
// = min(nbd, tiles left) tiles below the diagonal:
//   diag:   one workgroup (four wavefronts) factors A_kk and forms W_kk = L_kk^-1 (chol_tile.hpp);
//   panel:  A_ik <- A_ik W_kk^T for the nr tiles below;
//   update: A_22 -= L_21 L_21^T on the nr x nr band block (one workgroup per
//           lower tile: one launch where rocBLAS dsyrk took five).
// rocSOLVER trtri then overwrites L with Y = L^-1, and the coarse solve is
// x = Y^T (Y b): two triangular products reading n^2 / 2 values each.

__global__ __launch_bounds__(kCholTileThreads) void bchol_diag_kernel(double* A, int ld, int o, double* Wk, int* info) {
  chol_tile64(A, ld, o, Wk, info);
}

// panel: tiles i = kb+1 .. kb+nr: A_ik <- A_ik W^T
__global__ __launch_bounds__(256) void bchol_panel_kernel(double* A, int ld, int kb, const double* Wk) {
  __shared__ double Xs[NB][NB + 1];  // [row][m]
  __shared__ double Ws[NB][NB + 1];  // [c][m] = W(c, m)
  const int o = kb * NB, row0 = o + NB * (1 + blockIdx.x);
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int rr = t % NB, cc = t / NB;  // consecutive threads, consecutive rows (column-major)
    Xs[rr][cc] = A[(int64_t)(o + cc) * ld + row0 + rr];
    Ws[rr][cc] = Wk[rr + cc * NB];
  }
  __syncthreads();
  const int r0 = (threadIdx.x % 16) * 4, c0 = (threadIdx.x / 16) * 4;
  double acc[4][4] = {};
  for (int m = 0; m < NB; ++m) {
    double xa[4], wb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xa[u] = Xs[r0 + u][m];
      wb[u] = Ws[c0 + u][m];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += xa[p] * wb[q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int p = 0; p < 4; ++p) A[(int64_t)(o + c0 + q) * ld + row0 + r0 + p] = acc[p][q];
}

// update: A_22 -= L_21 L_21^T on the lower tiles of the nr x nr band block below
// the diagonal tile kb (L_21: the nr panel tiles).  One workgroup per lower tile
// (i, j), j <= i; each sums over the 64 panel columns in order.  Upper halves of
// the diagonal tiles are left as they are (never read as L).
//
// The workgroup of tile (0, 0) -- the next diagonal tile -- then factors it
// (chol_tile64 into Wk), so the next step's diagonal factor overlaps the other
// workgroups' updates instead of waiting for a launch of its own.
__global__ __launch_bounds__(256) void bchol_syrk_kernel(double* A, int ld, int kb, double* Wk, int* info) {
  __shared__ double Xs[NB][NB + 1];  // L_i [row][m]
  __shared__ double Ys[NB][NB + 1];  // L_j [row][m]
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= (int)blockIdx.x) ++i;
  const int j = (int)blockIdx.x - i * (i + 1) / 2;
  const int o = kb * NB, ri = o + NB * (1 + i), rj = o + NB * (1 + j);
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int rr = t % NB, cc = t / NB;  // consecutive threads, consecutive rows (column-major)
    Xs[rr][cc] = A[(int64_t)(o + cc) * ld + ri + rr];
    Ys[rr][cc] = A[(int64_t)(o + cc) * ld + rj + rr];
  }
  __syncthreads();
  const int r0 = (threadIdx.x % 16) * 4, c0 = (threadIdx.x / 16) * 4;
  double acc[4][4] = {};
  for (int m = 0; m < NB; ++m) {
    double xa[4], yb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xa[u] = Xs[r0 + u][m];
      yb[u] = Ys[c0 + u][m];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += xa[p] * yb[q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (i > j || r0 + p >= c0 + q) A[(int64_t)(rj + c0 + q) * ld + ri + r0 + p] -= acc[p][q];
  if (blockIdx.x == 0) {
    __syncthreads();  // the updated tile is read back by other threads of this workgroup
    chol_tile64(A, ld, o + NB, Wk, info);
  }
}

// Z = Y row-major (Z[i ld + j] = Y_ij) for the tiles on and below the diagonal
// (above the diagonal of the diagonal tiles: zero).  One workgroup per tile.
__global__ __launch_bounds__(256) void lower_transpose_kernel(const double* __restrict__ Y, int ld,
                                                              double* __restrict__ Z) {
  __shared__ double Ts[NB][NB + 1];
  int ib, jb;
  tri_index(blockIdx.x, ib, jb);
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int rr = t % NB, cc = t / NB;
    const double v = Y[(int64_t)(jb * NB + cc) * ld + ib * NB + rr];
    Ts[rr][cc] = (ib == jb && cc > rr) ? 0.0 : v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int cc = t % NB, rr = t / NB;  // consecutive threads, consecutive columns of a Z row
    Z[(int64_t)(ib * NB + rr) * ld + jb * NB + cc] = Ts[rr][cc];
  }
}

// t_i = sum_{j <= i} Y_ij b_j  (row i of Y = row i of Z); one wavefront per row
template <typename TM, typename TI>
__global__ __launch_bounds__(256) void trmv_lower_kernel(const TM* __restrict__ Z, int ld, int n,
                                                         const TI* __restrict__ b, double* __restrict__ t) {
  using V2 = typename std::conditional<sizeof(TM) == 8, double2, float2>::type;
  const int i = n - 1 - (blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (i < 0) return;
  const TM* z = Z + (int64_t)i * ld;
  // pairs of consecutive values per lane (rows start aligned to the pair, ld even)
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int j = 2 * lane;
  for (; j + 257 <= i; j += 256) {
    const V2 za = *reinterpret_cast<const V2*>(z + j);
    const V2 zb = *reinterpret_cast<const V2*>(z + j + 128);
    s0 += za.x * (double)b[j];
    s1 += za.y * (double)b[j + 1];
    s2 += zb.x * (double)b[j + 128];
    s3 += zb.y * (double)b[j + 129];
  }
  for (; j <= i; j += 128) {
    s0 += (double)z[j] * (double)b[j];
    if (j + 1 <= i) s1 += (double)z[j + 1] * (double)b[j + 1];
  }
  const double s = wave_sum((s0 + s1) + (s2 + s3));
  if (lane == 0) t[i] = s;
}

// x_j = sum_{i >= j} Y_ij t_i  (column j of Y); one wavefront per column
template <typename TM, typename TO>
__global__ __launch_bounds__(256) void trmv_upper_kernel(const TM* __restrict__ Y, int ld, int n,
                                                         const double* __restrict__ t, TO* __restrict__ x) {
  using V2 = typename std::conditional<sizeof(TM) == 8, double2, float2>::type;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;
  const TM* y = Y + (int64_t)j * ld;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  // start at the even index at or below j; the entry above the diagonal (i = j - 1) is zero
  int i = (j & ~1) + 2 * lane;
  for (; i + 129 < n; i += 256) {
    const V2 ya = *reinterpret_cast<const V2*>(y + i);
    const V2 yb = *reinterpret_cast<const V2*>(y + i + 128);
    s0 += ya.x * t[i];
    s1 += ya.y * t[i + 1];
    s2 += yb.x * t[i + 128];
    s3 += yb.y * t[i + 129];
  }
  for (; i < n; i += 128) {
    s0 += (double)y[i] * t[i];
    if (i + 1 < n) s1 += (double)y[i + 1] * t[i + 1];
  }
  const double s = wave_sum((s0 + s1) + (s2 + s3));
  if (lane == 0) x[j] = (TO)s;
}

// -------------------------------------------------------- coarse solve ---

// x = Ainv b (first n entries); one wavefront per row.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void symv_kernel(const double* __restrict__ Ainv, int ld, int n,
                                                   const TI* __restrict__ b, TO* __restrict__ x) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  const double* a = Ainv + (int64_t)row * ld;
  for (int c = lane; c < n; c += 64) s += a[c] * (double)b[c];
  s = wave_sum(s);
  if (lane == 0) x[row] = (TO)s;
}

// full symmetric matrix from the column-major lower triangle rocSOLVER leaves
// (the row-major upper triangle): A[i][j] = A[j][i] for j < i
__global__ __launch_bounds__(256) void mirror_upper_kernel(double* A, int ld, int n) {
  const int i = blockIdx.x;
  const double* src = A + i;  // column i of the row-major matrix: A[j][i] at src[j * ld]
  for (int j = threadIdx.x; j < i && i < n; j += blockDim.x) A[(int64_t)i * ld + j] = src[(int64_t)j * ld];
}

static int g_coarse_lib = 2;
void set_coarse_lib(int mode) { g_coarse_lib = mode >= 0 && mode <= 2 ? mode : 2; }
int get_coarse_lib() { return g_coarse_lib; }

__global__ void f64_to_f32_kernel(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

int CoarseSolver::init(int n_, hipStream_t s, bool f32_copy) {
  n = n_;
  ld = (n + NB - 1) / NB * NB;
  nbt = ld / NB;
  lib_mode = g_coarse_lib;
  use_lib = lib_mode != 0;
  const size_t bytes = (size_t)ld * ld * sizeof(double);
  ET_CHECK_HIP(hipMalloc(&A, bytes));
  if (lib_mode == 0) ET_CHECK_HIP(hipMalloc(&Y, bytes));
  if (lib_mode == 2) {
    ET_CHECK_HIP(hipMalloc(&Z, bytes));
    ET_CHECK_HIP(hipMalloc(&Wd, (size_t)NB * NB * sizeof(double)));
    ET_CHECK_HIP(hipMalloc(&tv, (size_t)ld * sizeof(double)));
    if (f32_copy) {
      ET_CHECK_HIP(hipMalloc(&Zf, (size_t)ld * ld * sizeof(float)));
      ET_CHECK_HIP(hipMalloc(&Af, (size_t)ld * ld * sizeof(float)));
    }
  }
  ET_CHECK_HIP(hipMalloc(&info, 2 * sizeof(int)));
  ET_CHECK_HIP(hipHostMalloc(&info_h, 2 * sizeof(int)));
  if (use_lib) {
    rocblas_handle h = nullptr;
    ET_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, "rocblas_create_handle failed");
    blas = h;
  }
  (void)s;
  return ET_OK;
}

void CoarseSolver::release() {
  if (blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(blas));
  blas = nullptr;
  if (A) (void)hipFree(A);
  if (Y) (void)hipFree(Y);
  if (Wd) (void)hipFree(Wd);
  if (tv) (void)hipFree(tv);
  if (Z) (void)hipFree(Z);
  if (Zf) (void)hipFree(Zf);
  if (Af) (void)hipFree(Af);
  Wd = tv = Z = nullptr;
  Zf = Af = nullptr;
  if (info) (void)hipFree(info);
  if (info_h) (void)hipHostFree(info_h);
  A = Y = nullptr;
  info = nullptr;
  info_h = nullptr;
}

int CoarseSolver::factor(const Grid& g, int dof, const double* S, hipStream_t s) {
  const size_t bytes = (size_t)ld * ld * sizeof(double);
  ET_CHECK_HIP(hipMemsetAsync(A, 0, bytes, s));
  if (Y && lib_mode == 0) ET_CHECK_HIP(hipMemsetAsync(Y, 0, bytes, s));
  ET_CHECK_HIP(hipMemsetAsync(info, 0, 2 * sizeof(int), s));
  const int NO = g.dim == 3 ? 27 : 9;
  const unsigned grid = grid_for(g.n_nodes() * NO);
  if (g.dim == 3)
    dense_from_stencil_kernel<3, 3><<<grid, kBlock, 0, s>>>(g, S, ld, A);
  else
    dense_from_stencil_kernel<2, 2><<<grid, kBlock, 0, s>>>(g, S, ld, A);
  ET_CHECK_LAUNCH();
  if (ld > n) {
    pad_identity_kernel<<<grid_for(ld - n), kBlock, 0, s>>>(n, ld, A);
    ET_CHECK_LAUNCH();
  }
  if (lib_mode == 2) {
    rocblas_handle h = static_cast<rocblas_handle>(blas);
    ET_REQUIRE(rocblas_set_stream(h, s) == rocblas_status_success, "rocblas_set_stream failed");
    // half-bandwidth of the node ordering (j nx + i) nz + k, in tiles
    const int64_t bw = g.dim == 3 ? (int64_t)dof * ((int64_t)g.nx() * g.nz() + g.nz() + 1) + dof - 1
                                  : (int64_t)dof * (g.nx() + 1) + dof - 1;
    const int nbd = (int)std::min<int64_t>((bw + NB - 1) / NB, nbt - 1);
    // tile 0 here; tile kb + 1 by the band update of step kb
    bchol_diag_kernel<<<1, kCholTileThreads, 0, s>>>(A, ld, 0, Wd, info);
    ET_CHECK_LAUNCH();
    for (int kb = 0; kb < nbt; ++kb) {
      const int nr = std::min(nbd, nbt - kb - 1);
      if (nr > 0) {
        bchol_panel_kernel<<<nr, 256, 0, s>>>(A, ld, kb, Wd);
        ET_CHECK_LAUNCH();
        bchol_syrk_kernel<<<nr * (nr + 1) / 2, kCholTileThreads, 0, s>>>(A, ld, kb, Wd, info);
        ET_CHECK_LAUNCH();
      }
    }
    ET_REQUIRE(rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, ld, A, ld, info + 1) ==
                   rocblas_status_success,
               "rocsolver_dtrtri failed");
    lower_transpose_kernel<<<nbt * (nbt + 1) / 2, 256, 0, s>>>(A, ld, Z);
    ET_CHECK_LAUNCH();
    if (Zf != nullptr) {
      const int64_t cnt = (int64_t)ld * ld;
      f64_to_f32_kernel<<<grid_for(cnt), kBlock, 0, s>>>(cnt, A, Af);
      f64_to_f32_kernel<<<grid_for(cnt), kBlock, 0, s>>>(cnt, Z, Zf);
      ET_CHECK_LAUNCH();
    }
    ET_CHECK_HIP(hipMemcpyAsync(info_h, info, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    return ET_OK;
  }
  if (use_lib) {
    rocblas_handle h = static_cast<rocblas_handle>(blas);
    ET_REQUIRE(rocblas_set_stream(h, s) == rocblas_status_success, "rocblas_set_stream failed");
    ET_REQUIRE(rocsolver_dpotrf(h, rocblas_fill_lower, ld, A, ld, info) == rocblas_status_success,
               "rocsolver_dpotrf failed");
    ET_REQUIRE(rocsolver_dpotri(h, rocblas_fill_lower, ld, A, ld, info + 1) == rocblas_status_success,
               "rocsolver_dpotri failed");
    mirror_upper_kernel<<<ld, 256, 0, s>>>(A, ld, ld);
    ET_CHECK_LAUNCH();
    ET_CHECK_HIP(hipMemcpyAsync(info_h, info, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    return ET_OK;
  }
  for (int kb = 0; kb < nbt; ++kb) {
    potrf_diag_kernel<<<1, 256, 0, s>>>(A, ld, kb, info);
    ET_CHECK_LAUNCH();
    const int rest = nbt - kb - 1;
    if (rest > 0) {
      trsm_panel_kernel<<<rest, 64, 0, s>>>(A, ld, kb);
      ET_CHECK_LAUNCH();
      syrk_update_kernel<<<rest * (rest + 1) / 2, 256, 0, s>>>(A, ld, kb);
      ET_CHECK_LAUNCH();
    }
  }
  // A now holds L in its lower triangle (upper part stale); Y = L^-1
  for (int ib = 0; ib < nbt; ++ib) {
    trtri_row_kernel<<<ib + 1, 256, 0, s>>>(A, Y, ld, ib);
    ET_CHECK_LAUNCH();
  }
  // A <- Y^T Y = A_coarse^-1
  yty_kernel<<<nbt * (nbt + 1) / 2, 256, 0, s>>>(Y, A, ld, nbt);
  ET_CHECK_LAUNCH();
  ET_CHECK_HIP(hipMemcpyAsync(info_h, info, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  return ET_OK;
}

int CoarseSolver::check_info(hipStream_t s) {
  ET_CHECK_HIP(hipStreamSynchronize(s));
  if (info_h[0] != 0) {
    set_error("coarse-grid Cholesky: matrix not positive definite at row " + std::to_string(info_h[0] - 1));
    return ET_ENONFINITE;
  }
  if (info_h[1] != 0) {
    set_error("coarse-grid inverse: singular factor at row " + std::to_string(info_h[1] - 1));
    return ET_ENONFINITE;
  }
  return ET_OK;
}

template <typename TI, typename TO>
int CoarseSolver::solve(const TI* b, TO* x, hipStream_t s) const {
  if (lib_mode == 2) {
    if (sizeof(TO) == 4 && Zf != nullptr) {
      // FP32 V-cycles apply the FP32 copy of the factors (half the bytes), FP64 sums
      trmv_lower_kernel<float, TI><<<(n + 3) / 4, 256, 0, s>>>(Zf, ld, n, b, tv);
      ET_CHECK_LAUNCH();
      trmv_upper_kernel<float, TO><<<(n + 3) / 4, 256, 0, s>>>(Af, ld, n, tv, x);
    } else {
      trmv_lower_kernel<double, TI><<<(n + 3) / 4, 256, 0, s>>>(Z, ld, n, b, tv);
      ET_CHECK_LAUNCH();
      trmv_upper_kernel<double, TO><<<(n + 3) / 4, 256, 0, s>>>(A, ld, n, tv, x);
    }
    ET_CHECK_LAUNCH();
    return ET_OK;
  }
  symv_kernel<TI, TO><<<(n + 3) / 4, 256, 0, s>>>(A, ld, n, b, x);
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int CoarseSolver::solve<double, double>(const double*, double*, hipStream_t) const;
template int CoarseSolver::solve<float, float>(const float*, float*, hipStream_t) const;

}  // namespace et
