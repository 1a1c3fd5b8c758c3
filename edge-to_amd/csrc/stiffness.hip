// Structured Q4/H8 stiffness operator kernels: connectivity, constrained
// matrix-free matvec, diagonal, element energy / SIMP sensitivity and the
// fine-grid block stencil used to seed the Galerkin hierarchy.
//
// The operator is A_hat = M K(E) M + I_c with K(E) = sum_e E_e Q_e^T K0 Q_e.
// Kernels are node-centric: one thread owns one node and gathers the 2^dim
// elements around it, so every output is written exactly once, the summation
// order is fixed (run-to-run reproducible) and no atomics are needed.  K0 is
// indexed with compile-time offsets only, so its entries are fetched with
// scalar loads shared by the whole wavefront.
#include "common.hpp"
#include "stiffness.hpp"

namespace et {

// ----------------------------------------------------------- connectivity ---

__global__ void gen_elements_2d(int nelx, int nely, int32_t* __restrict__ el) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nelx * nely) return;
  const int nx = nelx + 1;
  const int i = (int)(e / nely), j = (int)(e % nely);
  int32_t* o = el + 4 * e;
  o[0] = j * nx + i;
  o[1] = j * nx + i + 1;
  o[2] = (j + 1) * nx + i + 1;
  o[3] = (j + 1) * nx + i;
}

__global__ void gen_elements_3d(int nelx, int nely, int nelz, int32_t* __restrict__ el) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nelx * nely * nelz) return;
  const int nx = nelx + 1, nz = nelz + 1;
  const int i = (int)(e / ((int64_t)nely * nelz));
  const int r = (int)(e % ((int64_t)nely * nelz));
  const int j = r / nelz, k = r % nelz;
  const int b = nz * i + nz * nx * j + k;
  int32_t* o = el + 8 * e;
  o[0] = b;
  o[1] = b + nz;
  o[2] = b + nz + nz * nx;
  o[3] = b + nz * nx;
  o[4] = b + 1;
  o[5] = b + nz + 1;
  o[6] = b + nz + nz * nx + 1;
  o[7] = b + nz * nx + 1;
}

// ---------------------------------------------------------------- matvec ---

// Accumulate (M K(E) M u) at node (i, j[, k]) into out[0..dim).  cons may be null.
template <typename T, int DIM>
__device__ __forceinline__ void node_Ku(const Grid g, int i, int j, int k, const T* __restrict__ K0,
                                        const T* __restrict__ E, const uint8_t* __restrict__ cons,
                                        const T* __restrict__ u, T (&out)[DIM]) {
  constexpr int NE = DIM == 3 ? 8 : 4;
  constexpr int ND = NE * DIM;
  const int nx = g.nx(), nz = g.nz();
#pragma unroll
  for (int r = 0; r < DIM; ++r) out[r] = T(0);
#pragma unroll
  for (int a = 0; a < NE; ++a) {
    const int ox = DIM == 3 ? hx(a) : qx(a);
    const int oy = DIM == 3 ? hy(a) : qy(a);
    const int oz = DIM == 3 ? hz(a) : 0;
    const int ex = i - ox, ey = j - oy, ez = k - oz;
    bool valid = ex >= 0 && ex < g.ex && ey >= 0 && ey < g.ey;
    if (DIM == 3) valid = valid && ez >= 0 && ez < g.ez;
    if (!valid) continue;
    const int64_t e = DIM == 3 ? ((int64_t)ex * g.ey + ey) * g.ez + ez : (int64_t)ex * g.ey + ey;
    const T Ee = E[e];
    T s[DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r) s[r] = T(0);
#pragma unroll
    for (int b = 0; b < NE; ++b) {
      const int bx = ex + (DIM == 3 ? hx(b) : qx(b));
      const int by = ey + (DIM == 3 ? hy(b) : qy(b));
      const int bz = ez + (DIM == 3 ? hz(b) : 0);
      const int64_t m = DIM == 3 ? ((int64_t)by * nx + bx) * nz + bz : (int64_t)by * nx + bx;
      T um[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        um[c] = u[m * DIM + c];
        if (cons != nullptr && cons[m * DIM + c]) um[c] = T(0);
      }
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) s[r] += K0[(DIM * a + r) * ND + DIM * b + c] * um[c];
    }
#pragma unroll
    for (int r = 0; r < DIM; ++r) out[r] += Ee * s[r];
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void matvec_kernel(const Grid g, const T* __restrict__ K0,
                                                     const T* __restrict__ E,
                                                     const uint8_t* __restrict__ cons,
                                                     const T* __restrict__ u, T* __restrict__ y) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes()) return;
  const int nz = g.nz(), nx = g.nx();
  const int k = DIM == 3 ? (int)(n % nz) : 0;
  const int64_t t = DIM == 3 ? n / nz : n;
  const int i = (int)(t % nx), j = (int)(t / nx);
  T acc[DIM];
  node_Ku<T, DIM>(g, i, j, k, K0, E, cons, u, acc);
#pragma unroll
  for (int r = 0; r < DIM; ++r) {
    const int64_t d = n * DIM + r;
    y[d] = (cons != nullptr && cons[d]) ? u[d] : acc[r];
  }
}

template <typename T>
int matvec(const Grid& g, const T* K0, const T* E, const uint8_t* cons, const T* u, T* y,
           hipStream_t s, const WhtTable* fine) {
  const unsigned grid = grid_for(g.n_nodes());
  if (g.dim == 3 && fine != nullptr && fine_supported(g, sizeof(T)))
    return op0_fine<T>(0, g, (int64_t)g.ey * g.ez, g.ez, K0, *fine, E, cons, u, nullptr, nullptr, nullptr,
                        1.0, nullptr, y, s);
  if (g.dim == 3)
    matvec_kernel<T, 3><<<grid, kBlock, 0, s>>>(g, K0, E, cons, u, y);
  else
    matvec_kernel<T, 2><<<grid, kBlock, 0, s>>>(g, K0, E, cons, u, y);
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int matvec<double>(const Grid&, const double*, const double*, const uint8_t*,
                            const double*, double*, hipStream_t, const WhtTable*);
template int matvec<float>(const Grid&, const float*, const float*, const uint8_t*, const float*,
                           float*, hipStream_t, const WhtTable*);

// -------------------------------------------------------------- diagonal ---

template <typename T, int DIM>
__global__ __launch_bounds__(256) void diagonal_kernel(const Grid g, const T* __restrict__ K0,
                                                       const T* __restrict__ E,
                                                       const uint8_t* __restrict__ cons,
                                                       T* __restrict__ diag) {
  constexpr int NE = DIM == 3 ? 8 : 4;
  constexpr int ND = NE * DIM;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes()) return;
  const int nz = g.nz(), nx = g.nx();
  const int k = DIM == 3 ? (int)(n % nz) : 0;
  const int64_t t = DIM == 3 ? n / nz : n;
  const int i = (int)(t % nx), j = (int)(t / nx);
  T acc[DIM];
#pragma unroll
  for (int r = 0; r < DIM; ++r) acc[r] = T(0);
#pragma unroll
  for (int a = 0; a < NE; ++a) {
    const int ex = i - (DIM == 3 ? hx(a) : qx(a));
    const int ey = j - (DIM == 3 ? hy(a) : qy(a));
    const int ez = k - (DIM == 3 ? hz(a) : 0);
    bool valid = ex >= 0 && ex < g.ex && ey >= 0 && ey < g.ey;
    if (DIM == 3) valid = valid && ez >= 0 && ez < g.ez;
    if (!valid) continue;
    const int64_t e = DIM == 3 ? ((int64_t)ex * g.ey + ey) * g.ez + ez : (int64_t)ex * g.ey + ey;
    const T Ee = E[e];
#pragma unroll
    for (int r = 0; r < DIM; ++r) acc[r] += K0[(DIM * a + r) * ND + DIM * a + r] * Ee;
  }
#pragma unroll
  for (int r = 0; r < DIM; ++r) {
    const int64_t d = n * DIM + r;
    diag[d] = (cons != nullptr && cons[d]) ? T(1) : acc[r];
  }
}

template <typename T>
int diagonal(const Grid& g, const T* K0, const T* E, const uint8_t* cons, T* diag, hipStream_t s) {
  const unsigned grid = grid_for(g.n_nodes());
  if (g.dim == 3)
    diagonal_kernel<T, 3><<<grid, kBlock, 0, s>>>(g, K0, E, cons, diag);
  else
    diagonal_kernel<T, 2><<<grid, kBlock, 0, s>>>(g, K0, E, cons, diag);
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int diagonal<double>(const Grid&, const double*, const double*, const uint8_t*, double*,
                              hipStream_t);

// ------------------------------------------- element energy / sensitivity ---

// mode 0: out = -u^T K0 u ; mode 1: out = p*clip(xt,void,1)^(p-1) * (-u^T K0 u)
template <int DIM, int MODE>
__global__ __launch_bounds__(256) void energy_kernel(const Grid g, const double* __restrict__ K0,
                                                     const double* __restrict__ U,
                                                     const double* __restrict__ xt, double penal,
                                                     double vmin, double* __restrict__ out) {
  constexpr int NE = DIM == 3 ? 8 : 4;
  constexpr int ND = NE * DIM;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_elem()) return;
  int ex, ey, ez = 0;
  if (DIM == 3) {
    ex = (int)(e / ((int64_t)g.ey * g.ez));
    const int r = (int)(e % ((int64_t)g.ey * g.ez));
    ey = r / g.ez;
    ez = r % g.ez;
  } else {
    ex = (int)(e / g.ey);
    ey = (int)(e % g.ey);
  }
  const int nx = g.nx(), nz = g.nz();
  double ue[ND];
#pragma unroll
  for (int b = 0; b < NE; ++b) {
    const int bx = ex + (DIM == 3 ? hx(b) : qx(b));
    const int by = ey + (DIM == 3 ? hy(b) : qy(b));
    const int bz = ez + (DIM == 3 ? hz(b) : 0);
    const int64_t m = DIM == 3 ? ((int64_t)by * nx + bx) * nz + bz : (int64_t)by * nx + bx;
#pragma unroll
    for (int c = 0; c < DIM; ++c) ue[DIM * b + c] = U[m * DIM + c];
  }
  double acc = 0.0;
#pragma unroll
  for (int p = 0; p < ND; ++p) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < ND; ++q) v += K0[p * ND + q] * ue[q];
    acc -= v * ue[p];
  }
  if (MODE == 1) {
    const double x = fmin(fmax(xt[e], vmin), 1.0);
    acc *= penal * pow(x, penal - 1.0);
  }
  out[e] = acc;
}

int elem_energy(const Grid& g, const double* K0, const double* U, const double* xt, double penal,
                double vmin, int mode, double* out, hipStream_t s) {
  const unsigned grid = grid_for(g.n_elem());
  if (g.dim == 3) {
    if (mode == 0)
      energy_kernel<3, 0><<<grid, kBlock, 0, s>>>(g, K0, U, xt, penal, vmin, out);
    else
      energy_kernel<3, 1><<<grid, kBlock, 0, s>>>(g, K0, U, xt, penal, vmin, out);
  } else {
    if (mode == 0)
      energy_kernel<2, 0><<<grid, kBlock, 0, s>>>(g, K0, U, xt, penal, vmin, out);
    else
      energy_kernel<2, 1><<<grid, kBlock, 0, s>>>(g, K0, U, xt, penal, vmin, out);
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

__global__ void simp_forward_kernel(int64_t n, const double* __restrict__ xt, double penal,
                                    double vmin, int plain, double* __restrict__ E) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double x = xt[e];
  const double v = plain ? x : pow(x, penal);
  E[e] = fmin(fmax(v, vmin), 1.0);
}

int simp_forward(int64_t n, const double* xt, double penal, double vmin, int plain, double* E,
                 hipStream_t s) {
  simp_forward_kernel<<<grid_for(n), kBlock, 0, s>>>(n, xt, penal, vmin, plain, E);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

// ------------------------------------------------ fine-grid block stencil ---

// One thread per (node, neighbour offset): the dof x dof block A_hat[n, n+o].
template <int DIM>
__global__ __launch_bounds__(256) void stencil0_kernel(const Grid g, const double* __restrict__ K0,
                                                       const double* __restrict__ E,
                                                       const uint8_t* __restrict__ cons,
                                                       double* __restrict__ S) {
  constexpr int NE = DIM == 3 ? 8 : 4;
  constexpr int ND = NE * DIM;
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
  double blk[DIM][DIM];
#pragma unroll
  for (int r = 0; r < DIM; ++r)
#pragma unroll
    for (int c = 0; c < DIM; ++c) blk[r][c] = 0.0;
  const bool in = mi >= 0 && mi < nx && mj >= 0 && mj < g.ny() && mk >= 0 && mk < nz;
  if (in) {
    const int64_t m = DIM == 3 ? ((int64_t)mj * nx + mi) * nz + mk : (int64_t)mj * nx + mi;
#pragma unroll
    for (int a = 0; a < NE; ++a) {
      const int ax = DIM == 3 ? hx(a) : qx(a), ay = DIM == 3 ? hy(a) : qy(a);
      const int az = DIM == 3 ? hz(a) : 0;
      const int bx = ax + oi, by = ay + oj, bz = az + ok;
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
      const int b = DIM == 3 ? (bz ? 4 : 0) + (by ? (bx ? 2 : 3) : (bx ? 1 : 0))
                             : (by ? (bx ? 2 : 3) : (bx ? 1 : 0));
      const int ex = i - ax, ey = j - ay, ez = k - az;
      bool valid = ex >= 0 && ex < g.ex && ey >= 0 && ey < g.ey;
      if (DIM == 3) valid = valid && ez >= 0 && ez < g.ez;
      if (!valid) continue;
      const int64_t e = DIM == 3 ? ((int64_t)ex * g.ey + ey) * g.ez + ez : (int64_t)ex * g.ey + ey;
      const double Ee = E[e];
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) blk[r][c] += K0[(DIM * a + r) * ND + DIM * b + c] * Ee;
    }
    if (cons != nullptr) {
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          if (cons[n * DIM + r] || cons[m * DIM + c]) blk[r][c] = 0.0;
      if (oi == 0 && oj == 0 && ok == 0)
#pragma unroll
        for (int r = 0; r < DIM; ++r)
          if (cons[n * DIM + r]) blk[r][r] = 1.0;
    }
  }
#pragma unroll
  for (int r = 0; r < DIM; ++r)
#pragma unroll
    for (int c = 0; c < DIM; ++c) S[((int64_t)(o * DIM + r) * DIM + c) * N + n] = blk[r][c];
}

int stencil_from_operator(const Grid& g, const double* K0, const double* E, const uint8_t* cons,
                          double* S, hipStream_t s) {
  const int NO = g.dim == 3 ? 27 : 9;
  const unsigned grid = grid_for(g.n_nodes() * NO);
  if (g.dim == 3)
    stencil0_kernel<3><<<grid, kBlock, 0, s>>>(g, K0, E, cons, S);
  else
    stencil0_kernel<2><<<grid, kBlock, 0, s>>>(g, K0, E, cons, S);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int generate_elements(const Grid& g, int32_t* el, hipStream_t s) {
  const unsigned grid = grid_for(g.n_elem());
  if (g.dim == 3)
    gen_elements_3d<<<grid, kBlock, 0, s>>>(g.ex, g.ey, g.ez, el);
  else
    gen_elements_2d<<<grid, kBlock, 0, s>>>(g.ex, g.ey, el);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

}  // namespace et
