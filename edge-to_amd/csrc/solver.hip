// GMG-preconditioned conjugate gradients for A_hat(E) u = f on a structured
// grid: the behaviour of CuMultiGrid.solve (solvers/_multigrid_cuda.py:188-242)
// with the hierarchy built on the device.
//
//   level 0   matrix-free operator (fused Jacobi / residual kernels)
//   level l   Galerkin block stencil R S_{l-1} R^T, built every solve
//   level L   dense FP64 inverse (CoarseSolver)
//
// The PCG loop keeps alpha, beta, omega and the dot products in device
// memory; the host reads |r|^2 once per iteration for the convergence test.
// precision 1 runs the V-cycle in FP32 (operators cast from the FP64
// hierarchy) inside the FP64 outer loop.
#include <algorithm>
#include <cmath>
#include <vector>

#include "blas.hpp"
#include "coarse.hpp"
#include "common.hpp"
#include "galerkin_elem.hpp"
#include "solver.hpp"
#include "stiffness.hpp"

namespace et {

// ------------------------------------------------------------ kernels ----

// OP_JAC0: first Jacobi sweep from the zero guess fused with it (fine grid,
// transformed-basis operator only); x is not read
enum OpMode { OP_Y = 0, OP_RES = 1, OP_JAC = 2, OP_JAC_ADD = 3, OP_JAC0 = 5 };

template <typename T, int DIM, int MODE>
__device__ __forceinline__ void op_epilogue(int64_t n, const T (&ax)[DIM],
                                            const T* __restrict__ x, const T* __restrict__ b,
                                            const T* __restrict__ Dinv, T w,
                                            const T* __restrict__ xadd, T* __restrict__ y) {
#pragma unroll
  for (int r = 0; r < DIM; ++r) {
    const int64_t d = n * DIM + r;
    T v;
    if (MODE == OP_Y)
      v = ax[r];
    else if (MODE == OP_RES)
      v = b[d] - ax[r];
    else
      v = x[d] + w * Dinv[d] * (b[d] - ax[r]);
    if (MODE == OP_JAC_ADD) v += xadd[d];
    y[d] = v;
  }
}

// Level-0 operator, matrix free.  omega is read from device memory.
template <typename T, int DIM, int MODE>
__global__ __launch_bounds__(256) void op0_kernel(const Grid g, const T* __restrict__ K0,
                                                  const T* __restrict__ E,
                                                  const uint8_t* __restrict__ cons,
                                                  const T* __restrict__ x, const T* __restrict__ b,
                                                  const T* __restrict__ Dinv,
                                                  const double* __restrict__ omega, double wscale,
                                                  const T* __restrict__ xadd,
                                                  T* __restrict__ y) {
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
    T s[DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r) s[r] = T(0);
#pragma unroll
    for (int b2 = 0; b2 < NE; ++b2) {
      const int bx = ex + (DIM == 3 ? hx(b2) : qx(b2));
      const int by = ey + (DIM == 3 ? hy(b2) : qy(b2));
      const int bz = ez + (DIM == 3 ? hz(b2) : 0);
      const int64_t m = DIM == 3 ? ((int64_t)by * nx + bx) * nz + bz : (int64_t)by * nx + bx;
      T um[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) um[c] = (cons != nullptr && cons[m * DIM + c]) ? T(0) : x[m * DIM + c];
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) s[r] += K0[(DIM * a + r) * ND + DIM * b2 + c] * um[c];
    }
#pragma unroll
    for (int r = 0; r < DIM; ++r) acc[r] += Ee * s[r];
  }
#pragma unroll
  for (int r = 0; r < DIM; ++r)
    if (cons != nullptr && cons[n * DIM + r]) acc[r] = x[n * DIM + r];
  const T w = MODE >= OP_JAC ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);
  op_epilogue<T, DIM, MODE>(n, acc, x, b, Dinv, w, xadd, y);
}

// Coarse level operator: symmetric block stencil stored by halves.  S holds
// the centre and the forward offsets o = centre+1 .. 26 (3-D; 5 .. 8 in 2-D)
// of the full offset numbering o = (di+1)*9 + (dj+1)*3 + (dk+1); the backward
// blocks follow from symmetry, A(n, n-o) = A(n-o, n)^T = S_o(n-o)^T.
// Layout S[((h*dof + r)*dof + c)*N + n], h = o - centre.
template <typename T, int DIM, int MODE>
__global__ __launch_bounds__(256) void opS_kernel(const Grid g, const T* __restrict__ S,
                                                  const T* __restrict__ x, const T* __restrict__ b,
                                                  const T* __restrict__ Dinv,
                                                  const double* __restrict__ omega, double wscale,
                                                  const T* __restrict__ xadd,
                                                  T* __restrict__ y) {
  constexpr int OC = DIM == 3 ? 13 : 4;  // centre offset
  constexpr int NH = OC + 1;             // stored blocks per node
  const int64_t N = g.n_nodes();
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int nz = g.nz(), nx = g.nx(), ny = g.ny();
  const int k = DIM == 3 ? (int)(n % nz) : 0;
  const int64_t t = DIM == 3 ? n / nz : n;
  const int i = (int)(t % nx), j = (int)(t / nx);
  T acc[DIM];
#pragma unroll
  for (int r = 0; r < DIM; ++r) acc[r] = T(0);
  // two offsets per step: ~50 independent loads in flight without spilling
#pragma unroll 2
  for (int h = 0; h < NH; ++h) {
    const int o = OC + h;
    const int oi = DIM == 3 ? o / 9 - 1 : o / 3 - 1;
    const int oj = DIM == 3 ? (o / 3) % 3 - 1 : o % 3 - 1;
    const int ok = DIM == 3 ? o % 3 - 1 : 0;
    const int64_t off = DIM == 3 ? ((int64_t)oj * nx + oi) * nz + ok : (int64_t)oj * nx + oi;
    // Branch-free: an out-of-grid neighbour reads node n itself with its x
    // weighted by zero, so all loads of the thread can be in flight at once.
    // forward block S_o(n) x_{n+o}
    {
      const int mi = i + oi, mj = j + oj, mk = k + ok;
      const bool in = mi >= 0 && mi < nx && mj >= 0 && mj < ny && (DIM == 2 || (mk >= 0 && mk < nz));
      const int64_t m = in ? n + off : n;
#pragma unroll
      for (int q = 0; q < DIM; ++q) {
        const T xv = in ? x[m * DIM + q] : T(0);
#pragma unroll
        for (int r = 0; r < DIM; ++r) acc[r] += S[((int64_t)(h * DIM + r) * DIM + q) * N + n] * xv;
      }
    }
    // backward block S_o(n-o)^T x_{n-o}
    if (h > 0) {
      const int mi = i - oi, mj = j - oj, mk = k - ok;
      const bool in = mi >= 0 && mi < nx && mj >= 0 && mj < ny && (DIM == 2 || (mk >= 0 && mk < nz));
      const int64_t m = in ? n - off : n;
#pragma unroll
      for (int q = 0; q < DIM; ++q) {
        const T xv = in ? x[m * DIM + q] : T(0);
#pragma unroll
        for (int r = 0; r < DIM; ++r) acc[r] += S[((int64_t)(h * DIM + q) * DIM + r) * N + m] * xv;
      }
    }
  }
  const T w = MODE >= OP_JAC ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);
  op_epilogue<T, DIM, MODE>(n, acc, x, b, Dinv, w, xadd, y);
}

// The same operator with one thread per (node, component): three times the
// threads of opS_kernel, for the small coarse levels where one thread per node
// leaves the GPU latency-bound.  Output component r = t / N of node n = t % N.
template <typename T, int DIM, int MODE>
__global__ __launch_bounds__(256) void opS_row_kernel(const Grid g, const T* __restrict__ S,
                                                      const T* __restrict__ x, const T* __restrict__ b,
                                                      const T* __restrict__ Dinv,
                                                      const double* __restrict__ omega, double wscale,
                                                      const T* __restrict__ xadd,
                                                      T* __restrict__ y) {
  constexpr int OC = DIM == 3 ? 13 : 4;
  constexpr int NH = OC + 1;
  const int N = (int)g.n_nodes();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * DIM) return;
  const int r = t / N, n = t - r * N;
  const int nz = g.nz(), nx = g.nx(), ny = g.ny();
  const int k = DIM == 3 ? n % nz : 0;
  const int ij = DIM == 3 ? n / nz : n;
  const int i = ij % nx, j = ij / nx;
  T acc = T(0);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int o = OC + h;
    const int oi = DIM == 3 ? o / 9 - 1 : o / 3 - 1;
    const int oj = DIM == 3 ? (o / 3) % 3 - 1 : o % 3 - 1;
    const int ok = DIM == 3 ? o % 3 - 1 : 0;
    const int off = DIM == 3 ? (oj * nx + oi) * nz + ok : oj * nx + oi;
    // branch-free as in opS_kernel
    {
      const int mi = i + oi, mj = j + oj, mk = k + ok;
      const bool in = mi >= 0 && mi < nx && mj >= 0 && mj < ny && (DIM == 2 || (mk >= 0 && mk < nz));
      const int m = in ? n + off : n;
#pragma unroll
      for (int q = 0; q < DIM; ++q) {
        const T xv = in ? ld32(x, m * DIM, q) : T(0);
        acc += ld32(S, ((h * DIM + r) * DIM + q) * N + n) * xv;
      }
    }
    if (h > 0) {
      const int mi = i - oi, mj = j - oj, mk = k - ok;
      const bool in = mi >= 0 && mi < nx && mj >= 0 && mj < ny && (DIM == 2 || (mk >= 0 && mk < nz));
      const int m = in ? n - off : n;
#pragma unroll
      for (int q = 0; q < DIM; ++q) {
        const T xv = in ? ld32(x, m * DIM, q) : T(0);
        acc += ld32(S, ((h * DIM + q) * DIM + r) * N + m) * xv;
      }
    }
  }
  const int d = n * DIM + r;
  const T w = MODE >= OP_JAC ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);
  T v;
  if (MODE == OP_Y)
    v = acc;
  else if (MODE == OP_RES)
    v = ld32(b, d) - acc;
  else
    v = ld32(x, d) + w * ld32(Dinv, d) * (ld32(b, d) - acc);
  if (MODE == OP_JAC_ADD) v += ld32(xadd, d);
  st32(y, d, 0, v);
}

// coarse levels up to this size use opS_row_kernel (32-bit stencil offsets: also below 2^31 / 126 nodes)
static int64_t g_row_stencil_max_nodes = 300000;
void set_row_stencil_nodes(int64_t n) { g_row_stencil_max_nodes = n; }
static int g_fused_rz = 1;
void set_fused_rz(int on) { g_fused_rz = on ? 1 : 0; }
int get_fused_rz() { return g_fused_rz; }
int64_t get_row_stencil_nodes() { return g_row_stencil_max_nodes; }

// y = omega*wscale * Dinv * b  (first Jacobi sweep from a zero guess)
template <typename T>
__global__ void scale_kernel(int64_t n, const T* __restrict__ Dinv, const T* __restrict__ b,
                             const double* __restrict__ omega, double wscale, T* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const T w = (T)(*omega * wscale);
  y[t] = w * Dinv[t] * b[t];
}

template <typename TI, typename TO>
__global__ void convert_kernel(int64_t n, const TI* __restrict__ a, TO* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (TO)a[t];
}

template <typename T>
__global__ void axpy_kernel(int64_t n, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] += x[t];
}

template <typename T>
__global__ void recip_kernel(int64_t n, const double* __restrict__ d, T* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (T)(1.0 / d[t]);
}

// inverse diagonal of a block stencil (centre block diagonal)
template <typename T, int DIM>
__global__ void stencil_dinv_kernel(int64_t N, const double* __restrict__ S, T* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  constexpr int oc = DIM == 3 ? 13 : 4;
#pragma unroll
  for (int r = 0; r < DIM; ++r) out[n * DIM + r] = (T)(1.0 / S[((int64_t)(oc * DIM + r) * DIM + r) * N + n]);
}

// PCG: r -= alpha q (alpha = rho / pq); optionally a working-precision copy of
// r for the V-cycle; per-block partial sums of r.r.  x += alpha p is deferred
// to the p update, which reads p anyway: alpha goes to SC_ALPHA (zero when
// this update did not run).  A no-op once the loop has stopped (SC_STOP).
template <typename TB>
__global__ __launch_bounds__(256) void pcg_update_xr_kernel(int64_t n, double* __restrict__ sc, int cur,
                                                            const double* __restrict__ pq_part, int n_pq,
                                                            const double* __restrict__ q,
                                                            double* __restrict__ r,
                                                            TB* __restrict__ rb,
                                                            double* __restrict__ partial) {
  if (sc[SC_STOP] != 0.0) {
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_ALPHA] = 0.0;
    return;
  }
  // p.q from its partials: every block sums them in the same fixed order
  double pq = 0.0;
  for (int t = threadIdx.x; t < n_pq; t += blockDim.x) pq += pq_part[t];
  pq = block_sum256_all(pq);
  const double alpha = sc[SC_RHO + cur] / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[SC_PQ] = pq;
    sc[SC_ALPHA] = alpha;
  }
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double rv = r[t] - alpha * q[t];
    r[t] = rv;
    if (rb != nullptr) rb[t] = (TB)rv;
    s += rv * rv;
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// per-block partial sums of r.z with z in the V-cycle's precision
template <typename TZ>
__global__ __launch_bounds__(256) void pcg_dot_kernel(int64_t n, const double* __restrict__ r,
                                                      const TZ* __restrict__ z,
                                                      double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    s += r[t] * (double)z[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// x += alpha p (the deferred update of this iteration, alpha from SC_ALPHA),
// then p = z + beta p, beta = rho_new / rho (Fletcher-Reeves); p = z and no x
// update when first.  rho_new = r.z is summed from its partials by every block
// in the same fixed order and stored in slot SC_RHO + nxt; the previous rho is
// in the other slot.
template <typename TZ>
__global__ __launch_bounds__(256) void pcg_update_p_kernel(int64_t n, double* __restrict__ sc, int nxt,
                                                           const double* __restrict__ rz_part, int n_rz,
                                                           const TZ* __restrict__ z,
                                                           double* __restrict__ p, double* __restrict__ x,
                                                           int first) {
  double rz = 0.0;
  for (int t = threadIdx.x; t < n_rz; t += blockDim.x) rz += rz_part[t];
  rz = block_sum256_all(rz);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_RHO + nxt] = rz;
  const double beta = first ? 0.0 : rz / sc[SC_RHO + (nxt ^ 1)];
  const double alpha = first ? 0.0 : sc[SC_ALPHA];
  const bool upd_x = alpha != 0.0;  // zero: the x/r update of this iteration did not run
  // four consecutive values per step (z, p and x are 16-byte aligned allocations)
  const int64_t n4 = n >> 2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n4;
       t += (int64_t)gridDim.x * blockDim.x) {
    TZ zv[4];
    double pv[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) zv[u] = z[4 * t + u];
    if (!first) {
#pragma unroll
      for (int u = 0; u < 4; ++u) pv[u] = p[4 * t + u];
    }
    if (upd_x) {
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[u] = x[4 * t + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[4 * t + u] = xv[u] + alpha * pv[u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) p[4 * t + u] = first ? (double)zv[u] : (double)zv[u] + beta * pv[u];
  }
  for (int64_t t = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (upd_x) x[t] += alpha * p[t];
    p[t] = first ? (double)z[t] : (double)z[t] + beta * p[t];
  }
}

// Convergence / breakdown test of one iteration, evaluated on the device so the
// host never waits on the current iteration: r.r from the update's partials,
// R = |r| / |b|; non-finite p.q or r.r, p.q <= 0 and R < tol stop the loop
// (first reason wins, as in a plain sequential loop).  While running, omega =
// min + d/2 exp((-clip(R, tol, 0.1) + tol) * 500) for the next V-cycle.
// {R, status, stop} is published to pinned host memory.  One block of 256.
__global__ __launch_bounds__(256) void pcg_check_kernel(double* sc, const double* __restrict__ rr_part, int n_rr,
                                                        double tol, double wmin, double dw, double* ring) {
  double rr = 0.0;
  for (int t = threadIdx.x; t < n_rr; t += blockDim.x) rr += rr_part[t];
  rr = block_sum256(rr);
  if (threadIdx.x != 0) return;
  if (sc[SC_STOP] == 0.0) {
    sc[SC_RR] = rr;
    const double pq = sc[SC_PQ];
    const double R = sqrt(rr) / sqrt(sc[SC_BB]);
    double status = (double)ET_OK, stop = 0.0;
    if (!isfinite(pq) || !isfinite(rr)) {
      status = (double)ET_ENONFINITE;
      stop = 1.0;
    } else if (!(pq > 0.0)) {
      status = (double)ET_EBREAKDOWN;
      stop = 1.0;
    } else if (R < tol) {
      stop = 1.0;
    }
    sc[SC_R] = R;
    sc[SC_STATUS] = status;
    sc[SC_STOP] = stop;
    if (stop == 0.0) {
      const double Rc = fmin(fmax(R, tol), 1e-1);
      sc[SC_OMEGA] = wmin + dw / 2.0 * exp((-Rc + tol) * 500.0);
    }
  }
  ring[0] = sc[SC_R];
  ring[1] = sc[SC_STATUS];
  ring[2] = sc[SC_STOP];
  __threadfence_system();
}

__global__ void pcg_init_kernel(double* sc, double omega) {
  sc[SC_OMEGA] = omega;
  sc[SC_R] = 0.0;
  sc[SC_STATUS] = 0.0;
  sc[SC_STOP] = 0.0;
}

__global__ void set_scalar_kernel(double* sc, int slot, double v) { sc[slot] = v; }

// ------------------------------------------------------------ host side ----

static int op_kind(int level, int mode) {
  if (level > 0) return ET_PROF_COARSE_OPS;
  if (mode == OP_JAC0) return ET_PROF_L0_JAC0;
  return mode == OP_RES ? ET_PROF_L0_RES : (mode == OP_JAC_ADD ? ET_PROF_L0_JACADD : ET_PROF_L0_JAC);
}

template <typename T>
static int op_launch(const MGLevel<T>& L, const MGShared& sh, int mode, const T* x, const T* b,
                     const double* omega, double wscale, const T* xadd, T* y, hipStream_t s) {
  const Grid& g = L.g;
  sh.prof->begin(op_kind(L.level, mode), s);
  if (L.level == 1 && sh.l1 != nullptr) {
    const int st = sh.l1->apply<T>(mode, x, b, L.Dinv, omega, wscale, xadd, y, s);
    sh.prof->end(s);
    return st;
  }
  if (L.level == 0 && sizeof(T) == 4 && sh.mfma) {
    const int st = op0_mfma(mode, g, (int64_t)g.ey * g.ez, g.ez, sh.kscale, sh.K0f, sh.Ef, sh.cons,
                            (const float*)x, (const float*)b, (const float*)L.Dinv, omega, wscale,
                            (const float*)xadd, (float*)y, s);
    sh.prof->end(s);
    return st;
  }
  if (L.level == 0 && sh.tiled) {
    const int st = op0_fine<T>(mode, g, (int64_t)g.ey * g.ez, g.ez, sh.K0<T>(), sh.wht, sh.E<T>(), sh.cons,
                                x, b, L.Dinv, omega, wscale, xadd, y, s);
    sh.prof->end(s);
    return st;
  }
  const unsigned grid = grid_for(g.n_nodes());
#define ET_OP0(DIM, MODE)                                                                   \
  op0_kernel<T, DIM, MODE><<<grid, kBlock, 0, s>>>(g, sh.K0<T>(), sh.E<T>(), sh.cons, x, b, \
                                                   L.Dinv, omega, wscale, xadd, y)
#define ET_OPS(DIM, MODE)                                                                            \
  do {                                                                                               \
    if (g.n_nodes() <= g_row_stencil_max_nodes && g.n_nodes() * 126 * (int64_t)sizeof(T) < ((int64_t)1 << 32)) \
      opS_row_kernel<T, DIM, MODE><<<grid_for(g.n_nodes() * DIM), kBlock, 0, s>>>(g, L.S, x, b, L.Dinv, \
                                                                                  omega, wscale, xadd, y); \
    else                                                                                             \
      opS_kernel<T, DIM, MODE><<<grid, kBlock, 0, s>>>(g, L.S, x, b, L.Dinv, omega, wscale, xadd, y); \
  } while (0)
#define ET_DISPATCH(M)                \
  do {                                \
    if (L.level == 0) {               \
      if (g.dim == 3)                 \
        ET_OP0(3, M);                 \
      else                            \
        ET_OP0(2, M);                 \
    } else {                          \
      if (g.dim == 3)                 \
        ET_OPS(3, M);                 \
      else                            \
        ET_OPS(2, M);                 \
    }                                 \
  } while (0)
  switch (mode) {
    case OP_Y: ET_DISPATCH(OP_Y); break;
    case OP_RES: ET_DISPATCH(OP_RES); break;
    case OP_JAC: ET_DISPATCH(OP_JAC); break;
    default: ET_DISPATCH(OP_JAC_ADD); break;
  }
#undef ET_DISPATCH
#undef ET_OPS
#undef ET_OP0
  sh.prof->end(s);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

template <typename T>
int MGHierarchy<T>::vcycle(int l, MGShared& sh, hipStream_t s, bool x0_ready) {
  MGLevel<T>& L = lv[l];
  if (l == n_level) {
    sh.prof->begin(ET_PROF_COARSE_SOLVE, s);
    const int st = coarse->solve<T, T>(L.b, L.x, s);
    sh.prof->end(s);
    return st;
  }
  const double ws = std::pow(1.06, l);
  const double* om = sh.scal + SC_OMEGA;
  const int64_t n = L.n;
  if (n_smooth == 0) {
    ET_CHECK_HIP(hipMemsetAsync(L.x, 0, n * sizeof(T), s));
  } else {
    int it = 1;
    // both precisions (the FP64 sweeps run on 8-slice tiles, where the fused pass is faster)
    if (l == 0 && n_smooth >= 2 && sh.tiled && !(sizeof(T) == 4 && sh.mfma) && fine_jac0_supported(L.g, sizeof(T))) {
      // x0 = w D^-1 b and the first sweep from it in one pass
      ET_TRY(op_launch<T>(L, sh, OP_JAC0, nullptr, L.b, om, ws, nullptr, L.t, s));
      std::swap(L.x, L.t);
      it = 2;
    } else if (!x0_ready) {
      sh.prof->begin(l == 0 ? ET_PROF_L0_SCALE : ET_PROF_COARSE_OPS, s);
      scale_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, L.Dinv, L.b, om, ws, L.x);
      sh.prof->end(s);
      ET_CHECK_LAUNCH();
    }
    for (; it < n_smooth; ++it) {
      ET_TRY(op_launch<T>(L, sh, OP_JAC, L.x, L.b, om, ws, nullptr, L.t, s));
      std::swap(L.x, L.t);
    }
  }
  const int passes = (cycle_w && ((w_level_mask >> l) & 1)) ? 2 : 1;
  for (int pass = 0; pass < passes; ++pass) {
    ET_TRY(op_launch<T>(L, sh, OP_RES, L.x, L.b, om, ws, nullptr, L.r, s));
    sh.prof->begin(ET_PROF_TRANSFER, s);
    // a smoothed coarse level starts from x = w D^-1 b: written by the restriction
    const bool x0 = l + 1 < n_level && n_smooth > 0;
    MGLevel<T>& C = lv[l + 1];
    ET_TRY(restrict_op<T>(L.g, dof, L.r, C.b, s, x0 ? C.Dinv : nullptr, om, std::pow(1.06, l + 1),
                          x0 ? C.x : nullptr));
    sh.prof->end(s);
    ET_TRY(vcycle(l + 1, sh, s, x0));
    sh.prof->begin(ET_PROF_TRANSFER, s);
    ET_TRY(prolong_op<T>(L.g, dof, lv[l + 1].x, L.e, s));
    sh.prof->end(s);
    if (n_smooth == 0) {
      axpy_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, L.e, L.x);
      ET_CHECK_LAUNCH();
    }
    for (int it = 0; it < n_smooth; ++it) {
      if (it + 1 < n_smooth) {
        ET_TRY(op_launch<T>(L, sh, OP_JAC, L.e, L.r, om, ws, nullptr, L.t, s));
        std::swap(L.e, L.t);
      } else if (l == 0 && pass + 1 == passes && sh.want_dot) {
        // the preconditioned residual z is final here: r.z partials in the same pass
        sh.prof->begin(ET_PROF_L0_JACADD, s);
        const int st = op0_wht_jac_add_dot<T>(L.g, (int64_t)L.g.ey * L.g.ez, L.g.ez, sh.wht, sh.E<T>(),
                                              sh.cons, L.e, L.r, L.Dinv, om, ws, L.x, L.t, L.b,
                                              sh.dot_part, kRedBlocks, s);
        sh.prof->end(s);
        ET_TRY(st);
        sh.dot_done = true;
        std::swap(L.x, L.t);
      } else {
        ET_TRY(op_launch<T>(L, sh, OP_JAC_ADD, L.e, L.r, om, ws, L.x, L.t, s));
        std::swap(L.x, L.t);
      }
    }
  }
  return ET_OK;
}

template <typename T>
int MGHierarchy<T>::alloc(const Grid& g0, int dof_, int n_level_, hipStream_t s, bool skip_s1) {
  dof = dof_;
  n_level = n_level_;
  lv.resize(n_level + 1);
  Grid g = g0;
  for (int l = 0; l <= n_level; ++l) {
    MGLevel<T>& L = lv[l];
    L.level = l;
    L.g = g;
    L.n = g.n_nodes() * dof;
    const size_t vb = (size_t)L.n * sizeof(T);
    for (T** p : {&L.b, &L.x, &L.r, &L.e, &L.t, &L.Dinv}) ET_CHECK_HIP(hipMalloc(p, vb));
    if (l > 0 && l < n_level && !(l == 1 && skip_s1)) {
      const int NH = g.dim == 3 ? 14 : 5;  // centre + forward offsets
      ET_CHECK_HIP(hipMalloc(&L.S, (size_t)NH * dof * dof * g.n_nodes() * sizeof(T)));
    }
    g = g.coarse();
  }
  (void)s;
  return ET_OK;
}

template <typename T>
void MGHierarchy<T>::release() {
  for (auto& L : lv)
    for (T* p : {L.b, L.x, L.r, L.e, L.t, L.Dinv, L.S})
      if (p) (void)hipFree(p);
  lv.clear();
}

template struct MGHierarchy<double>;
template struct MGHierarchy<float>;

// Level-0 operator outside the solver (C ABI): the fine-grid kernels in 3-D,
// the node-gather kernel otherwise.
template <typename T>
int level0_op(int mode, const Grid& g, const T* K0, const T* E, const uint8_t* cons, const T* x,
              const T* b, const T* Dinv, const double* omega, double wscale, const T* xadd, T* y,
              hipStream_t s, const WhtTable* fine) {
  if (g.dim == 3 && fine != nullptr && fine_supported(g, sizeof(T)))
    return op0_fine<T>(mode, g, (int64_t)g.ey * g.ez, g.ez, K0, *fine, E, cons, x, b, Dinv, omega, wscale, xadd, y,
                       s);
  const unsigned grid = grid_for(g.n_nodes());
#define ET_L0(DIM, MODE) \
  op0_kernel<T, DIM, MODE><<<grid, kBlock, 0, s>>>(g, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y)
#define ET_L0D(MODE)     \
  do {                   \
    if (g.dim == 3)      \
      ET_L0(3, MODE);    \
    else                 \
      ET_L0(2, MODE);    \
  } while (0)
  switch (mode) {
    case OP_Y: ET_L0D(OP_Y); break;
    case OP_RES: ET_L0D(OP_RES); break;
    case OP_JAC: ET_L0D(OP_JAC); break;
    default: ET_L0D(OP_JAC_ADD); break;
  }
#undef ET_L0D
#undef ET_L0
  ET_CHECK_LAUNCH();
  return ET_OK;
}

template int level0_op<double>(int, const Grid&, const double*, const double*, const uint8_t*,
                               const double*, const double*, const double*, const double*, double,
                               const double*, double*, hipStream_t, const WhtTable*);
template int level0_op<float>(int, const Grid&, const float*, const float*, const uint8_t*,
                              const float*, const float*, const float*, const double*, double,
                              const float*, float*, hipStream_t, const WhtTable*);

// ------------------------------------------------------------ et_mg ----

int MGSolver::coarse_solve(const void* b, void* x) {
  ET_REQUIRE(built, "et_mg_coarse_solve: no hierarchy yet (call et_mg_solve first)");
  if (prm.precision == 1)
    ET_TRY((coarse.solve<float, float>((const float*)b, (float*)x, stream)));
  else
    ET_TRY((coarse.solve<double, double>((const double*)b, (double*)x, stream)));
  ET_CHECK_HIP(hipStreamSynchronize(stream));
  return ET_OK;
}

int MGSolver::apply_level(int level, const void* x, void* y) {
  ET_REQUIRE(built, "et_mg_apply_level: no hierarchy yet (call et_mg_solve first)");
  ET_REQUIRE(level >= 0 && level < prm.n_level, "et_mg_apply_level: level must be in [0, n_level)");
  if (prm.precision == 1)
    ET_TRY(op_launch<float>(h32.lv[level], sh, OP_Y, (const float*)x, nullptr, nullptr, 1.0, nullptr,
                            (float*)y, stream));
  else
    ET_TRY(op_launch<double>(h64.lv[level], sh, OP_Y, (const double*)x, nullptr, nullptr, 1.0, nullptr,
                             (double*)y, stream));
  ET_CHECK_HIP(hipStreamSynchronize(stream));
  return ET_OK;
}

int MGSolver::profile_read(double* ms, int64_t* launches) {
  ET_CHECK_HIP(hipStreamSynchronize(stream));
  for (int k = 0; k < ET_PROF_KINDS; ++k) {
    ms[k] = 0.0;
    launches[k] = 0;
  }
  for (const auto& r : prof.recs) {
    float t = 0.f;
    ET_CHECK_HIP(hipEventElapsedTime(&t, r.a, r.b));
    if (r.kind >= 0 && r.kind < ET_PROF_KINDS) {
      ms[r.kind] += t;
      launches[r.kind] += 1;
    }
  }
  return ET_OK;
}

int MGSolver::create(const et_mg_params& p, const double* K0_host, hipStream_t s) {
  prm = p;
  sh.prof = &prof;
  stream = s;
  ET_TRY(make_grid(p.dim, p.nel, &g0));
  dof = p.dim;
  ET_REQUIRE(p.n_level >= 0 && p.n_level <= 12, "n_level out of range");
  ET_REQUIRE(p.n_smooth >= 0 && p.n_smooth <= 64, "n_smooth out of range");
  ET_REQUIRE(p.precision == 0 || p.precision == 1, "precision must be 0 (fp64) or 1 (mixed)");
  {
    Grid g = g0;
    for (int l = 0; l < p.n_level; ++l) {
      ET_REQUIRE(g.ex >= 2 && g.ey >= 2 && (g.dim == 2 || g.ez >= 2),
                 "n_level too deep for this mesh (a level would have fewer than one element)");
      g = g.coarse();
    }
  }
  n_dof = g0.n_nodes() * dof;
  n_elem = g0.n_elem();
  // the fast 3-D fine-grid kernels build K0's blocks from the brick's reflections:
  // any other element matrix takes the general node-gather kernel
  k0_flags = k0_structure(p.dim, K0_host, kK0Rtol);
  sh.tiled = g0.dim == 3 && k0_fast3d(k0_flags) && fine_supported(g0, sizeof(double)) &&
             fine_supported(g0, sizeof(float));
  sh.mfma = p.precision == 1 && sh.tiled && get_mixed_mfma() && mfma_supported(g0);
  if (sh.tiled) wht_table(K0_host, &sh.wht);
  sh.kscale = k0_scale(K0_host, (p.dim == 3 ? 24 * 24 : 8 * 8));
  const int nd = (p.dim == 3 ? 8 : 4) * dof;
  ET_CHECK_HIP(hipMalloc(&sh.K0d, nd * nd * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&sh.K0f, nd * nd * sizeof(float)));
  ET_CHECK_HIP(hipMemcpyAsync(sh.K0d, K0_host, nd * nd * sizeof(double), hipMemcpyHostToDevice, s));
  {
    std::vector<float> kf(nd * nd);
    for (int t = 0; t < nd * nd; ++t) kf[t] = (float)K0_host[t];
    ET_CHECK_HIP(hipMemcpyAsync(sh.K0f, kf.data(), nd * nd * sizeof(float), hipMemcpyHostToDevice, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
  }
  ET_CHECK_HIP(hipMalloc(&sh.scal, SC_COUNT * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&partial, 2 * kRedBlocks * sizeof(double)));
  ET_CHECK_HIP(hipHostMalloc(&ring, 8 * sizeof(double), hipHostMallocCoherent));
  ET_CHECK_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&ring_dev), ring, 0));
  for (auto& e : ring_ev) ET_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (double** v : {&r, &z, &p_, &q, &diag64}) ET_CHECK_HIP(hipMalloc(v, n_dof * sizeof(double)));
  if (p.precision == 1) ET_CHECK_HIP(hipMalloc(&sh.Ef, n_elem * sizeof(float)));
  // FP64 stencils: level 0 (fine operator) and every coarse level
  const int NO = p.dim == 3 ? 27 : 9;
  S64.assign(p.n_level + 1, nullptr);
  elem_galerkin = p.n_level > 0 && galerkin_elem_supported(g0);
  if (elem_galerkin) {
    const std::vector<double> M = galerkin_child_tables(p.dim, K0_host);
    ET_CHECK_HIP(hipMalloc(&Mt, M.size() * sizeof(double)));
    ET_CHECK_HIP(hipMemcpy(Mt, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
    ET_CHECK_HIP(hipMalloc(&eflag, (size_t)n_elem));
  }
  {
    Grid g = g0;
    for (int l = 0; l <= p.n_level; ++l) {
      const bool need = p.n_level > 0 && !(l == 0 && elem_galerkin);
      if (need) ET_CHECK_HIP(hipMalloc(&S64[l], (size_t)NO * dof * dof * g.n_nodes() * sizeof(double)));
      if (l == p.n_level) coarse_grid = g;
      g = g.coarse();
    }
  }
  {
    // scratch of the three-pass Galerkin products (the largest of levels built from a stencil);
    // above 4 GiB the one-pass kernels run instead
    int64_t need = 0;
    Grid g = g0;
    for (int l = 1; l <= p.n_level; ++l) {
      if (!(l == 1 && elem_galerkin)) need = std::max(need, galerkin_work_doubles(g));
      g = g.coarse();
    }
    if (need > 0 && need * (int64_t)sizeof(double) <= ((int64_t)1 << 32))
      ET_CHECK_HIP(hipMalloc(&gal_work, (size_t)need * sizeof(double)));
  }
  use_l1 = elem_galerkin && get_level1_mf() && level1_supported(g0, p.n_level) && (k0_flags & K0S_LEVEL1);
  if (use_l1) ET_TRY(l1.init(g0, K0_host, p.precision == 1, s));
  sh.l1 = use_l1 ? &l1 : nullptr;
  if (p.precision == 0)
    ET_TRY(h64.alloc(g0, dof, p.n_level, s, use_l1));
  else
    ET_TRY(h32.alloc(g0, dof, p.n_level, s, use_l1));
  ET_TRY(coarse.init((int)(coarse_grid.n_nodes() * dof), s, p.precision == 1));
  h64.coarse = &coarse;
  h32.coarse = &coarse;
  for (auto* h : {&h64.n_smooth, &h32.n_smooth}) *h = p.n_smooth;
  h64.cycle_w = h32.cycle_w = p.cycle_w;
  h64.w_level_mask = h32.w_level_mask = p.w_level_mask;
  max_omega = p.omega;
  min_omega = std::fmax(p.omega / 2.0, p.min_omega);
  d_omega = max_omega - min_omega;
  ET_CHECK_HIP(hipEventCreate(&ev0));
  ET_CHECK_HIP(hipEventCreate(&ev1));
  ET_CHECK_HIP(hipEventCreate(&ev2));
  return ET_OK;
}

void MGSolver::destroy() {
  prof.release();
  l1.release();
  h64.release();
  h32.release();
  coarse.release();
  for (double* v : S64)
    if (v) (void)hipFree(v);
  for (void* v : {(void*)sh.K0d, (void*)sh.K0f, (void*)sh.scal, (void*)partial, (void*)r, (void*)z,
                  (void*)p_, (void*)q, (void*)diag64, (void*)sh.Ef, (void*)Mt, (void*)eflag, (void*)gal_work})
    if (v) (void)hipFree(v);
  if (ring) (void)hipHostFree(ring);
  for (auto& e : ring_ev)
    if (e) (void)hipEventDestroy(e);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (ev2) (void)hipEventDestroy(ev2);
}

template <typename T>
int MGSolver::setup_levels(MGHierarchy<T>& h) {
  hipStream_t s = stream;
  const int L = prm.n_level;
  // level 0: inverse diagonal
  ET_TRY(diagonal<double>(g0, sh.K0d, sh.Ed, sh.cons, diag64, s));
  recip_kernel<T><<<grid_for(n_dof), kBlock, 0, s>>>(n_dof, diag64, h.lv[0].Dinv);
  ET_CHECK_LAUNCH();
  if (L == 0) {
    return ET_EINVAL;  // a direct solve of the fine grid is not supported
  }
  if (!elem_galerkin) ET_TRY(stencil_from_operator(g0, sh.K0d, sh.Ed, sh.cons, S64[0], s));
  Grid g = g0;
  for (int l = 1; l <= L; ++l) {
    if (l == 1 && elem_galerkin) {
      ET_TRY(galerkin_from_elements(g0, sh.K0d, Mt, sh.Ed, sh.cons, eflag, S64[1], s));
      if (use_l1) ET_TRY(l1.setup(sh.Ed, Mt, eflag, S64[1], s));
    } else
      ET_TRY(galerkin(g, dof, S64[l - 1], S64[l], s, gal_work));
    g = g.coarse();
    MGLevel<T>& lev = h.lv[l];
    const int64_t N = g.n_nodes();
    if (g.dim == 3)
      stencil_dinv_kernel<T, 3><<<grid_for(N), kBlock, 0, s>>>(N, S64[l], lev.Dinv);
    else
      stencil_dinv_kernel<T, 2><<<grid_for(N), kBlock, 0, s>>>(N, S64[l], lev.Dinv);
    ET_CHECK_LAUNCH();
    if (l < L && !(l == 1 && use_l1)) {
      // the centre and forward offsets are the trailing blocks of the full stencil
      const int OC = g.dim == 3 ? 13 : 4;
      const int64_t cnt = (int64_t)(OC + 1) * dof * dof * N;
      convert_kernel<double, T><<<grid_for(cnt), kBlock, 0, s>>>(cnt, S64[l] + (int64_t)OC * dof * dof * N,
                                                                 lev.S);
      ET_CHECK_LAUNCH();
    }
  }
  ET_TRY(coarse.factor(coarse_grid, dof, S64[L], s));
  return ET_OK;
}

template <typename T>
int MGSolver::precondition(MGHierarchy<T>& h, bool rhs_ready, const T** zout) {
  // z = V(r).  FP64 V-cycles read r in place (level-0 b is only read);
  // FP32 V-cycles take the FP32 copy of r that the x/r update wrote.
  hipStream_t s = stream;
  MGLevel<T>& L0 = h.lv[0];
  if constexpr (sizeof(T) == sizeof(double)) {
    T* saved = L0.b;
    L0.b = r;
    const int st = h.vcycle(0, sh, s);
    L0.b = saved;
    ET_TRY(st);
  } else {
    if (!rhs_ready) {
      prof.begin(ET_PROF_PCG_VEC, s);
      convert_kernel<double, T><<<grid_for(n_dof), kBlock, 0, s>>>(n_dof, r, L0.b);
      prof.end(s);
      ET_CHECK_LAUNCH();
    }
    ET_TRY(h.vcycle(0, sh, s));
  }
  *zout = L0.x;
  return ET_OK;
}

int MGSolver::apply_A64(const double* x, double* y, int mode, const double* b) {
  hipStream_t s = stream;
  const unsigned grid = grid_for(g0.n_nodes());
  prof.begin(ET_PROF_MATVEC64, s);
  if (sh.tiled) {
    const int st = op0_fine<double>(mode, g0, (int64_t)g0.ey * g0.ez, g0.ez, sh.K0d, sh.wht, sh.Ed, sh.cons,
                                     x, b, nullptr, sh.scal, 1.0, nullptr, y, s);
    prof.end(s);
    return st;
  }
  if (g0.dim == 3) {
    if (mode == OP_Y)
      op0_kernel<double, 3, OP_Y><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
    else
      op0_kernel<double, 3, OP_RES><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
  } else {
    if (mode == OP_Y)
      op0_kernel<double, 2, OP_Y><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
    else
      op0_kernel<double, 2, OP_RES><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
  }
  prof.end(s);
  ET_CHECK_LAUNCH();
  return ET_OK;
}


int MGSolver::solve(const double* E, const uint8_t* cons, const double* b, double* x, double tol,
                    int maxiter, int* iters, double* residual) {
  ET_REQUIRE(E && cons && b && x, "et_mg_solve: null pointer");
  ET_REQUIRE(maxiter >= 1, "et_mg_solve: maxiter must be >= 1");
  ET_REQUIRE(prm.n_level >= 1, "et_mg_solve: n_level must be >= 1");
  hipStream_t s = stream;
  sh.Ed = E;
  sh.cons = cons;
  ET_CHECK_HIP(hipEventRecord(ev0, s));
  prof.begin(ET_PROF_SETUP, s);
  if (prm.precision == 1) {
    convert_kernel<double, float><<<grid_for(n_elem), kBlock, 0, s>>>(n_elem, E, sh.Ef);
    ET_CHECK_LAUNCH();
    ET_TRY(setup_levels(h32));
  } else {
    ET_TRY(setup_levels(h64));
  }
  prof.end(s);
  ET_TRY(coarse.check_info(s));
  built = true;
  ET_CHECK_HIP(hipEventRecord(ev1, s));
  // omega <- min_omega at the start of every solve (GPU semantics)
  pcg_init_kernel<<<1, 1, 0, s>>>(sh.scal, min_omega);
  ET_CHECK_LAUNCH();
  int status = prm.precision == 1 ? pcg(h32, b, x, tol, maxiter, iters, residual)
                                  : pcg(h64, b, x, tol, maxiter, iters, residual);
  ET_CHECK_HIP(hipEventRecord(ev2, s));
  ET_CHECK_HIP(hipEventSynchronize(ev2));
  float a = 0.f, c = 0.f;
  ET_CHECK_HIP(hipEventElapsedTime(&a, ev0, ev1));
  ET_CHECK_HIP(hipEventElapsedTime(&c, ev1, ev2));
  t_setup_ms = a;
  t_pcg_ms = c;
  return status;
}

// Preconditioned CG (Fletcher-Reeves, as the reference) with the convergence
// test evaluated on the device.  The host enqueues iteration k+1 before it
// reads the outcome of iteration k, so the GPU never idles on a host round
// trip; once the device has stopped the loop, the already-enqueued work of the
// next iteration leaves x and r untouched.
template <typename T>
int MGSolver::pcg(MGHierarchy<T>& h, const double* b, double* x, double tol, int maxiter, int* iters,
                   double* residual) {
  hipStream_t s = stream;
  T* rb = nullptr;
  if constexpr (sizeof(T) == sizeof(float)) rb = h.lv[0].b;
  // r = b - A x ; |b|^2
  ET_TRY(apply_A64(x, r, OP_RES, b));
  ET_TRY(dot(n_dof, b, b, partial, sh.scal + SC_BB, s));
  {
    // b = 0: the solution is x = 0 (the reference's relative test divides by
    // |b| = 0 and its iterates become NaN); one host read per solve
    double bb = 0.0;
    ET_CHECK_HIP(hipMemcpyAsync(&bb, sh.scal + SC_BB, sizeof(double), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    if (bb == 0.0) {
      ET_CHECK_HIP(hipMemsetAsync(x, 0, n_dof * sizeof(double), s));
      if (iters) *iters = 0;
      if (residual) *residual = 0.0;
      return ET_OK;
    }
  }
  const T* z = nullptr;
  ET_TRY(precondition(h, false, &z));
  prof.begin(ET_PROF_PCG_VEC, s);
  // partial: first-stage sums of p.q and r.z; partial2: of r.r
  double* partial2 = partial + kRedBlocks;
  pcg_dot_kernel<T><<<kRedBlocks, 256, 0, s>>>(n_dof, r, z, partial);
  pcg_update_p_kernel<T><<<kRedBlocks, 256, 0, s>>>(n_dof, sh.scal, 0, partial, kRedBlocks, z, p_, x, 1);
  prof.end(s);
  ET_CHECK_LAUNCH();
  // the fused q = A p, p.q path exists for the transformed-basis operator
  const bool fused_pq = sh.tiled && g0.dim == 3 && get_wht() && wht_supported(g0, sizeof(double)) &&
                        wht_partials(g0, sizeof(double), TM_YDOT) <= kRedBlocks;
  const int n_pq_fused = fused_pq ? wht_partials(g0, sizeof(double), TM_YDOT) : 0;
  // the last fine-grid sweep of the V-cycle (same operator) yields r.z, with r
  // in the V-cycle's precision (FP32 V-cycles: the copy of r they read)
  const bool fused_rz = fused_pq && get_fused_rz() &&
                        wht_supported(g0, sizeof(T)) && wht_partials(g0, sizeof(T), TM_JAC_ADD_DOT) <= kRedBlocks &&
                        !sh.mfma && h.n_smooth > 0;
  const int n_rz_fused = fused_rz ? wht_partials(g0, sizeof(T), TM_JAC_ADD_DOT) : kRedBlocks;
  int last = 0;
  auto read_ring = [&](int k, double* R, int* st) -> bool {
    ET_CHECK_HIP(hipEventSynchronize(ring_ev[k & 1]));
    const double* slot = ring + 4 * (k & 1);
    *R = slot[0];
    *st = (int)slot[1];
    return slot[2] != 0.0;
  };
  double R = 0.0;
  int status = ET_OK;
  bool stopped = false;
  for (int it = 1; it <= maxiter; ++it) {
    // rho of this iteration is in slot SC_RHO + cur; the new one goes to the other
    const int cur = (it - 1) & 1;
    int n_pq = kRedBlocks;
    if (fused_pq) {
      // q = A p with p.q summed per workgroup inside the operator
      prof.begin(ET_PROF_MATVEC64, s);
      ET_TRY(op0_wht_ydot<double>(g0, (int64_t)g0.ey * g0.ez, g0.ez, sh.wht, sh.Ed, sh.cons, p_, q, partial,
                                  kRedBlocks, s));
      prof.end(s);
      n_pq = n_pq_fused;
      prof.begin(ET_PROF_PCG_VEC, s);
    } else {
      ET_TRY(apply_A64(p_, q, OP_Y, nullptr));
      prof.begin(ET_PROF_PCG_VEC, s);
      ET_TRY(dot_partials(n_dof, p_, q, partial, s));
    }
    pcg_update_xr_kernel<T><<<kRedBlocks, 256, 0, s>>>(n_dof, sh.scal, cur, partial, n_pq, q, r, rb,
                                                       partial2);
    ET_CHECK_LAUNCH();
    pcg_check_kernel<<<1, 256, 0, s>>>(sh.scal, partial2, kRedBlocks, tol, min_omega, d_omega,
                                       ring_dev + 4 * (it & 1));
    ET_CHECK_LAUNCH();
    ET_CHECK_HIP(hipEventRecord(ring_ev[it & 1], s));
    prof.end(s);
    last = it;
    if (it > 1 && read_ring(it - 1, &R, &status)) {
      last = it - 1;
      stopped = true;
      break;
    }
    // r.z partials from the V-cycle's last fine sweep where it can produce them
    sh.want_dot = fused_rz;
    sh.dot_part = partial;
    sh.dot_done = false;
    const int pst = precondition(h, true, &z);
    sh.want_dot = false;
    ET_TRY(pst);
    prof.begin(ET_PROF_PCG_VEC, s);
    int n_rz = n_rz_fused;
    if (!sh.dot_done) {
      pcg_dot_kernel<T><<<kRedBlocks, 256, 0, s>>>(n_dof, r, z, partial);
      n_rz = kRedBlocks;
    }
    pcg_update_p_kernel<T><<<kRedBlocks, 256, 0, s>>>(n_dof, sh.scal, it & 1, partial, n_rz, z, p_, x, 0);
    prof.end(s);
    ET_CHECK_LAUNCH();
  }
  if (!stopped) stopped = read_ring(last, &R, &status);
  if (!stopped && status == ET_OK) {
    set_error("PCG: maxiter reached without meeting the tolerance");
    status = ET_ENOCONV;
  }
  if (status == ET_ENONFINITE) set_error("PCG: non-finite value (p.q or r.r)");
  if (status == ET_EBREAKDOWN) set_error("PCG breakdown: p^T A p <= 0");
  if (iters) *iters = last;
  if (residual) *residual = R;
  return status;
}

}  // namespace et
